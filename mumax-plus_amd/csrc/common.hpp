// Shared device/host helpers for the MI355X micromagnetic core.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include "../../include/mumaxplus_amd.h"

namespace mpa {

// Errors are C++ exceptions; the C-ABI layer turns them into return codes and
// a thread-local message.  (The reference exits the process on a CUDA error,
// src/cudautil/cudaerror.hpp:9-18.)
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

// Error codes MPA_ERR_* come from the C ABI header.

inline void hipCheck(hipError_t e, const char* what) {
  if (e != hipSuccess)
    throw Error(MPA_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
#define HIP_CHECK(x) ::mpa::hipCheck((x), #x)

inline void checkLaunch(const char* name) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    throw Error(MPA_ERR_HIP, std::string("launch of ") + name + ": " + hipGetErrorString(e));
}

// RAII device allocation.
template <typename T>
struct DeviceArray {
  T* ptr = nullptr;
  size_t n = 0;
  DeviceArray() = default;
  explicit DeviceArray(size_t count) { allocate(count); }
  DeviceArray(const DeviceArray&) = delete;
  DeviceArray& operator=(const DeviceArray&) = delete;
  DeviceArray(DeviceArray&& o) noexcept : ptr(o.ptr), n(o.n) { o.ptr = nullptr; o.n = 0; }
  DeviceArray& operator=(DeviceArray&& o) noexcept {
    if (this != &o) { release(); ptr = o.ptr; n = o.n; o.ptr = nullptr; o.n = 0; }
    return *this;
  }
  ~DeviceArray() { release(); }
  void allocate(size_t count) {
    release();
    if (count == 0) return;
    HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&ptr), count * sizeof(T)));
    n = count;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    n = 0;
  }
  T* get() const { return ptr; }
  size_t bytes() const { return n * sizeof(T); }
};

// Scalar parameter as seen by kernels: uniform value, or per-cell array.
struct PS {
  const float* p;  // nullptr -> uniform
  float v;
  __device__ __forceinline__ float at(int i) const { return p ? p[i] : v; }
};

// Vector parameter: uniform (x,y,z) or SoA per-cell array of 3*N floats.
struct PV {
  const float* p;
  float x, y, z;
  __device__ __forceinline__ float3 at(int i, int n) const {
    return p ? make_float3(p[i], p[n + i], p[2 * n + i]) : make_float3(x, y, z);
  }
};

// Value between two regions i != j (CuInterParameter::valueBetween,
// inter_parameter.hpp:71-90): uniform, or a lookup table over the pairs in
// lower-triangle order, i > j at i (i - 1) / 2 + j.
struct IP {
  const float* p;  // nullptr -> uniform
  float v;
  __host__ __device__ __forceinline__ static size_t index(unsigned i, unsigned j) {
    return i > j ? (size_t)i * (i - 1) / 2 + j : (size_t)j * (j - 1) / 2 + i;
  }
  __device__ __forceinline__ float between(unsigned i, unsigned j) const { return p ? p[index(i, j)] : v; }
};

// i mod d for i within a few periods of [0, d) (stencil offsets): repeated
// add / subtract instead of an integer division.
__host__ __device__ __forceinline__ int wrap_near(int i, int d) {
  while (i < 0) i += d;
  while (i >= d) i -= d;
  return i;
}

__host__ __device__ __forceinline__ int imod(int i, int d) {
  i %= d;
  return i >= 0 ? i : i + d;
}

__host__ __device__ __forceinline__ float3 f3(float a, float b, float c) { return make_float3(a, b, c); }
__host__ __device__ __forceinline__ float3 operator+(float3 a, float3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
__host__ __device__ __forceinline__ float3 operator-(float3 a, float3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__host__ __device__ __forceinline__ float3 operator*(float s, float3 a) { return f3(s * a.x, s * a.y, s * a.z); }
__host__ __device__ __forceinline__ float3 operator/(float3 a, float s) { return f3(a.x / s, a.y / s, a.z / s); }
__host__ __device__ __forceinline__ float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ __forceinline__ float3 cross3(float3 a, float3 b) {
  return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// Remap a linear workgroup index so that consecutive logical tiles land on
// the same XCD (8 XCDs; hardware dispatches workgroups round-robin over them).
// Each XCD then walks a contiguous range of the grid and its own L2 sees the
// stencil neighbours of the tiles it processes.
__device__ __forceinline__ int xcd_remap(int bid, int nblocks) {
  constexpr int NXCD = 8;
  int per = nblocks / NXCD;
  int full = per * NXCD;
  if (bid >= full) return bid;  // tail that does not divide evenly stays put
  return (bid % NXCD) * per + bid / NXCD;
}

}  // namespace mpa
