// Demag (self stray field) of one magnet by zero-padded FFT convolution.
//
// Replaces StrayFieldFFTExecutor (src/physics/strayfieldfft.cu:149-246) and
// StrayFieldKernel (src/physics/strayfieldkernel.cu:11-99).
#pragma once

#include <vector>

#include "common.hpp"
#include "profiler.hpp"
#include "rk.hpp"

namespace mpa {

struct DemagGeometry {
  int n[3];          // magnet grid size (cells)
  double cs[3];      // cell size (m)
  int mastergrid[3]; // periodic box (0 = open axis)
  int pbc[3];        // image repetitions per axis
};

// FFT length used for an axis with n cells: 1 for n == 1, otherwise the
// smallest power of two >= 2n-1 (any length >= 2n-1 gives the same linear
// convolution; the reference uses 2n, strayfieldkernel.cu:84-96).
int demag_fft_length(int n);

class DemagPlan {
 public:
  explicit DemagPlan(const DemagGeometry& g);
  // h = -mu0 N * (Ms m) on the magnet grid; zero outside the geometry.
  // m, h: 3*N floats (SoA, x fastest).  geom: N bytes or nullptr.
  // prof (optional) times each pass with HIP events.
  // With m2 != nullptr the source is Ms m + Ms2 m2 (the two sublattices of an
  // antiferromagnet, strayfieldfft.cu:206-213).
  // The FFT intermediates A and B are the plan's own, allocated on the
  // first exec.
  void exec(const float* m, PS msat, const uint8_t* geom, float* h, hipStream_t s, Profiler* prof = nullptr,
            const float* m2 = nullptr, PS msat2 = PS{nullptr, 0.f});
  // exec of the Runge-Kutta stage state formed in the first pass:
  // x = normalise(x0 + sum_j co.c[j] co.k[j]), zero outside geom, is written
  // to x (3N floats, not aliasing x0 or any co.k[j]) and its field to h.
  void exec_combine(const float* x0, const Combo& co, float* x, PS msat, const uint8_t* geom, float* h,
                    hipStream_t s, Profiler* prof = nullptr);
  // The same transform with A and B in the library's static device scratch
  // array instead (scratch_capacity() bytes per device, shared by every call
  // on that device): allocates nothing.  Throws when workspace_bytes()
  // exceeds the scratch.  Concurrent calls on one device race on the scratch.
  void exec_noalloc(const float* m, PS msat, const uint8_t* geom, float* h, hipStream_t s);
  static size_t scratch_capacity();
  // Device bytes of the intermediates A (3 Nz Hx Ny complex) and, for a 3-D
  // grid, B (3 Hx Ky Nz complex), each 256-byte aligned.
  size_t workspace_bytes() const;
  // Allocates A and B if the plan does not hold them yet.
  void ensure_workspace();
  // Pass names and algorithmic HBM bytes per pass.
  int npasses() const { return is2d() ? 3 : 5; }
  const char* pass_name(int i) const;
  double pass_bytes(int i) const;

  const DemagGeometry& geometry() const { return g_; }
  int K(int a) const { return K_[a]; }
  bool is2d() const { return K_[2] == 1; }
  // Algorithmic HBM bytes of one exec (see the byte model in the design notes).
  double algorithmic_bytes() const;
  size_t device_bytes() const;
  // Real kernel spectrum table (Hx x Hy x Hz entries of 6 floats xx, yy, zz,
  // xy, xz, yz; pre-scaled), for tests.
  void copy_spectrum(float* host) const;
  int spectrum_dims(int a) const { return a == 0 ? Hx_ : (a == 1 ? Hy_ : Hz_); }

 private:
  friend class DemagBatch;
  void build_kernel(hipStream_t s);
  struct StageIn {
    const float* x0;
    const Combo* co;
    float* x;
  };
  void run(float2* A, float2* B, const float* m, PS msat, const uint8_t* geom, float* h, hipStream_t s,
           Profiler* prof, const float* m2, PS msat2, const StageIn* in = nullptr);
  size_t a_count() const;
  size_t b_count() const;
  DemagGeometry g_;
  int N_ = 0;
  int K_[3];
  int log2K_[3];
  int Hx_, Hy_, Hz_;
  DeviceArray<float2> A_, B_;
  DeviceArray<float> T_;
  DeviceArray<float2> tw_[3];
};

// Demag of many independent magnets at once, each with its own plan.  The
// magnetisations and fields are packed SoA arrays: component c of cell i of
// member j is at c * cs + off_j + i.  Every pass is one launch over all
// members, whatever their FFT line lengths (a block finds its member through
// a block table), so a batch of small magnets fills the GPU in a handful of
// launches instead of a few per magnet.
class DemagBatch {
 public:
  struct Member {
    DemagPlan* plan;
    long long off;
    PS msat;
    const uint8_t* geom;  // the member's own mask or nullptr
  };
  DemagBatch(const std::vector<Member>& members, long long cs);
  // H = demag field of X for every member (the plans' A, B buffers are the
  // scratch space).  With a second stream and two events, the single-layer
  // members' passes run on s2 concurrently with the others' (fork and join
  // through the events; works under graph capture).
  void exec(const float* X, float* H, hipStream_t s, Profiler* prof, hipStream_t s2 = nullptr,
            hipEvent_t fork = nullptr, hipEvent_t join = nullptr) const;
  // The passes of one group only (0: members with a z transform, 1: single
  // layers), on stream s.
  void exec_group(const float* X, float* H, int group, hipStream_t s, Profiler* prof) const;
  bool has_group(int group) const;
  double algorithmic_bytes() const;
  int nlaunches() const { return (int)launches_.size(); }
  // The launches' block tables, in launch order (each group's passes in
  // pass order), for the persistent step kernel.
  struct PassTable {
    int pass;   // 0 xfwd, 1 yfwd, 2 zpass, 3 yinv, 4 ypass2d, 5 xinv
    int group;  // 0: members with a z transform, 1: single layers
    unsigned nblocks;
    size_t shm;
    const void* segs;
    const int* b2s;
  };
  std::vector<PassTable> pass_tables() const;

 private:
  struct Launch {
    int pass;   // 0 xfwd, 1 yfwd, 2 zpass, 3 yinv, 4 ypass2d, 5 xinv
    int group;  // 0: members with a z transform, 1: single layers
    size_t shm;
    unsigned nblocks;
    DeviceArray<char> segs;
    DeviceArray<int> b2s;
    double bytes;
  };
  std::vector<Launch> launches_;
};

// One radix-8 stage of the FFT passes (8-point DFTs of nlines lines of 16
// floats, re/im interleaved) done `rounds` times in LDS, by vector-ALU
// butterflies (variant 0) or as a 16 x 16 matrix product on the matrix cores
// (variant 1); returns the kernel's milliseconds, the results in host_out.
double bench_dft8(int variant, int nlines, int rounds, const float* host_in, float* host_out);

// The real-space kernel exactly as the reference lays it out for a magnet
// (wrap_magnet.cpp:56-59 `_demag_kernel`): shape (6, Kz, Ky, Kx) on the
// reference kernel grid (strayfieldkernel.cu:80-99), computed in FP64 on the
// device and returned as float.
void reference_kernel(const DemagGeometry& g, int ksize[3], float* host_out);
void reference_kernel_size(const DemagGeometry& g, int ksize[3]);
// The six kernel components N(origin + j) for j on a ksize grid, FP64
// (with the images of g's periodic axes), stored as float (6, Kz, Ky, Kx) in
// device memory; g.n is not used.
void newell_grid_device(float* out, const int ksize[3], const int origin[3], const DemagGeometry& g, hipStream_t s);

}  // namespace mpa
