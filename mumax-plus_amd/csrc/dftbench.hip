// Measurement behind a design choice of the demag FFTs: one radix-8 stage
// done as butterflies on the vector ALUs (the passes' own dft<8>), or posed
// as a dense real 16 x 16 matrix product on the matrix cores
// (v_mfma_f32_16x16x4f32).  Both variants read each round's inputs from LDS
// and write the outputs back, as a stage between two exchanges of a pass
// does; a block barrier separates rounds.  No counterpart in the reference.
#include "common.hpp"
#include "fft.hpp"

namespace mpa {

namespace {

typedef float f4v __attribute__((ext_vector_type(4)));

constexpr int DB_THREADS = 256;     // 4 wavefronts, one line per thread
constexpr int DB_LINE = 17;         // floats per line in LDS (16 and a pad)

__device__ __forceinline__ void load_lines(const float* __restrict__ in, float* s, size_t line0) {
  // the block's 256 lines of 16 floats, coalesced
  for (int e = threadIdx.x; e < DB_THREADS * 16; e += DB_THREADS)
    s[(e >> 4) * DB_LINE + (e & 15)] = in[line0 * 16 + e];
}

__device__ __forceinline__ void store_lines(float* __restrict__ out, const float* s, size_t line0) {
  for (int e = threadIdx.x; e < DB_THREADS * 16; e += DB_THREADS) out[line0 * 16 + e] = s[(e >> 4) * DB_LINE + (e & 15)];
}

// Each thread transforms its own line per round.
__global__ void __launch_bounds__(DB_THREADS) k_dft8_valu(const float* __restrict__ in, float* __restrict__ out,
                                                          int rounds) {
  __shared__ float s[DB_THREADS * DB_LINE];
  const size_t line0 = (size_t)blockIdx.x * DB_THREADS;
  load_lines(in, s, line0);
  __syncthreads();
  float* my = s + threadIdx.x * DB_LINE;
  for (int r = 0; r < rounds; ++r) {
    float2 a[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) a[q] = make_float2(my[2 * q], my[2 * q + 1]);
    dft<8, false>(a);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      my[2 * q] = a[q].x;
      my[2 * q + 1] = a[q].y;
    }
    __syncthreads();
  }
  store_lines(out, s, line0);
}

// Each wavefront transforms its 64 lines per round as four 16-line products
// D = In W^T with In (16 lines x 16 reals) and W the real form of the 8-point
// DFT (row 2o / 2o+1: real / imaginary part of output o; column 2n / 2n+1:
// real / imaginary part of input n).  Operand layout of the 16 x 16 x 4
// instruction: A lane l = row l % 16, k = l / 16; B lane l = column l % 16,
// k = l / 16; D lane l, register i = row 4 (l / 16) + i, column l % 16.
__global__ void __launch_bounds__(DB_THREADS) k_dft8_mfma(const float* __restrict__ in, float* __restrict__ out,
                                                          const float* __restrict__ W, int rounds) {
  __shared__ float s[DB_THREADS * DB_LINE];
  const size_t line0 = (size_t)blockIdx.x * DB_THREADS;
  load_lines(in, s, line0);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c16 = lane & 15, k4 = lane >> 4;
  float b[4];
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) b[kb] = W[c16 * 16 + 4 * kb + k4];  // W^T[4kb + k4][c16]
  __syncthreads();
  float* ws = s + wave * 64 * DB_LINE;
  for (int r = 0; r < rounds; ++r) {
    f4v acc[4];
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      acc[g] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < 4; ++kb) {
        const float a = ws[(g * 16 + c16) * DB_LINE + 4 * kb + k4];
        acc[g] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b[kb], acc[g], 0, 0, 0);
      }
    }
    __syncthreads();
#pragma unroll
    for (int g = 0; g < 4; ++g)
#pragma unroll
      for (int i = 0; i < 4; ++i) ws[(g * 16 + 4 * k4 + i) * DB_LINE + c16] = acc[g][i];
    __syncthreads();
  }
  store_lines(out, s, line0);
}

}  // namespace

double bench_dft8(int variant, int nlines, int rounds, const float* host_in, float* host_out) {
  if (variant < 0 || variant > 1) throw Error(MPA_ERR_INVALID, "dft8 bench: variant must be 0 (VALU) or 1 (MFMA)");
  if (nlines <= 0 || nlines % DB_THREADS != 0)
    throw Error(MPA_ERR_INVALID, "dft8 bench: nlines must be a positive multiple of 256");
  if (rounds < 1 || rounds > 40) throw Error(MPA_ERR_INVALID, "dft8 bench: rounds must be in 1..40");
  const size_t n = (size_t)nlines * 16;
  DeviceArray<float> din(n), dout(n), dw(256);
  HIP_CHECK(hipMemcpy(din.get(), host_in, n * sizeof(float), hipMemcpyHostToDevice));
  float w[256];
  for (int o = 0; o < 8; ++o)
    for (int q = 0; q < 8; ++q) {
      const double th = -2.0 * M_PI * (double)(o * q % 8) / 8.0;
      const float c = (float)std::cos(th), sn = (float)std::sin(th);
      // (re + i im)(c + i sn): re' = c re - sn im, im' = sn re + c im
      w[(2 * o) * 16 + 2 * q] = c;
      w[(2 * o) * 16 + 2 * q + 1] = -sn;
      w[(2 * o + 1) * 16 + 2 * q] = sn;
      w[(2 * o + 1) * 16 + 2 * q + 1] = c;
    }
  HIP_CHECK(hipMemcpy(dw.get(), w, sizeof(w), hipMemcpyHostToDevice));
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  const unsigned blocks = (unsigned)(nlines / DB_THREADS);
  auto run = [&]() {
    if (variant == 0) k_dft8_valu<<<blocks, DB_THREADS>>>(din.get(), dout.get(), rounds);
    else k_dft8_mfma<<<blocks, DB_THREADS>>>(din.get(), dout.get(), dw.get(), rounds);
  };
  run();  // warm-up
  HIP_CHECK(hipEventRecord(e0, nullptr));
  run();
  HIP_CHECK(hipEventRecord(e1, nullptr));
  HIP_CHECK(hipEventSynchronize(e1));
  checkLaunch("dft8 bench");
  float ms = 0.f;
  HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  HIP_CHECK(hipMemcpy(host_out, dout.get(), n * sizeof(float), hipMemcpyDeviceToHost));
  return (double)ms;
}

}  // namespace mpa
