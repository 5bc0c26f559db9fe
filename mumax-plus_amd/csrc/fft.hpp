// In-LDS complex FFT used by the demag convolution passes.
//
// Self-sorting Stockham formulation (radix-4 stages plus one radix-2 stage
// when log2(L) is odd).  A stage with current sub-transform size Ns reads, for
// butterfly j of a line, the elements j + r*L/R, multiplies element r by
// w^(r*(j mod Ns)) with w = exp(-2 pi i/(Ns R)), applies the R-point DFT and
// writes element r to ((j / Ns) * Ns * R) + (j mod Ns) + r*Ns.  Output is in
// natural order.  The transform is unnormalised.
//
// Each stage first loads every butterfly's inputs into registers, then
// synchronises, then writes, so the transform runs in place.  A thread holds
// at most E complex values: the caller guarantees nlines * L <= E * blockDim.
#pragma once

#include "common.hpp"

namespace mpa {

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }

// One Stockham stage of radix R over `nlines` lines of length L = 2^log2L.
template <int R, int E, bool INV>
__device__ __forceinline__ void fft_stage(float2* buf, int log2L, int log2Ns, int nlines, int lstride,
                                          const float2* __restrict__ tw) {
  constexpr int LOG2R = (R == 4) ? 2 : 1;
  constexpr int NB = E / R;  // butterflies per thread
  const int log2nbf = log2L - LOG2R;  // butterflies per line = L/R
  const int nbf = 1 << log2nbf;
  const int total = nbf * nlines;
  const int Ns = 1 << log2Ns;
  // twiddle table stride: w^(r*k) = tw[r*k*(L/(Ns*R))]
  const int twshift = log2L - log2Ns - LOG2R;

  float2 v[NB][R];
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int b = threadIdx.x + i * blockDim.x;
    if (b < total) {
      const int line = b >> log2nbf;
      const int j = b & (nbf - 1);
      const int k = j & (Ns - 1);
      const float2* src = buf + line * lstride;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        float2 a = src[j + r * nbf];
        if (r > 0 && k != 0) {
          float2 w = tw[(r * k) << twshift];
          if (INV) w.y = -w.y;
          a = cmul(a, w);
        }
        v[i][r] = a;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int b = threadIdx.x + i * blockDim.x;
    if (b < total) {
      const int line = b >> log2nbf;
      const int j = b & (nbf - 1);
      const int k = j & (Ns - 1);
      const int idxD = ((j >> log2Ns) << (log2Ns + LOG2R)) + k;
      float2* dst = buf + line * lstride + idxD;
      if (R == 2) {
        dst[0] = cadd(v[i][0], v[i][1]);
        dst[Ns] = csub(v[i][0], v[i][1]);
      } else {
        const float2 t0 = cadd(v[i][0], v[i][2]);
        const float2 t1 = csub(v[i][0], v[i][2]);
        const float2 t2 = cadd(v[i][1], v[i][3]);
        const float2 t3 = csub(v[i][1], v[i][3]);
        // forward: X1 = t1 - i t3, X3 = t1 + i t3; inverse swaps the signs
        const float2 mit3 = INV ? make_float2(-t3.y, t3.x) : make_float2(t3.y, -t3.x);
        dst[0] = cadd(t0, t2);
        dst[Ns] = cadd(t1, mit3);
        dst[2 * Ns] = csub(t0, t2);
        dst[3 * Ns] = csub(t1, mit3);
      }
    }
  }
  __syncthreads();
}

// Full transform of `nlines` lines (line l at buf + l*lstride).  Every thread
// of the block must call this.
template <int E, bool INV>
__device__ void lds_fft(float2* buf, int log2L, int nlines, int lstride, const float2* __restrict__ tw) {
  int log2Ns = 0;
  if (log2L & 1) {
    fft_stage<2, E, INV>(buf, log2L, log2Ns, nlines, lstride, tw);
    log2Ns += 1;
  }
  while (log2Ns < log2L) {
    fft_stage<4, E, INV>(buf, log2L, log2Ns, nlines, lstride, tw);
    log2Ns += 2;
  }
}

}  // namespace mpa
