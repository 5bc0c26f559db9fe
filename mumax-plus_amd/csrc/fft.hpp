// Register-resident complex FFT of one line, cooperatively computed by
// T = L/8 threads, used by the demag convolution passes.
//
// Self-sorting Stockham formulation.  A pass of radix R with current
// sub-transform size Ns takes, for butterfly j, the elements j + r*L/R,
// multiplies element r by w^(r*(j mod Ns)) with w = exp(-2 pi i/(Ns R)),
// applies the R-point DFT and produces element r at
// ((j / Ns) * Ns * R) + (j mod Ns) + r*Ns.  Thread t owns butterflies
// j = t + m*T (m < 8/R), so its inputs are always the elements t + T*q
// (q = 0..7).  The first pass therefore reads straight from global memory,
// the outputs of the last pass are again the elements t + T*q, and only the
// passes in between exchange data through LDS.  Passes: one radix-2 or
// radix-4 pass when log2(L) is not a multiple of 3, then radix-8 passes.
// The transform is unnormalised.  L is 1 (no work) or a power of two >= 8.
#pragma once

#include "common.hpp"

namespace mpa {

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ float2 mul_mi(float2 a) {
  return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// LDS index with one pad slot every 16 elements (spreads the strided
// Stockham writes over the banks).
__device__ __forceinline__ int lpad(int i) { return i + (i >> 4); }
__host__ __device__ __forceinline__ int padded_len(int L) { return L + (L >> 4) + 1; }

template <int R, bool INV>
__device__ __forceinline__ void dft(float2* a) {
  if (R == 2) {
    const float2 t = a[1];
    a[1] = csub(a[0], t);
    a[0] = cadd(a[0], t);
  } else if (R == 4) {
    const float2 t0 = cadd(a[0], a[2]), t1 = csub(a[0], a[2]);
    const float2 t2 = cadd(a[1], a[3]), t3 = mul_mi<INV>(csub(a[1], a[3]));
    a[0] = cadd(t0, t2);
    a[2] = csub(t0, t2);
    a[1] = cadd(t1, t3);
    a[3] = csub(t1, t3);
  } else {  // R == 8: split into even / odd halves of two 4-point DFTs
    constexpr float S = 0.70710678118654752f;
    float2 b[4], d[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      b[k] = cadd(a[k], a[k + 4]);
      d[k] = csub(a[k], a[k + 4]);
    }
    // d_k *= W8^k (forward) or W8^-k (inverse)
    if (INV) {
      d[1] = make_float2(S * (d[1].x - d[1].y), S * (d[1].x + d[1].y));
      d[3] = make_float2(-S * (d[3].x + d[3].y), S * (d[3].x - d[3].y));
    } else {
      d[1] = make_float2(S * (d[1].x + d[1].y), S * (d[1].y - d[1].x));
      d[3] = make_float2(S * (d[3].y - d[3].x), -S * (d[3].x + d[3].y));
    }
    d[2] = mul_mi<INV>(d[2]);
    dft<4, INV>(b);
    dft<4, INV>(d);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      a[2 * k] = b[k];
      a[2 * k + 1] = d[k];
    }
  }
}

// One Stockham pass for NC lines held by the same thread group.  Line c of
// the group lives at lds + c*cstride (padded).  `last` passes keep their
// outputs in registers; the others exchange through LDS (two barriers, so
// every thread of the block must call this with the same arguments).
template <int R, bool INV, int NC>
__device__ __forceinline__ void fft_pass(float2 (&v)[NC][8], int t, int T, int log2L, int log2Ns, bool last,
                                         float2* lds, int cstride, const float2* __restrict__ tw) {
  constexpr int M = 8 / R;
  constexpr int LOG2R = R == 8 ? 3 : (R == 4 ? 2 : 1);
  const int Ns = 1 << log2Ns;
  const int twsh = log2L - log2Ns - LOG2R;
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const int j = t + m * T;
    const int k = j & (Ns - 1);
    float2 w[R];
#pragma unroll
    for (int r = 1; r < R; ++r) {
      w[r] = make_float2(1.f, 0.f);
      if (log2Ns > 0 && k != 0) {
        w[r] = tw[(r * k) << twsh];
        if (INV) w[r].y = -w[r].y;
      }
    }
    const int base = ((j >> log2Ns) << (log2Ns + LOG2R)) + k;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      float2 a[R];
#pragma unroll
      for (int r = 0; r < R; ++r) a[r] = v[c][m + r * M];
#pragma unroll
      for (int r = 1; r < R; ++r) a[r] = cmul(a[r], w[r]);
      dft<R, INV>(a);
      if (last) {
#pragma unroll
        for (int r = 0; r < R; ++r) v[c][m + r * M] = a[r];
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) lds[c * cstride + lpad(base + r * Ns)] = a[r];
      }
    }
  }
  if (!last) {
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int q = 0; q < 8; ++q) v[c][q] = lds[c * cstride + lpad(t + T * q)];
    __syncthreads();
  }
}

// Full transform: on entry v[c][q] holds element t + T*q of line c, on exit
// the transform's element t + T*q.  log2L >= 3.
template <bool INV, int NC>
__device__ __forceinline__ void line_fft(float2 (&v)[NC][8], int t, int log2L, float2* lds, int cstride,
                                         const float2* __restrict__ tw) {
  const int T = 1 << (log2L - 3);
  const int rem = log2L % 3;
  const int npass = log2L / 3 + (rem ? 1 : 0);
  int p = 0, log2Ns = 0;
  if (rem == 1) {
    fft_pass<2, INV, NC>(v, t, T, log2L, 0, npass == 1, lds, cstride, tw);
    log2Ns = 1;
    p = 1;
  } else if (rem == 2) {
    fft_pass<4, INV, NC>(v, t, T, log2L, 0, npass == 1, lds, cstride, tw);
    log2Ns = 2;
    p = 1;
  }
  for (; p < npass; ++p) {
    fft_pass<8, INV, NC>(v, t, T, log2L, log2Ns, p == npass - 1, lds, cstride, tw);
    log2Ns += 3;
  }
}

}  // namespace mpa
