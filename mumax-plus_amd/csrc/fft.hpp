// Register-resident complex FFT of one line of L = 2^L2 points, computed by
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
// radix-4 pass when L2 is not a multiple of 3, then radix-8 passes.
//
// L2 is a template parameter, so every pass's radix, Ns and twiddle-table
// offset are compile-time constants.  The exchange after pass p stores
// element i at slot i + PM (i >> PS) (fft_pad_*), an affine function of the
// per-thread base, so every LDS access has an immediate offset.  PS and PM
// are chosen per exchange so that, with 32 banks of 4 bytes and 8-byte
// accesses served 16 lanes at a time, the 16 lanes of a group touch 16
// different bank pairs both when the butterflies store (element stride 8,
// or blocks of Ns consecutive elements) and when the next pass loads
// (consecutive elements): no bank conflicts for lines of 128 points and
// more.  Twiddles come
// from a per-pass table (stage_twiddles) holding, for each k < Ns, the R-1
// factors w^(r k) consecutively.  The transform is unnormalised; L is 1 (no
// work) or a power of two >= 8.
#pragma once

#include <vector>

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

// LDS slot i + M (i >> S): M pad slots every 2^S elements.
template <int S, int M>
__host__ __device__ constexpr int padsm(int i) { return i + M * (i >> S); }

// Synchronisation of an exchange between the T threads of one line.  Lines
// are aligned to wavefronts (thread t of line g is threadIdx.x = g*T + t), so
// with T <= 64 a line lives in one wavefront: its LDS accesses execute in
// issue order, and ordering them at compile time (a wavefront-scope fence
// around the wave barrier) is enough.  Longer lines span several wavefronts
// and need the block barrier.
template <int T>
__device__ __forceinline__ void line_sync() {
  if constexpr (T <= 64) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// Pass schedule of an L = 2^L2 transform.
__host__ __device__ constexpr int fft_npass(int L2) { return L2 < 3 ? 0 : L2 / 3 + (L2 % 3 ? 1 : 0); }
__host__ __device__ constexpr int fft_radix(int L2, int p) {
  return p == 0 ? (L2 % 3 == 1 ? 2 : (L2 % 3 == 2 ? 4 : 8)) : 8;
}
__host__ __device__ constexpr int log2i(int r) { return r == 8 ? 3 : (r == 4 ? 2 : (r == 2 ? 1 : 0)); }
__host__ __device__ constexpr int fft_l2ns(int L2, int p) {
  int s = 0;
  for (int q = 0; q < p; ++q) s += log2i(fft_radix(L2, q));
  return s;
}
__host__ __device__ constexpr int fft_twoff(int L2, int p) {
  int o = 0;
  for (int q = 0; q < p; ++q) o += (1 << fft_l2ns(L2, q)) * (fft_radix(L2, q) - 1);
  return o;
}
__host__ __device__ constexpr int fft_twsize(int L2) { return fft_twoff(L2, fft_npass(L2)); }

// Pad of the exchange after pass p (slot = i + PM (i >> PS)).  Lines of 128
// points and more: one slot per 16 elements after a pass with Ns = 1 or
// Ns >= 16, two after Ns = 2, 4 or 8.  Shorter lines share a lane group with
// other lines; their pads and line strides (Line::fft_target) come from the
// same bank model.
__host__ __device__ constexpr int fft_pad_shift(int L2, int p) { return (L2 == 4 || L2 == 6) ? 3 : 4; }
__host__ __device__ constexpr int fft_pad_mul(int L2, int p) {
  return (L2 >= 7 && fft_l2ns(L2, p) >= 1 && fft_l2ns(L2, p) <= 3) ? 2 : 1;
}

// Host: per-pass twiddle table of an L = 2^L2 transform (FP64 -> FP32).
inline std::vector<float2> stage_twiddles(int L2) {
  std::vector<float2> t((size_t)(fft_twsize(L2) > 0 ? fft_twsize(L2) : 1), make_float2(1.f, 0.f));
  for (int p = 0; p < fft_npass(L2); ++p) {
    const int R = fft_radix(L2, p), Ns = 1 << fft_l2ns(L2, p), off = fft_twoff(L2, p);
    for (int k = 0; k < Ns; ++k)
      for (int r = 1; r < R; ++r) {
        const double a = -2.0 * M_PI * (double)(r * k) / (double)(Ns * R);
        t[(size_t)off + (size_t)k * (R - 1) + (r - 1)] = make_float2((float)std::cos(a), (float)std::sin(a));
      }
  }
  return t;
}

template <int R, bool INV>
__device__ __forceinline__ void dft(float2* a) {
  if constexpr (R == 2) {
    const float2 t = a[1];
    a[1] = csub(a[0], t);
    a[0] = cadd(a[0], t);
  } else if constexpr (R == 4) {
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

// Radix-R butterfly m of component c in pass P: twiddle, DFT, and either keep
// the outputs (last pass) or store them into the line buffer `buf`.
template <int L2, int P, bool INV, int NC>
__device__ __forceinline__ void fft_butterfly(float2 (&v)[NC][8], int c, int m, const float2* w, int base,
                                              float2* buf) {
  constexpr int R = fft_radix(L2, P), M = 8 / R;
  constexpr int NS = 1 << fft_l2ns(L2, P);
  constexpr bool LAST = P == fft_npass(L2) - 1;
  constexpr int PS = fft_pad_shift(L2, P), PM = fft_pad_mul(L2, P);
  float2 a[R];
#pragma unroll
  for (int r = 0; r < R; ++r) a[r] = v[c][m + r * M];
  if constexpr (NS > 1) {
#pragma unroll
    for (int r = 1; r < R; ++r) a[r] = cmul(a[r], w[r]);
  }
  dft<R, INV>(a);
  if constexpr (LAST) {
#pragma unroll
    for (int r = 0; r < R; ++r) v[c][m + r * M] = a[r];
  } else {
#pragma unroll
    for (int r = 0; r < R; ++r) buf[base + r * NS + PM * ((r * NS) >> PS)] = a[r];
  }
}

// One Stockham pass for NC lines held by the same thread group, exchanged
// through NBUF line buffers at lds + b*bstride.  NBUF >= NC: all lines are
// stored, then read back (two barriers).  NBUF < NC: lines go one at a time,
// line i = P*NC + c through buffer i % NBUF; with NBUF >= 2 one barrier per
// line suffices, because a buffer is rewritten only after the barrier that
// follows the next line's stores.  The last pass keeps its outputs in
// registers.  Every thread of the block must call this; the barriers are
// line_sync (wavefront-level for lines of at most 64 threads).
template <int L2, int P, bool INV, int NC, int NBUF>
__device__ __forceinline__ void fft_pass(float2 (&v)[NC][8], int t, float2* lds, int bstride,
                                         const float2* __restrict__ tw) {
  constexpr int L = 1 << L2, T = L / 8;
  constexpr int R = fft_radix(L2, P), M = 8 / R, LOG2R = log2i(R);
  constexpr int L2NS = fft_l2ns(L2, P), NS = 1 << L2NS;
  constexpr int TWOFF = fft_twoff(L2, P);
  constexpr bool LAST = P == fft_npass(L2) - 1;
  constexpr int PS = fft_pad_shift(L2, P), PM = fft_pad_mul(L2, P);
  float2 w[M][R];
  int base[M];
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const int j = t + m * T;
    const int k = j & (NS - 1);
    if constexpr (NS > 1) {
      const float2* tp = tw + TWOFF + k * (R - 1);
#pragma unroll
      for (int r = 1; r < R; ++r) {
        w[m][r] = tp[r - 1];
        if (INV) w[m][r].y = -w[m][r].y;
      }
    }
    base[m] = padsm<PS, PM>(((j >> L2NS) << (L2NS + LOG2R)) + k);
  }
  const int rb = padsm<PS, PM>(t);
  if constexpr (LAST) {
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int m = 0; m < M; ++m) fft_butterfly<L2, P, INV, NC>(v, c, m, w[m], base[m], nullptr);
  } else if constexpr (NBUF >= NC) {
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int m = 0; m < M; ++m) fft_butterfly<L2, P, INV, NC>(v, c, m, w[m], base[m], lds + c * bstride);
    line_sync<T>();
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int q = 0; q < 8; ++q) v[c][q] = lds[c * bstride + rb + T * q + PM * ((T * q) >> PS)];
    line_sync<T>();
  } else {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      float2* buf = lds + ((P * NC + c) % NBUF) * bstride;
#pragma unroll
      for (int m = 0; m < M; ++m) fft_butterfly<L2, P, INV, NC>(v, c, m, w[m], base[m], buf);
      line_sync<T>();
#pragma unroll
      for (int q = 0; q < 8; ++q) v[c][q] = buf[rb + T * q + PM * ((T * q) >> PS)];
      if constexpr (NBUF == 1) line_sync<T>();
    }
  }
}

template <int L2, int P, bool INV, int NC, int NBUF>
__device__ __forceinline__ void fft_passes(float2 (&v)[NC][8], int t, float2* lds, int bstride,
                                           const float2* __restrict__ tw) {
  if constexpr (P < fft_npass(L2)) {
    fft_pass<L2, P, INV, NC, NBUF>(v, t, lds, bstride, tw);
    fft_passes<L2, P + 1, INV, NC, NBUF>(v, t, lds, bstride, tw);
  }
}

// Full transform: on entry v[c][q] holds element t + T*q of line c, on exit
// the transform's element t + T*q.  No work for L2 == 0.  Lines exchange
// through NBUF buffers of bstride float2 each (NBUF = NC: one per line).  On
// return every LDS access of the transform has completed block-wide, so the
// caller may reuse the buffers.
template <int L2, bool INV, int NC, int NBUF = NC>
__device__ __forceinline__ void line_fft(float2 (&v)[NC][8], int t, float2* lds, int bstride,
                                         const float2* __restrict__ tw) {
  static_assert(L2 == 0 || L2 >= 3, "line length must be 1 or >= 8");
  static_assert(NBUF >= 1, "at least one exchange buffer");
  fft_passes<L2, 0, INV, NC, NBUF>(v, t, lds, bstride, tw);
  // sequential exchanges through two or more buffers end without a barrier
  if constexpr (NBUF >= 2 && NBUF < NC && fft_npass(L2) > 1) line_sync<(1 << L2) / 8>();
}

}  // namespace mpa
