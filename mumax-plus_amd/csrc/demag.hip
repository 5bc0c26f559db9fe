// Demag convolution on MI355X: FP64 Newell spectrum build plus five
// hand-written FFT passes (three for a single-layer grid).
//
// Data path of one exec() for an Nx x Ny x Nz magnet padded to Kx x Ky x Kz
// (Hx = Kx/2+1, Hy = Ky/2+1, Hz = Kz/2+1), all complex buffers FP32:
//
//   xfwd : m (x-lines)          -> A[c][z][kx][y]   R2C of two real lines at a time,
//                                                   Ms*m and the zero pad fused in
//   yfwd : A (y-lines)          -> B[c][kx][ky][z]  only Ny inputs per line are non-zero
//   zpass: B (z-lines)          -> B                forward z, 3x3 real symmetric kernel
//                                                   multiply in registers, inverse z;
//                                                   keeps Nz outputs
//   yinv : B                    -> A                keeps Ny outputs
//   xinv : A                    -> h (x-lines)      C2R of two lines at a time, crop and
//                                                   geometry mask fused in
//
// For Kz == 1 the y pass does forward y, the kernel multiply and inverse y
// (ypass2d) and B is not used.  Every pass reads or writes whole lines
// contiguously on one side; the transposing side goes through LDS.  The
// line transforms are register-resident (fft.hpp): a 512-point line is 64
// threads x 8 values with two LDS exchanges.  Each pass is instantiated per
// line length, so thread counts, pass schedules and LDS strides are
// compile-time constants.  Since K >= 2N, the non-zero inputs of a forward
// transform and the kept outputs of an inverse one are all in the first
// K/2 points, i.e. in the first 4 of a thread's 8 values; the others are
// compile-time zeros or never computed.
//
// Kernel spectrum: the Newell kernel on offsets -(N-1)..(N-1) placed
// circularly is even (xx, yy, zz) or odd (xy: x,y; xz: x,z; yz: y,z) along
// each axis, so its DFT is real and only the octant kx<Hx, ky<Hy, kz<Hz is
// stored (6 floats per entry), with the sign of odd components flipped for
// ky > Ky/2 and kz > Kz/2.  It is built in FP64 as separable cosine / sine
// sums of the FP64 octant of the real-space kernel, pre-scaled by
// -mu0/(Kx Ky Kz) (the unnormalised inverse), and rounded to FP32 once.
#include "demag.hpp"
#include "batch_kernels.hpp"
#include "field_cell.hpp"
#include "fft.hpp"
#include "newell.hpp"
#include "rk.hpp"

#include <cmath>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

namespace mpa {

namespace {

constexpr double MU0 = 1.25663706212e-6;  // src/physics/constants.hpp:22
constexpr int BLOCK = 256;
constexpr int MAXL2 = 11;  // longest FFT line supported per axis: 2048 points
constexpr size_t LDS_LIMIT = 163840;  // LDS a gfx950 workgroup may allocate
// Block sizes of the line passes (threads), chosen by timing on MI355X.
#ifndef MPA_POINT_THREADS
#define MPA_POINT_THREADS 64
#endif
#ifndef MPA_XFWD_THREADS
#define MPA_XFWD_THREADS 512
#endif
#ifndef MPA_XINV_THREADS
#define MPA_XINV_THREADS 512
#endif
#ifndef MPA_YINV_THREADS
#define MPA_YINV_THREADS 512
#endif
#ifndef MPA_YFWD_THREADS
#define MPA_YFWD_THREADS 1024
#endif

int ilog2(int v) {
  int l = 0;
  while ((1 << l) < v) ++l;
  return l;
}

constexpr int clampc(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
constexpr int ilog2c(int v) { return v <= 1 ? 0 : 1 + ilog2c(v >> 1); }

struct DD {
  int nx, ny, nz;
  int Kx, Ky, Kz;
  int Hx, Hy, Hz;
  int l2x, l2y, l2z;
  int nlines;    // 3*ny*nz real x-lines
  long long cs;  // component stride of m and h (N, or the packed length of a batch)
};

// One magnet of a batched launch: its sizes and buffers, where its cells
// start in the packed state (component stride d.cs), and the first block of
// the launch that works on it.
struct BSeg {
  DD d;
  long long off;
  PS msat;
  const uint8_t* geom;  // the magnet's own mask (cell 0 at its first cell) or nullptr
  float2* A;
  float2* B;
  const float* T;
  const float2* tw[3];
  int first;
};

// Thread layout of one line of 2^L2 points (L2 == 0: a single point).
template <int L2>
struct Line {
  static_assert(L2 == 0 || (L2 >= 3 && L2 <= MAXL2), "unsupported line length");
  static constexpr int L = 1 << L2;
  static constexpr int L2T = L2 >= 3 ? L2 - 3 : 0;
  static constexpr int T = 1 << L2T;          // threads per line
  static constexpr int E = L2 >= 3 ? 8 : 1;   // values per thread
  static constexpr int EH = L2 >= 3 ? 4 : 1;  // values covering the first L/2 points
  static constexpr int H = L2 >= 3 ? L / 2 + 1 : 1;
  // LDS stride of one line: room for every pad of the transform and of the
  // staging (at most L + L/8 - 2 slots), rounded up to `target` modulo 32
  // float2 (target < 0: no rounding).
  static constexpr int stride(int target) {
    const int base = L2 >= 3 ? L + L / 8 - 1 : 1;
    return target < 0 ? base : base + (((target - base) % 32) + 32) % 32;
  }
  // Line stride for the transform's own exchanges: lines of fewer than 16
  // threads share a group of 16 lanes, and these residues modulo 32 keep
  // their stores and loads on different bank pairs (2 and 3 leave one 2-way
  // conflict per exchange); longer lines need none.
  static constexpr int fft_target() { return L2 == 4 ? 2 : (L2 == 5 ? 3 : (L2 == 6 ? 8 : -1)); }
  // Line stride when a block also reads its lines across (one element of
  // each of up to 16 lines per lane group): lines of 16 threads or more
  // take `across`, shorter ones the transform's own residue.
  static constexpr int target(int across) { return T >= 16 ? across : fft_target(); }
};

// ---------------------------------------------------------------- x forward

template <int L2>
struct XfwdCfg {
  using C = Line<L2>;
  static constexpr int PB = clampc(MPA_XFWD_THREADS / C::T, 1, 32);  // line pairs per block
  static constexpr int L2PB = ilog2c(PB);
  static constexpr int THREADS = PB * C::T;
  // the split reads one element of PB lines per lane group: odd stride
  static constexpr int LP = C::stride(C::target(1));
  // the lines, then the A offsets of the block's real lines
  static constexpr size_t DATA = (size_t)PB * LP * sizeof(float2);
  static constexpr size_t SHM = DATA + 2 * PB * sizeof(long long);
  static_assert(SHM <= LDS_LIMIT, "k_xfwd LDS");
};

// Batched launches run every line length with 256 threads per block, so
// one launch covers magnets of all sizes.
template <int L2>
struct XfwdCfgB {
  using C = Line<L2>;
  static constexpr int PB = 256 / C::T;
  static constexpr int L2PB = ilog2c(PB);
  static constexpr int THREADS = PB * C::T;
  static constexpr int LP = C::stride(C::target(1));
  static constexpr size_t DATA = (size_t)PB * LP * sizeof(float2);
  static constexpr size_t SHM = DATA + 2 * PB * sizeof(long long);
  static_assert(THREADS == 256 && SHM <= LDS_LIMIT, "batched k_xfwd");
};

// Slot of element i of a staged line (the stores after a transform and the
// loads before one go to consecutive elements): the pad of the transform's
// first exchange with one slot per group.
template <int L2>
__device__ __forceinline__ int iopad(int i) {
  return padsm<fft_pad_shift(L2 < 3 ? 3 : L2, 0), 1>(i);
}

// Block-wide transfer of TOTAL elements in two phases: each thread first
// issues all its reads (elements e = threadIdx.x + i THREADS), then all its
// writes, so its global loads (or LDS loads) are in flight together instead
// of one at a time.
template <int TOTAL, int THREADS, class Rd, class Wr>
__device__ __forceinline__ void staged_copy(const Rd& rd, const Wr& wr) {
  constexpr int NIT = (TOTAL + THREADS - 1) / THREADS;
  float2 buf[NIT];
#pragma unroll
  for (int i = 0; i < NIT; ++i) {
    const int e = threadIdx.x + i * THREADS;
    buf[i] = e < TOTAL ? rd(e) : make_float2(0.f, 0.f);
  }
#pragma unroll
  for (int i = 0; i < NIT; ++i) {
    const int e = threadIdx.x + i * THREADS;
    if (e < TOTAL) wr(e, buf[i]);
  }
}

// The split of PB complex lines (two real lines packed each) into their half
// spectra, stored to A[...][kx][y] at abase[li] + c_off: elements e = (kx, li)
// with li (the real line) fastest.
template <int L2, class K>
__device__ __forceinline__ void xfwd_split_store(const DD& d, const float2* lds, const long long* abase,
                                                 long long c_off, float2* __restrict__ A) {
  using C = Line<L2>;
  constexpr int L2NL = K::L2PB + 1;
  staged_copy<(C::H << L2NL), K::THREADS>(
      [&](int e) {
        const int kx = e >> L2NL, li = e & ((1 << L2NL) - 1);
        const float2* pl = lds + (li >> 1) * K::LP;
        const float2 z1 = pl[iopad<L2>(kx)];
        const float2 z2 = pl[iopad<L2>((C::L - kx) & (C::L - 1))];
        return (li & 1) ? make_float2(0.5f * (z1.y + z2.y), -0.5f * (z1.x - z2.x))
                        : make_float2(0.5f * (z1.x + z2.x), 0.5f * (z1.y - z2.y));
      },
      [&](int e, float2 X) {
        const int kx = e >> L2NL, li = e & ((1 << L2NL) - 1);
        const long long base = abase[li];
        if (base >= 0) A[c_off + base + (long long)kx * d.ny] = X;
      });
}

// Offset of x-line l (c, z, y) in A[c][z][kx][y] at kx = 0; -1 past the end.
__device__ __forceinline__ long long a_line_base(const DD& d, int l) {
  if (l >= d.nlines) return -1;
  const int yz = l / d.ny;
  return (long long)yz * d.Hx * d.ny + (l - yz * d.ny);
}

// PLAIN: no geometry mask and uniform Ms of both sublattices (geom, msat.p
// and msat2.p null; PLAIN 1: no second sublattice, PLAIN 2: m2 set), so the
// rows are loaded and scaled with nothing else live in registers; PLAIN 0
// takes any magnet.
template <int L2, class K = XfwdCfg<L2>, int PLAIN = 0>
__device__ __forceinline__ void xfwd_body(const DD& d, const float* __restrict__ m, PS msat,
                                          const float* __restrict__ m2, PS msat2, const uint8_t* __restrict__ geom,
                                          float2* __restrict__ A, const float2* __restrict__ tw, int blk) {
  using C = Line<L2>;
  extern __shared__ float2 lds[];
  long long* abase = reinterpret_cast<long long*>(lds + K::DATA / sizeof(float2));
  const int g = threadIdx.x >> C::L2T, t = threadIdx.x & (C::T - 1);
  const int l0 = 2 * K::PB * blk;
  for (int i = threadIdx.x; i < 2 * K::PB; i += K::THREADS) abase[i] = a_line_base(d, l0 + i);
  const int la = l0 + 2 * g, lb = la + 1;
  const int nyz = d.ny * d.nz;
  const bool va = la < d.nlines, vb = lb < d.nlines;
  const int cella = (la % nyz) * d.nx, cellb = (lb % nyz) * d.nx;
  // element offsets of the two rows; a row past the end reads row 0
  const size_t ia = va ? (size_t)(la / nyz) * d.cs + cella : 0, ib = vb ? (size_t)(lb / nyz) * d.cs + cellb : 0;
  const int ca = va ? cella : 0, cb = vb ? cellb : 0;
  // M = Ms m, or Ms1 m1 + Ms2 m2 for the two sublattices of an antiferromagnet
  // (k_pad, strayfieldfft.cu:51-73), zero outside the geometry.  Every load
  // is issued unconditionally from an in-range address (x clamped to 0 past
  // nx) and the row, x and geometry tests select afterwards, so all of a
  // thread's loads are in flight together.
  float ma[C::EH], mb[C::EH], m2a[C::EH], m2b[C::EH];
  if (d.nx >= C::T * C::EH) {
    // rows as long as the half line: fixed offsets from one address per row
#pragma unroll
    for (int q = 0; q < C::EH; ++q) {
      ma[q] = m[ia + t + C::T * q];
      mb[q] = m[ib + t + C::T * q];
      if constexpr (PLAIN == 2) {
        m2a[q] = m2[ia + t + C::T * q];
        m2b[q] = m2[ib + t + C::T * q];
      }
    }
  } else {
#pragma unroll
    for (int q = 0; q < C::EH; ++q) {
      const int x = t + C::T * q, xs = x < d.nx ? x : 0;
      ma[q] = m[ia + xs];
      mb[q] = m[ib + xs];
      if constexpr (PLAIN == 2) {
        m2a[q] = m2[ia + xs];
        m2b[q] = m2[ib + xs];
      }
    }
  }
  bool ina[C::EH], inb[C::EH];
  if constexpr (PLAIN != 0) {
#pragma unroll
    for (int q = 0; q < C::EH; ++q) {
      const int x = t + C::T * q;
      ina[q] = va && x < d.nx;
      inb[q] = vb && x < d.nx;
      ma[q] *= msat.v;
      mb[q] *= msat.v;
      if constexpr (PLAIN == 2) {
        ma[q] += msat2.v * m2a[q];
        mb[q] += msat2.v * m2b[q];
      }
    }
  } else {
    // geometry flags as bits (q: row a, 8 + q: row b)
    unsigned gm = 0xFFFFu;
    if (geom) {
      gm = 0;
#pragma unroll
      for (int q = 0; q < C::EH; ++q) {
        const int x = t + C::T * q, xs = x < d.nx ? x : 0;
        gm |= (geom[ca + xs] ? 1u : 0u) << q;
        gm |= (geom[cb + xs] ? 1u : 0u) << (8 + q);
      }
    }
    if (msat.p) {
#pragma unroll
      for (int q = 0; q < C::EH; ++q) {
        const int x = t + C::T * q, xs = x < d.nx ? x : 0;
        ma[q] *= msat.p[ca + xs];
        mb[q] *= msat.p[cb + xs];
      }
    } else {
#pragma unroll
      for (int q = 0; q < C::EH; ++q) {
        ma[q] *= msat.v;
        mb[q] *= msat.v;
      }
    }
#pragma unroll
    for (int q = 0; q < C::EH; ++q) {
      const int x = t + C::T * q;
      ina[q] = va && x < d.nx && ((gm >> q) & 1u);
      inb[q] = vb && x < d.nx && ((gm >> (8 + q)) & 1u);
    }
  }
  if (PLAIN == 0 && m2) {
#pragma unroll
    for (int q = 0; q < C::EH; ++q) {
      const int x = t + C::T * q, xs = x < d.nx ? x : 0;
      ma[q] += msat2.at(ca + xs) * m2[ia + xs];
      mb[q] += msat2.at(cb + xs) * m2[ib + xs];
    }
  }
  float2 v[1][8];
#pragma unroll
  for (int q = 0; q < 8; ++q) v[0][q] = make_float2(0.f, 0.f);
#pragma unroll
  for (int q = 0; q < C::EH; ++q) v[0][q] = make_float2(ina[q] ? ma[q] : 0.f, inb[q] ? mb[q] : 0.f);
  float2* line = lds + g * K::LP;
  line_fft<L2, false, 1>(v, t, line, 0, tw);
#pragma unroll
  for (int q = 0; q < C::E; ++q) line[iopad<L2>(t + C::T * q)] = v[0][q];
  __syncthreads();
  xfwd_split_store<L2, K>(d, lds, abase, 0, A);
}

template <int L2, int PLAIN>
__global__ void __launch_bounds__(XfwdCfg<L2>::THREADS)
    k_xfwd(DD d, const float* __restrict__ m, PS msat, const float* __restrict__ m2, PS msat2,
           const uint8_t* __restrict__ geom, float2* __restrict__ A, const float2* __restrict__ tw) {
  xfwd_body<L2, XfwdCfg<L2>, PLAIN>(d, m, msat, m2, msat2, geom, A, tw, blockIdx.x);
}

// x forward with the Runge-Kutta stage combination in front.  The rows of
// the magnet are r = z ny + y; thread group g of block blk takes rows
// ra = 2 (PB blk + g) and ra + 1.  For its cells it forms the stage state
// x = normalise(x0 + sum_j c_j k_j), zero outside the geometry (the
// arithmetic of k_combine_batch), writes it as the new state and keeps it in
// registers; then it transforms Ms x one component at a time (two real rows
// per complex transform, as k_xfwd) and the block stores each component's
// half spectra into A.  The state is thus read once for both the combination
// and the transform.
template <int L2, class K = XfwdCfg<L2>>
__device__ __forceinline__ void xfwd_combine_body(const DD& d, const float* __restrict__ x0, const Combo& co,
                                                  float* __restrict__ xout, PS msat, const uint8_t* __restrict__ geom,
                                                  float2* __restrict__ A, const float2* __restrict__ tw, int blk) {
  using C = Line<L2>;
  extern __shared__ float2 lds[];
  long long* abase = reinterpret_cast<long long*>(lds + K::DATA / sizeof(float2));
  const int nrows = d.ny * d.nz;
  const int g = threadIdx.x >> C::L2T, t = threadIdx.x & (C::T - 1);
  const int r0 = 2 * K::PB * blk;
  for (int i = threadIdx.x; i < 2 * K::PB; i += K::THREADS) {
    const int r = r0 + i;
    abase[i] = r < nrows ? (long long)(r / d.ny) * d.Hx * d.ny + (r % d.ny) : -1;
  }
  const size_t N = (size_t)d.cs;
  const int ra = r0 + 2 * g;
  // stage state of the group's cells: s[row][c][q]
  float sv[2][3][C::EH];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int row = ra + h;
#pragma unroll
    for (int q = 0; q < C::EH; ++q) {
      const int x = t + C::T * q;
      float v3[3] = {0.f, 0.f, 0.f};
      if (row < nrows && x < d.nx) {
        const size_t idx = (size_t)row * d.nx + x;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          float a = x0[c * N + idx];
          for (int j = 0; j < co.nk; ++j) a += co.c[j] * co.k[j][c * N + idx];
          v3[c] = a;
        }
        const float inv = rsqrtf(v3[0] * v3[0] + v3[1] * v3[1] + v3[2] * v3[2]);
#pragma unroll
        for (int c = 0; c < 3; ++c) v3[c] *= inv;
        if (geom && !geom[idx]) v3[0] = v3[1] = v3[2] = 0.f;
#pragma unroll
        for (int c = 0; c < 3; ++c) xout[c * N + idx] = v3[c];
        const float ms = msat.at(idx);
#pragma unroll
        for (int c = 0; c < 3; ++c) v3[c] *= ms;
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) sv[h][c][q] = v3[c];
    }
  }
  float2* line = lds + g * K::LP;
  const long long cstride = (long long)d.nz * d.Hx * d.ny;  // component stride of A
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float2 v[1][8];
#pragma unroll
    for (int q = 0; q < 8; ++q) v[0][q] = q < C::EH ? make_float2(sv[0][c][q], sv[1][c][q]) : make_float2(0.f, 0.f);
    line_fft<L2, false, 1>(v, t, line, 0, tw);
#pragma unroll
    for (int q = 0; q < C::E; ++q) line[iopad<L2>(t + C::T * q)] = v[0][q];
    __syncthreads();
    xfwd_split_store<L2, K>(d, lds, abase, c * cstride, A);
    __syncthreads();
  }
}

template <int L2>
__global__ void __launch_bounds__(XfwdCfg<L2>::THREADS)
    k_xfwd_combine(DD d, const float* __restrict__ x0, Combo co, float* __restrict__ xout, PS msat,
                   const uint8_t* __restrict__ geom, float2* __restrict__ A, const float2* __restrict__ tw) {
  xfwd_combine_body<L2>(d, x0, co, xout, msat, geom, A, tw, blockIdx.x);
}

// ---------------------------------------------------------------- y forward / inverse (3-D)

template <int L2>
struct YCfg {
  using C = Line<L2>;
  // lines (c, kx, z) per block, z fastest: at least 16 (128-byte runs of z on
  // the strided B side) while a block stays within 1024 threads
  static constexpr int LB = C::T >= 64 ? clampc(MPA_YFWD_THREADS / C::T, 1, 16) : clampc(512 / C::T, 1, 32);
  static constexpr int L2LB = ilog2c(LB);
  static constexpr int THREADS = LB * C::T;
  // the transposed store reads one element of each of 16 lines per lane
  // group: odd stride
  static constexpr int LP = C::stride(C::target(1));
  // the lines, then the B offsets of the block's lines
  static constexpr size_t DATA = (size_t)LB * LP * sizeof(float2);
  static constexpr size_t SHM = DATA + LB * sizeof(long long);
  static_assert(SHM <= LDS_LIMIT, "k_yfwd LDS");
};

// The y inverse reads B with LB consecutive z per (c, kx) and writes A lines.
template <int L2>
struct YinvCfg {
  using C = Line<L2>;
  static constexpr int LB = clampc(MPA_YINV_THREADS / C::T, 1, 32);
  static constexpr int L2LB = ilog2c(LB);
  static constexpr int THREADS = LB * C::T;
  // the staging writes 8 lines x 2 consecutive elements per lane group:
  // stride 2 modulo 16
  static constexpr int LP = C::stride(C::target(2));
  static constexpr size_t DATA = (size_t)LB * LP * sizeof(float2);
  static constexpr size_t SHM = DATA + LB * sizeof(long long);
  static_assert(SHM <= LDS_LIMIT, "k_yinv LDS");
};

template <int L2>
struct YCfgB {
  using C = Line<L2>;
  static constexpr int LB = 256 / C::T;
  static constexpr int L2LB = ilog2c(LB);
  static constexpr int THREADS = LB * C::T;
  static constexpr int LP = C::stride(C::target(1));
  static constexpr size_t DATA = (size_t)LB * LP * sizeof(float2);
  static constexpr size_t SHM = DATA + LB * sizeof(long long);
  static_assert(THREADS == 256 && SHM <= LDS_LIMIT, "batched k_yfwd");
};

// Offset of y-line q = (c, kx, z) in B[c][kx][ky][z] at ky = 0; -1 past the end.
__device__ __forceinline__ long long b_line_base(const DD& d, int q) {
  if (q >= 3 * d.Hx * d.nz) return -1;
  const int ckx = q / d.nz, z = q - ckx * d.nz;
  return (long long)ckx * d.Ky * d.nz + z;
}

// Start of y-line q = (c, kx, z) in A[c][z][kx][y].
__device__ __forceinline__ size_t a_yline(const DD& d, int q) {
  const int ckx = q / d.nz, z = q - ckx * d.nz;
  const int c = ckx / d.Hx, kx = ckx - c * d.Hx;
  return ((size_t)(c * d.nz + z) * d.Hx + kx) * d.ny;
}

template <int L2, class K = YCfg<L2>>
__device__ __forceinline__ void yfwd_body(const DD& d, const float2* __restrict__ A, float2* __restrict__ B,
                                          const float2* __restrict__ tw, int blk) {
  using C = Line<L2>;
  extern __shared__ float2 lds[];
  long long* bbase = reinterpret_cast<long long*>(lds + K::DATA / sizeof(float2));
  const int nq = 3 * d.Hx * d.nz;
  const int g = threadIdx.x >> C::L2T, t = threadIdx.x & (C::T - 1);
  const int q0 = blk * K::LB;
  for (int i = threadIdx.x; i < K::LB; i += K::THREADS) bbase[i] = b_line_base(d, q0 + i);
  const int ql = q0 + g;
  float2 v[1][8];
#pragma unroll
  for (int q = 0; q < 8; ++q) v[0][q] = make_float2(0.f, 0.f);
  if (ql < nq) {
    const float2* src = A + a_yline(d, ql);
#pragma unroll
    for (int q = 0; q < C::EH; ++q) {
      const int y = t + C::T * q;
      if (y < d.ny) v[0][q] = src[y];
    }
  }
  float2* line = lds + g * K::LP;
  line_fft<L2, false, 1>(v, t, line, 0, tw);
#pragma unroll
  for (int q = 0; q < C::E; ++q) line[iopad<L2>(t + C::T * q)] = v[0][q];
  __syncthreads();
  staged_copy<(C::L << K::L2LB), K::THREADS>(
      [&](int e) {
        const int ky = e >> K::L2LB, li = e & (K::LB - 1);
        return lds[li * K::LP + iopad<L2>(ky)];
      },
      [&](int e, float2 v) {
        const int ky = e >> K::L2LB, li = e & (K::LB - 1);
        const long long base = bbase[li];
        if (base >= 0) B[base + (long long)ky * d.nz] = v;
      });
}

template <int L2>
__global__ void __launch_bounds__(YCfg<L2>::THREADS)
    k_yfwd(DD d, const float2* __restrict__ A, float2* __restrict__ B, const float2* __restrict__ tw) {
  yfwd_body<L2>(d, A, B, tw, xcd_remap(blockIdx.x, gridDim.x));
}

template <int L2, class K = YinvCfg<L2>>
__device__ __forceinline__ void yinv_body(const DD& d, const float2* __restrict__ B, float2* __restrict__ A,
                                          const float2* __restrict__ tw, int blk) {
  using C = Line<L2>;
  extern __shared__ float2 lds[];
  long long* bbase = reinterpret_cast<long long*>(lds + K::DATA / sizeof(float2));
  const int nq = 3 * d.Hx * d.nz;
  const int q0 = blk * K::LB;
  for (int i = threadIdx.x; i < K::LB; i += K::THREADS) bbase[i] = b_line_base(d, q0 + i);
  __syncthreads();
  staged_copy<(C::L << K::L2LB), K::THREADS>(
      [&](int e) {
        const int ky = e >> K::L2LB, li = e & (K::LB - 1);
        const long long base = bbase[li];
        return base >= 0 ? B[base + (long long)ky * d.nz] : make_float2(0.f, 0.f);
      },
      [&](int e, float2 v) {
        const int ky = e >> K::L2LB, li = e & (K::LB - 1);
        lds[li * K::LP + iopad<L2>(ky)] = v;
      });
  __syncthreads();
  const int g = threadIdx.x >> C::L2T, t = threadIdx.x & (C::T - 1);
  float2* line = lds + g * K::LP;
  float2 v[1][8];
#pragma unroll
  for (int q = 0; q < 8; ++q) v[0][q] = q < C::E ? line[iopad<L2>(t + C::T * q)] : make_float2(0.f, 0.f);
  __syncthreads();
  line_fft<L2, true, 1>(v, t, line, 0, tw);
  const int ql = q0 + g;
  if (ql < nq) {
    float2* dst = A + a_yline(d, ql);
#pragma unroll
    for (int q = 0; q < C::EH; ++q) {
      const int y = t + C::T * q;
      if (y < d.ny) dst[y] = v[0][q];
    }
  }
}

template <int L2>
__global__ void __launch_bounds__(YinvCfg<L2>::THREADS)
    k_yinv(DD d, const float2* __restrict__ B, float2* __restrict__ A, const float2* __restrict__ tw) {
  yinv_body<L2>(d, B, A, tw, xcd_remap(blockIdx.x, gridDim.x));
}

// ---------------------------------------------------------------- kernel multiply

// h = T . m at one spectral point.  `e` points at the octant entry (six
// adjacent floats xx, yy, zz, xy, xz, yz of the spectrum table
// T[Hx][Hy][Hz][6]); sy and sz are the signs of the components odd in y and
// z at the point (-1 where ky > Ky/2 or kz > Kz/2; at Ky/2 and Kz/2 those
// components are zero, so either sign is right there).
template <bool FLAT>
__device__ __forceinline__ void kernel_multiply(const float* __restrict__ e, float sy, float sz, float2& a,
                                                float2& b, float2& c) {
  const float2* e2 = reinterpret_cast<const float2*>(e);  // entries start on 8-byte boundaries
  const float2 t0 = e2[0], t1 = e2[1];
  const float nxx = t0.x, nyy = t0.y, nzz = t1.x;
  const float nxy = sy * t1.y;
  float nxz = 0.f, nyz = 0.f;  // single layer: xz and yz vanish
  if constexpr (!FLAT) {
    const float2 t2 = e2[2];
    nxz = sz * t2.x;
    nyz = sy * sz * t2.y;
  }
  const float2 ma = a, mb = b, mc = c;
  a = make_float2(nxx * ma.x + nxy * mb.x + nxz * mc.x, nxx * ma.y + nxy * mb.y + nxz * mc.y);
  b = make_float2(nxy * ma.x + nyy * mb.x + nyz * mc.x, nxy * ma.y + nyy * mb.y + nyz * mc.y);
  c = make_float2(nxz * ma.x + nyz * mb.x + nzz * mc.x, nxz * ma.y + nyz * mb.y + nzz * mc.y);
}

// ---------------------------------------------------------------- z pass (3-D) and 2-D y pass

template <int L2>
struct PointCfg {
  using C = Line<L2>;
  static constexpr int PT = clampc(MPA_POINT_THREADS / C::T, 1, 256);  // spectral points per block
  static constexpr int THREADS = PT * C::T;
  static constexpr int LP = C::stride(C::fft_target());
  static constexpr int NBUF = 2;  // exchange buffers per line for the three components
  static constexpr size_t SHM = (size_t)NBUF * PT * LP * sizeof(float2);
  static_assert(SHM <= LDS_LIMIT - 1024, "point pass LDS");
};

// Batched launches: 256 threads for every line length.
template <int L2>
struct PointCfgB {
  using C = Line<L2>;
  static constexpr int PT = clampc(256 / C::T, 1, 256);  // spectral points per block
  static constexpr int THREADS = PT * C::T;
  static constexpr int LP = C::stride(C::fft_target());
  static constexpr int NBUF = 2;  // exchange buffers per line for the three components
  static constexpr size_t SHM = (size_t)NBUF * PT * LP * sizeof(float2);
  static_assert(SHM <= LDS_LIMIT - 1024, "batched point pass LDS");
};

// PT spectral points (kx, ky) per block; each thread holds the three
// components of its 8 z values, so the kernel multiply between the forward
// and the inverse transform stays in registers.  Adjacent thread groups take
// the mirror points ky and Ky - ky (0 pairs with Ky/2), which read the same
// row of the octant spectrum, so each row comes from HBM once.
template <int L2, class K = PointCfg<L2>>
__device__ __forceinline__ void zpass_body(const DD& d, float2* __restrict__ B, const float* __restrict__ Tk,
                                           const float2* __restrict__ tw, int blk) {
  using C = Line<L2>;
  extern __shared__ float2 lds[];
  const int npt = d.Hx * d.Ky;
  const size_t cs = (size_t)npt;  // points per component
  const int g = threadIdx.x >> C::L2T, t = threadIdx.x & (C::T - 1);
  const int pt = blk * K::PT + g;  // pair-ordered point index
  const bool valid = pt < npt;
  const int kx = pt >> d.l2y, r = pt & (d.Ky - 1), kh = r >> 1;
  const int ky = (r & 1) ? (kh == 0 ? d.Ky >> 1 : d.Ky - kh) : kh;
  const size_t line0 = (size_t)kx * d.Ky + ky;  // z-line of (kx, ky) in component 0
  float2 v[3][8];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int q = 0; q < 8; ++q) v[c][q] = make_float2(0.f, 0.f);
  if (valid) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int q = 0; q < C::EH; ++q) {
        const int z = t + C::T * q;
        if (z < d.nz) v[c][q] = B[((size_t)c * cs + line0) * d.nz + z];
      }
  }
  float2* line = lds + g * K::LP;
  line_fft<L2, false, 3, K::NBUF>(v, t, line, K::PT * K::LP, tw);
  if (valid) {
    // values q < EH have kz < Kz/2; the others mirror to Kz - kz with the odd-z sign flipped
    const bool yflip = ky > (d.Ky >> 1);
    const float sy = yflip ? -1.f : 1.f;
    const float* row = Tk + ((size_t)kx * d.Hy + (yflip ? d.Ky - ky : ky)) * d.Hz * 6;
#pragma unroll
    for (int q = 0; q < C::E; ++q) {
      const int kz = t + C::T * q;
      const bool upper = q >= C::EH;
      kernel_multiply<false>(row + 6 * (upper ? C::L - kz : kz), sy, upper ? -1.f : 1.f, v[0][q], v[1][q], v[2][q]);
    }
  }
  line_fft<L2, true, 3, K::NBUF>(v, t, line, K::PT * K::LP, tw);
  if (valid) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int q = 0; q < C::EH; ++q) {
        const int z = t + C::T * q;
        if (z < d.nz) B[((size_t)c * cs + line0) * d.nz + z] = v[c][q];
      }
  }
}

template <int L2>
__global__ void __launch_bounds__(PointCfg<L2>::THREADS)
    k_zpass(DD d, float2* __restrict__ B, const float* __restrict__ Tk, const float2* __restrict__ tw) {
  zpass_body<L2>(d, B, Tk, tw, blockIdx.x);
}

// Single layer: forward y, kernel multiply, inverse y; PT points kx per block.
template <int L2, class K = PointCfg<L2>>
__device__ __forceinline__ void ypass2d_body(const DD& d, float2* __restrict__ A, const float* __restrict__ Tk,
                                             const float2* __restrict__ tw, int blk) {
  using C = Line<L2>;
  extern __shared__ float2 lds[];
  const int g = threadIdx.x >> C::L2T, t = threadIdx.x & (C::T - 1);
  const int kx = blk * K::PT + g;
  const bool valid = kx < d.Hx;
  float2 v[3][8];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int q = 0; q < 8; ++q) v[c][q] = make_float2(0.f, 0.f);
  if (valid) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int q = 0; q < C::EH; ++q) {
        const int y = t + C::T * q;
        if (y < d.ny) v[c][q] = A[((size_t)c * d.Hx + kx) * d.ny + y];
      }
  }
  float2* line = lds + g * K::LP;
  line_fft<L2, false, 3, K::NBUF>(v, t, line, K::PT * K::LP, tw);
  if (valid) {
    const float* row = Tk + (size_t)kx * d.Hy * 6;  // Hz == 1
#pragma unroll
    for (int q = 0; q < C::E; ++q) {
      const int ky = t + C::T * q;
      const bool upper = q >= C::EH;  // ky >= Ky/2: mirror with the odd-y sign flipped
      kernel_multiply<true>(row + 6 * (upper ? C::L - ky : ky), upper ? -1.f : 1.f, 1.f, v[0][q], v[1][q], v[2][q]);
    }
  }
  line_fft<L2, true, 3, K::NBUF>(v, t, line, K::PT * K::LP, tw);
  if (valid) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int q = 0; q < C::EH; ++q) {
        const int y = t + C::T * q;
        if (y < d.ny) A[((size_t)c * d.Hx + kx) * d.ny + y] = v[c][q];
      }
  }
}

template <int L2>
__global__ void __launch_bounds__(PointCfg<L2>::THREADS)
    k_ypass2d(DD d, float2* __restrict__ A, const float* __restrict__ Tk, const float2* __restrict__ tw) {
  ypass2d_body<L2>(d, A, Tk, tw, blockIdx.x);
}

// ---------------------------------------------------------------- x inverse

template <int L2>
struct XinvCfg {
  using C = Line<L2>;
  // line pairs per block: 16 lines give 128-byte runs on the strided A side
  static constexpr int PB = clampc(MPA_XINV_THREADS / C::T, 1, 32);
  static constexpr int L2PB = ilog2c(PB);
  static constexpr int THREADS = PB * C::T;
  static constexpr int XS = C::H;  // stride of a staged half spectrum (odd: no bank conflicts)
  static constexpr int LP = C::stride(C::fft_target());
  // the staged half spectra and the transform buffers share the LDS
  static constexpr size_t XB = (size_t)2 * PB * XS, ZB = (size_t)PB * LP;
  static constexpr size_t DATA = (XB > ZB ? XB : ZB) * sizeof(float2);
  static constexpr size_t SHM = DATA + 2 * PB * sizeof(long long);  // then the A offsets of the lines
  static_assert(SHM <= LDS_LIMIT, "k_xinv LDS");
};

template <int L2>
struct XinvCfgB {
  using C = Line<L2>;
  static constexpr int PB = 256 / C::T;
  static constexpr int L2PB = ilog2c(PB);
  static constexpr int THREADS = PB * C::T;
  static constexpr int XS = C::H;
  static constexpr int LP = C::stride(C::fft_target());
  static constexpr size_t XB = (size_t)2 * PB * XS, ZB = (size_t)PB * LP;
  static constexpr size_t DATA = (XB > ZB ? XB : ZB) * sizeof(float2);
  static constexpr size_t SHM = DATA + 2 * PB * sizeof(long long);
  static_assert(THREADS == 256 && SHM <= LDS_LIMIT, "batched k_xinv");
};

// PB line pairs per block; C2R of two lines at a time.
template <int L2, class K = XinvCfg<L2>>
__device__ __forceinline__ void xinv_body(const DD& d, const float2* __restrict__ A, const uint8_t* __restrict__ geom,
                                          float* __restrict__ h, const float2* __restrict__ tw, int blk) {
  using C = Line<L2>;
  extern __shared__ float2 lds[];
  long long* abase = reinterpret_cast<long long*>(lds + K::DATA / sizeof(float2));
  constexpr int L2NL = K::L2PB + 1;
  constexpr int NL = 2 * K::PB;
  float2* X = lds;
  float2* Z = lds;  // reused once the half spectra are in registers
  const int l0 = 2 * blk * K::PB;
  for (int i = threadIdx.x; i < NL; i += K::THREADS) abase[i] = a_line_base(d, l0 + i);
  // geometry flags of the thread's output cells (bit q: row a, bit 8 + q:
  // row b), loaded together up front so their latency hides behind the
  // transform (x clamped to 0 past nx: those cells are not stored)
  unsigned gmask = 0xFFFFu;
  if (geom) {
    const int g0 = threadIdx.x >> C::L2T, t0 = threadIdx.x & (C::T - 1);
    const int nyz0 = d.ny * d.nz;
    const int ca = ((l0 + 2 * g0) % nyz0) * d.nx, cb = ((l0 + 2 * g0 + 1) % nyz0) * d.nx;
    gmask = 0;
#pragma unroll
    for (int q = 0; q < C::EH; ++q) {
      const int x = t0 + C::T * q, xs = x < d.nx ? x : 0;
      gmask |= (geom[ca + xs] ? 1u : 0u) << q;
      gmask |= (geom[cb + xs] ? 1u : 0u) << (8 + q);
    }
  }
  __syncthreads();
  staged_copy<(C::H << L2NL), K::THREADS>(
      [&](int e) {
        const int kx = e >> L2NL, li = e & (NL - 1);
        const long long base = abase[li];
        return base >= 0 ? A[base + (long long)kx * d.ny] : make_float2(0.f, 0.f);
      },
      [&](int e, float2 v) {
        const int kx = e >> L2NL, li = e & (NL - 1);
        X[li * K::XS + kx] = v;
      });
  __syncthreads();
  const int g = threadIdx.x >> C::L2T, t = threadIdx.x & (C::T - 1);
  constexpr int HALF = C::L >> 1;
  const float2* xa_line = X + (2 * g) * K::XS;
  const float2* xb_line = xa_line + K::XS;
  float2 v[1][8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    float2 zv = make_float2(0.f, 0.f);
    if (q < C::E) {
      const int k = t + C::T * q;
      const bool upper = k > HALF;
      const int kk = upper ? C::L - k : k;
      float2 xa = xa_line[kk];
      float2 xb = xb_line[kk];
      if (kk == 0 || kk == HALF) { xa.y = 0.f; xb.y = 0.f; }
      zv = upper ? make_float2(xa.x + xb.y, xb.x - xa.y) : make_float2(xa.x - xb.y, xa.y + xb.x);
    }
    v[0][q] = zv;
  }
  __syncthreads();
  line_fft<L2, true, 1>(v, t, Z + g * K::LP, 0, tw);
  const int la = l0 + 2 * g, lb = la + 1;
  const int nyz = d.ny * d.nz;
  const int cella = (la % nyz) * d.nx, cellb = (lb % nyz) * d.nx;
#pragma unroll
  for (int q = 0; q < C::EH; ++q) {
    const int x = t + C::T * q;
    if (x >= d.nx) continue;
    if (la < d.nlines) h[(size_t)(la / nyz) * d.cs + cella + x] = (gmask >> q) & 1u ? v[0][q].x : 0.f;
    if (lb < d.nlines) h[(size_t)(lb / nyz) * d.cs + cellb + x] = (gmask >> (8 + q)) & 1u ? v[0][q].y : 0.f;
  }
}

template <int L2>
__global__ void __launch_bounds__(XinvCfg<L2>::THREADS)
    k_xinv(DD d, const float2* __restrict__ A, const uint8_t* __restrict__ geom, float* __restrict__ h,
           const float2* __restrict__ tw) {
  xinv_body<L2>(d, A, geom, h, tw, xcd_remap(blockIdx.x, gridDim.x));
}

// ------------------------------------------------ kernel spectrum (FP64)

__global__ void k_newell_octant(double* __restrict__ out, int nx, int ny, int nz, double cx, double cy, double cz,
                                int mgx, int mgy, int mgz, int px, int py, int pz) {
  const int n = nx * ny * nz;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const int x = idx % nx, y = (idx / nx) % ny, z = idx / (nx * ny);
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int i = -px; i <= px; ++i)
    for (int j = -py; j <= py; ++j)
      for (int k = -pz; k <= pz; ++k) {
        double v[6];
        newell6(x + i * mgx, y + j * mgy, z + k * mgz, cx, cy, cz, v);
        for (int c = 0; c < 6; ++c) acc[c] += v[c];
      }
  for (int c = 0; c < 6; ++c) out[(size_t)c * n + idx] = acc[c];
}

// out[(o*H + k)*inner + i] = sum_r in[(o*N + r)*inner + i] * w(r) * trig(2 pi k r / K)
// with cosine and weights 1,2,2,... for components even along the axis and
// sine with weight 2 for odd ones.  o runs over 6 components x `per` rows.
__global__ void k_trig_axis(const double* __restrict__ in, double* __restrict__ out, int per, int N, int inner,
                            int K, int H, int oddmask, const double* __restrict__ ct, const double* __restrict__ st) {
  const long long total = (long long)6 * per * H * inner;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int i = (int)(t % inner);
  const long long rest = t / inner;
  const int k = (int)(rest % H);
  const long long o = rest / H;
  const int comp = (int)(o / per);
  const bool odd = (oddmask >> comp) & 1;
  const double* src = in + (size_t)o * N * inner + i;
  double s = 0.0;
  if (odd) {
    for (int r = 1; r < N; ++r) s += 2.0 * src[(size_t)r * inner] * st[(int)(((long long)k * r) & (K - 1))];
  } else {
    s = src[0];
    for (int r = 1; r < N; ++r) s += 2.0 * src[(size_t)r * inner] * ct[(int)(((long long)k * r) & (K - 1))];
  }
  out[t] = s;
}

__global__ void k_finalize(const double* __restrict__ in, float* __restrict__ T, int Hx, int Hy, int Hz, double pref) {
  const long long total = (long long)6 * Hx * Hy * Hz;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  // output index ((kx*Hy + ky)*Hz + kz)*6 + c: the six components of an
  // octant entry are adjacent
  const int c = (int)(t % 6);
  const long long e = t / 6;
  const int kz = (int)(e % Hz);
  const int ky = (int)((e / Hz) % Hy);
  const int kx = (int)(e / ((long long)Hz * Hy));
  const double sign = (c >= 3) ? -1.0 : 1.0;  // (-i)^2 of the two odd axes
  const double v = in[(((size_t)c * Hz + kz) * Hy + ky) * Hx + kx];
  T[t] = (float)(sign * pref * v);
}

__global__ void k_newell_grid(float* __restrict__ out, int Kx, int Ky, int Kz, int ox, int oy, int oz, double cx,
                              double cy, double cz, int mgx, int mgy, int mgz, int px, int py, int pz) {
  const int n = Kx * Ky * Kz;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const int x = ox + idx % Kx, y = oy + (idx / Kx) % Ky, z = oz + idx / (Kx * Ky);
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int i = -px; i <= px; ++i)
    for (int j = -py; j <= py; ++j)
      for (int k = -pz; k <= pz; ++k) {
        double v[6];
        newell6(x + i * mgx, y + j * mgy, z + k * mgz, cx, cy, cz, v);
        for (int c = 0; c < 6; ++c) acc[c] += v[c];
      }
  for (int c = 0; c < 6; ++c) out[(size_t)c * n + idx] = (float)acc[c];
}

unsigned grid_for(long long n, int block = BLOCK) { return (unsigned)((n + block - 1) / block); }

void make_twiddles(DeviceArray<float2>& tw, int l2) {
  const std::vector<float2> h = stage_twiddles(l2);
  tw.allocate(h.size());
  HIP_CHECK(hipMemcpy(tw.get(), h.data(), h.size() * sizeof(float2), hipMemcpyHostToDevice));
}

void make_trig(DeviceArray<double>& ct, DeviceArray<double>& st, int K) {
  std::vector<double> c(K), s(K);
  for (int q = 0; q < K; ++q) {
    const double a = 2.0 * M_PI * (double)q / (double)K;
    c[q] = std::cos(a);
    s[q] = std::sin(a);
  }
  ct.allocate(K);
  st.allocate(K);
  HIP_CHECK(hipMemcpy(ct.get(), c.data(), K * sizeof(double), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(st.get(), s.data(), K * sizeof(double), hipMemcpyHostToDevice));
}

// ------------------------------------------------ launchers per line length

template <int L2>
struct RunXfwd {
  static void run(const DD& d, const float* m, PS msat, const float* m2, PS msat2, const uint8_t* geom, float2* A,
                  const float2* tw, hipStream_t s) {
    using K = XfwdCfg<L2>;
    const int npairs = (d.nlines + 1) / 2;
    const dim3 grid = grid_for(npairs, K::PB);
    if (!geom && !msat.p && !m2) k_xfwd<L2, 1><<<grid, K::THREADS, K::SHM, s>>>(d, m, msat, m2, msat2, geom, A, tw);
    else if (!geom && !msat.p && !msat2.p)
      k_xfwd<L2, 2><<<grid, K::THREADS, K::SHM, s>>>(d, m, msat, m2, msat2, geom, A, tw);
    else k_xfwd<L2, 0><<<grid, K::THREADS, K::SHM, s>>>(d, m, msat, m2, msat2, geom, A, tw);
    checkLaunch("k_xfwd");
  }
};

template <int L2>
struct RunXfwdCombine {
  static void run(const DD& d, const float* x0, const Combo& co, float* xout, PS msat, const uint8_t* geom,
                  float2* A, const float2* tw, hipStream_t s) {
    using K = XfwdCfg<L2>;
    const int npairs = (d.ny * d.nz + 1) / 2;
    k_xfwd_combine<L2><<<grid_for(npairs, K::PB), K::THREADS, K::SHM, s>>>(d, x0, co, xout, msat, geom, A, tw);
    checkLaunch("k_xfwd_combine");
  }
};

template <int L2>
struct RunYfwd {
  static void run(const DD& d, const float2* A, float2* B, const float2* tw, hipStream_t s) {
    using K = YCfg<L2>;
    k_yfwd<L2><<<grid_for(3LL * d.Hx * d.nz, K::LB), K::THREADS, K::SHM, s>>>(d, A, B, tw);
    checkLaunch("k_yfwd");
  }
};

template <int L2>
struct RunYinv {
  static void run(const DD& d, const float2* B, float2* A, const float2* tw, hipStream_t s) {
    using K = YinvCfg<L2>;
    k_yinv<L2><<<grid_for(3LL * d.Hx * d.nz, K::LB), K::THREADS, K::SHM, s>>>(d, B, A, tw);
    checkLaunch("k_yinv");
  }
};

template <int L2>
struct RunZpass {
  static void run(const DD& d, float2* B, const float* Tk, const float2* tw, hipStream_t s) {
    using K = PointCfg<L2>;
    k_zpass<L2><<<grid_for((long long)d.Hx * d.Ky, K::PT), K::THREADS, K::SHM, s>>>(d, B, Tk, tw);
    checkLaunch("k_zpass");
  }
};

template <int L2>
struct RunYpass2d {
  static void run(const DD& d, float2* A, const float* Tk, const float2* tw, hipStream_t s) {
    using K = PointCfg<L2>;
    k_ypass2d<L2><<<grid_for(d.Hx, K::PT), K::THREADS, K::SHM, s>>>(d, A, Tk, tw);
    checkLaunch("k_ypass2d");
  }
};

template <int L2>
struct RunXinv {
  static void run(const DD& d, const float2* A, const uint8_t* geom, float* h, const float2* tw, hipStream_t s) {
    using K = XinvCfg<L2>;
    const int npairs = (d.nlines + 1) / 2;
    k_xinv<L2><<<grid_for(npairs, K::PB), K::THREADS, K::SHM, s>>>(d, A, geom, h, tw);
    checkLaunch("k_xinv");
  }
};

// Batched passes: one launch of 256-thread blocks per pass over magnets of
// every size; a block finds its magnet through b2s and runs that magnet's
// line length.
// Block b of batched pass PASS (0 xfwd, 1 yfwd, 2 zpass, 3 yinv, 4 ypass2d,
// 5 xinv): the member b2s[b] at its own line length.  X: packed state (pass
// 0), H: packed field (pass 5).
// Lines of up to 2^PERSIST_MAXL2 points in the persistent step kernel (its
// register budget is that of its largest body).
constexpr int PERSIST_MAXL2 = 8;

template <int PASS, int LMAX = MAXL2>
__device__ __forceinline__ void batch_pass_block(const BSeg* __restrict__ segs, const int* __restrict__ b2s, int b,
                                                 const float* __restrict__ X, float* __restrict__ H) {
  const BSeg s = segs[b2s[b]];
  const int lb = b - s.first;
#define MPA_L2_CASES(l2, CALL)                       \
  switch (l2) {                                       \
    case 0: CALL(0); break;                           \
    case 3: CALL(3); break;                           \
    case 4: CALL(4); break;                           \
    case 5: CALL(5); break;                           \
    case 6: CALL(6); break;                           \
    case 7: CALL(7); break;                           \
    case 8: CALL(8); break;                           \
    case 9: if constexpr (LMAX >= 9) CALL(9); break;  \
    case 10: if constexpr (LMAX >= 10) CALL(10); break; \
    case 11: if constexpr (LMAX >= 11) CALL(11); break; \
    default: break;                                   \
  }
  if constexpr (PASS == 0) {
#define MPA_CALL(L) xfwd_body<L, XfwdCfgB<L>>(s.d, X + s.off, s.msat, nullptr, PS{nullptr, 0.f}, s.geom, s.A, s.tw[0], lb)
    MPA_L2_CASES(s.d.l2x, MPA_CALL)
#undef MPA_CALL
  } else if constexpr (PASS == 1) {
#define MPA_CALL(L) yfwd_body<L, YCfgB<L>>(s.d, s.A, s.B, s.tw[1], lb)
    MPA_L2_CASES(s.d.l2y, MPA_CALL)
#undef MPA_CALL
  } else if constexpr (PASS == 2) {
#define MPA_CALL(L) zpass_body<L, PointCfgB<L>>(s.d, s.B, s.T, s.tw[2], lb)
    MPA_L2_CASES(s.d.l2z, MPA_CALL)
#undef MPA_CALL
  } else if constexpr (PASS == 3) {
#define MPA_CALL(L) yinv_body<L, YCfgB<L>>(s.d, s.B, s.A, s.tw[1], lb)
    MPA_L2_CASES(s.d.l2y, MPA_CALL)
#undef MPA_CALL
  } else if constexpr (PASS == 4) {
#define MPA_CALL(L) ypass2d_body<L, PointCfgB<L>>(s.d, s.A, s.T, s.tw[1], lb)
    MPA_L2_CASES(s.d.l2y, MPA_CALL)
#undef MPA_CALL
  } else {
#define MPA_CALL(L) xinv_body<L, XinvCfgB<L>>(s.d, s.A, s.geom, H + s.off, s.tw[0], lb)
    MPA_L2_CASES(s.d.l2x, MPA_CALL)
#undef MPA_CALL
  }
}
#undef MPA_L2_CASES

__global__ void __launch_bounds__(256) k_xfwd_b(const BSeg* __restrict__ segs, const int* __restrict__ b2s,
                                                const float* __restrict__ X) {
  batch_pass_block<0>(segs, b2s, xcd_remap(blockIdx.x, gridDim.x), X, nullptr);
}

__global__ void __launch_bounds__(256) k_yfwd_b(const BSeg* __restrict__ segs, const int* __restrict__ b2s) {
  batch_pass_block<1>(segs, b2s, xcd_remap(blockIdx.x, gridDim.x), nullptr, nullptr);
}

__global__ void __launch_bounds__(256) k_zpass_b(const BSeg* __restrict__ segs, const int* __restrict__ b2s) {
  batch_pass_block<2>(segs, b2s, xcd_remap(blockIdx.x, gridDim.x), nullptr, nullptr);
}

__global__ void __launch_bounds__(256) k_yinv_b(const BSeg* __restrict__ segs, const int* __restrict__ b2s) {
  batch_pass_block<3>(segs, b2s, xcd_remap(blockIdx.x, gridDim.x), nullptr, nullptr);
}

__global__ void __launch_bounds__(256) k_ypass2d_b(const BSeg* __restrict__ segs, const int* __restrict__ b2s) {
  batch_pass_block<4>(segs, b2s, xcd_remap(blockIdx.x, gridDim.x), nullptr, nullptr);
}

__global__ void __launch_bounds__(256) k_xinv_b(const BSeg* __restrict__ segs, const int* __restrict__ b2s,
                                                float* __restrict__ H) {
  batch_pass_block<5>(segs, b2s, xcd_remap(blockIdx.x, gridDim.x), nullptr, H);
}

// ---------------------------------------------------------------------------
// Persistent batch step: every launch of the per-pass step (stage
// combination, demag passes, field, error norms, controller) is a phase of
// one kernel; the phases' blocks are shared out over a grid that is resident
// as a whole (a cooperative launch), and a grid barrier separates phases.

__device__ __forceinline__ unsigned ld_agent(const unsigned* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int ld_agent(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Grid barrier number `gen` (generations increase by one per barrier and
// never repeat between launches).  After the block's own barrier (its
// stores have then reached the L2), one thread per block publishes them
// (agent-scope release: the L2 written back) and stores gen into the
// block's arrival word; block 0 waits for every arrival word and stores gen
// into the release word, which the other blocks wait for; one thread per
// block then acquires (agent scope: the L1 and L2 invalidated) before the
// block's barrier.  The agent-scope fences act on the compute unit's and
// the XCD's caches, so one thread per block suffices, and is several times
// cheaper at a full grid than a fence in every thread.  Block 0 is the only
// one that decides: when an arrival does not come within a.spin polls it
// writes the error word instead of releasing, so either every block passes
// barrier gen or none does.  A block that waits far longer than that for
// the release (64 a.spin polls) gives up too.  Returns false when the
// kernel must end.
__device__ bool grid_barrier(const bk::PersistArgs& a, unsigned gen, int ordinal, int& fail) {
  if (threadIdx.x == 0) fail = 0;
  __syncthreads();
  if (blockIdx.x == 0) {
    bool ok = a.spin != 0;
    for (int i = 1 + (int)threadIdx.x; ok && i < (int)gridDim.x; i += (int)blockDim.x) {
      unsigned n = 0;
      while (ld_agent(a.flags + i) != gen) {
        if (++n > a.spin || ld_agent(a.err) != 0) {
          ok = false;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    if (!ok) fail = 1;
    __syncthreads();
    if (threadIdx.x == 0) {
      if (fail) {
        int expect = 0;
        __hip_atomic_compare_exchange_strong(a.err, &expect, 1 + ordinal, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
      } else {
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
        __hip_atomic_store(a.go, gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  } else if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_store(a.flags + blockIdx.x, gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long limit = 64ull * (a.spin ? a.spin : 1u);
    unsigned long long n = 0;
    while (ld_agent(a.go) != gen) {
      if (ld_agent(a.err) != 0) {
        fail = 1;
        break;
      }
      if (++n > limit) {
        int expect = 0;
        __hip_atomic_compare_exchange_strong(a.err, &expect, -1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
        fail = 1;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
  return fail == 0;
}

// Arguments of a called function arrive in vector registers; these put the
// block-uniform ones back into scalar registers, so the body's member table
// loads and address arithmetic stay scalar as in the per-pass kernels.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
template <class T>
__device__ __forceinline__ T* uni(T* p) {
  const unsigned long long u = reinterpret_cast<unsigned long long>(p);
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
  return reinterpret_cast<T*>(((unsigned long long)hi << 32) | lo);
}

template <int PASS>
__device__ __noinline__ void persist_pass(const BSeg* segs, const int* b2s, int b, const float* X, float* H) {
  batch_pass_block<PASS, PERSIST_MAXL2>(uni(segs), uni(b2s), uni(b), uni(X), uni(H));
}

__device__ __forceinline__ void persist_demag(const bk::PhasePass& P, int b, const float* X, float* H) {
  const BSeg* segs = static_cast<const BSeg*>(P.segs);
  switch (P.pass) {
    case 0: persist_pass<0>(segs, P.b2s, b, X, H); break;
    case 1: persist_pass<1>(segs, P.b2s, b, X, H); break;
    case 2: persist_pass<2>(segs, P.b2s, b, X, H); break;
    case 3: persist_pass<3>(segs, P.b2s, b, X, H); break;
    case 4: persist_pass<4>(segs, P.b2s, b, X, H); break;
    default: persist_pass<5>(segs, P.b2s, b, X, H); break;
  }
}

__device__ __noinline__ void persist_stage(const bk::PersistArgs* a, const bk::Phase* P, int v) {
  a = uni(a);
  P = uni(P);
  bk::bstage_block(P->sa, a->X, P->out, a->X0, a->K0, a->KL, a->G, a->c2m, a->st, a->cs, a->nch, uni(v));
}

__device__ __noinline__ void persist_field(const bk::PersistArgs* a, const bk::Phase* P, int v) {
  a = uni(a);
  P = uni(P);
  const FBInst* tab = static_cast<const FBInst*>(a->ftab);
  if (!P->has_next) {
    field_uniform_b_block<0>(tab, a->fb2i, P->x, a->HD, P->out, a->cs, uni(v), P->next);
    return;
  }
  switch (P->next.nk) {
    case 1: field_uniform_b_block<1>(tab, a->fb2i, P->x, a->HD, P->out, a->cs, uni(v), P->next); break;
    case 2: field_uniform_b_block<2>(tab, a->fb2i, P->x, a->HD, P->out, a->cs, uni(v), P->next); break;
    case 3: field_uniform_b_block<3>(tab, a->fb2i, P->x, a->HD, P->out, a->cs, uni(v), P->next); break;
    case 4: field_uniform_b_block<4>(tab, a->fb2i, P->x, a->HD, P->out, a->cs, uni(v), P->next); break;
    case 5: field_uniform_b_block<5>(tab, a->fb2i, P->x, a->HD, P->out, a->cs, uni(v), P->next); break;
    default: field_uniform_b_block<-1>(tab, a->fb2i, P->x, a->HD, P->out, a->cs, uni(v), P->next); break;
  }
}

__device__ __noinline__ void persist_err(const bk::PersistArgs* a, const bk::ErrArgs* ea, int v,
                                         float (*red)[bk::CH / 64]) {
  a = uni(a);
  bk::berr_block(*uni(ea), a->G, a->c2m, a->st, a->errbits, a->tqbits, a->cs, uni(v), red);
}

__device__ __noinline__ void persist_control(const bk::PersistArgs* a, const bk::CtlArgs* ca, int v) {
  a = uni(a);
  bk::bcontrol_member(a->st, a->errbits, a->tqbits, a->nslots, *uni(ca), uni(v) * 256 + (int)threadIdx.x);
}

// The launch's arguments are in device memory (pa), so the called phase
// bodies read them through a pointer instead of a private copy.
__global__ void __launch_bounds__(256, 3) k_bstep_persist(const bk::PersistArgs* __restrict__ pa, int nsteps,
                                                          unsigned gen0) {
  __shared__ float red[2][bk::CH / 64];
  __shared__ int fail;
  const bk::PersistArgs& a = *pa;
  int ordinal = 0;
  for (int step = 0; step < nsteps; ++step) {
    for (int p = 0; p < a.nph; ++p) {
      const bk::Phase& P = a.ph[p];
      const int kind = P.kind, nvb = P.nvb;
      for (int v = blockIdx.x; v < nvb; v += gridDim.x) {
        switch (kind) {
          case bk::PH_STAGE: persist_stage(pa, &P, v); break;
          case bk::PH_DEMAG:
            if (v < P.pass[0].nblocks) persist_demag(P.pass[0], v, P.x, a.HD);
            else persist_demag(P.pass[1], v - P.pass[0].nblocks, P.x, a.HD);
            break;
          case bk::PH_FIELD: persist_field(pa, &P, v); break;
          case bk::PH_ERR: persist_err(pa, &P.ea, v, red); break;
          default: persist_control(pa, &P.ca, v); break;
        }
        __syncthreads();  // the next virtual block reuses the LDS
      }
      if (!grid_barrier(a, gen0 + (unsigned)ordinal, ordinal, fail)) return;
      ++ordinal;
    }
  }
}

// Blocks a magnet needs in a batched pass, and the pass's LDS bytes at its
// line length.
template <int L2>
struct BatchShape {
  static void run(int pass, const DD& d, unsigned* blocks, size_t* shm) {
    switch (pass) {
      case 0:
        *blocks = grid_for((d.nlines + 1) / 2, XfwdCfgB<L2>::PB);
        *shm = XfwdCfgB<L2>::SHM;
        break;
      case 1:
      case 3:
        *blocks = grid_for(3LL * d.Hx * d.nz, YCfgB<L2>::LB);
        *shm = YCfgB<L2>::SHM;
        break;
      case 2:
        *blocks = grid_for((long long)d.Hx * d.Ky, PointCfgB<L2>::PT);
        *shm = PointCfgB<L2>::SHM;
        break;
      case 4:
        *blocks = grid_for(d.Hx, PointCfgB<L2>::PT);
        *shm = PointCfgB<L2>::SHM;
        break;
      default:
        *blocks = grid_for((d.nlines + 1) / 2, XinvCfgB<L2>::PB);
        *shm = XinvCfgB<L2>::SHM;
        break;
    }
  }
};

void run_batch_pass(int pass, unsigned nb, size_t shm, const BSeg* segs, const int* b2s, const float* X, float* H,
                    hipStream_t s) {
  switch (pass) {
    case 0: k_xfwd_b<<<nb, 256, shm, s>>>(segs, b2s, X); break;
    case 1: k_yfwd_b<<<nb, 256, shm, s>>>(segs, b2s); break;
    case 2: k_zpass_b<<<nb, 256, shm, s>>>(segs, b2s); break;
    case 3: k_yinv_b<<<nb, 256, shm, s>>>(segs, b2s); break;
    case 4: k_ypass2d_b<<<nb, 256, shm, s>>>(segs, b2s); break;
    default: k_xinv_b<<<nb, 256, shm, s>>>(segs, b2s, H); break;
  }
  checkLaunch("batched demag pass");
}

// Calls R<l2>::run(args...) for a supported line length 2^l2.
template <template <int> class R, typename... Args>
void with_line(int l2, Args&&... args) {
  switch (l2) {
    case 0: R<0>::run(std::forward<Args>(args)...); return;
    case 3: R<3>::run(std::forward<Args>(args)...); return;
    case 4: R<4>::run(std::forward<Args>(args)...); return;
    case 5: R<5>::run(std::forward<Args>(args)...); return;
    case 6: R<6>::run(std::forward<Args>(args)...); return;
    case 7: R<7>::run(std::forward<Args>(args)...); return;
    case 8: R<8>::run(std::forward<Args>(args)...); return;
    case 9: R<9>::run(std::forward<Args>(args)...); return;
    case 10: R<10>::run(std::forward<Args>(args)...); return;
    case 11: R<11>::run(std::forward<Args>(args)...); return;
    default: break;
  }
  throw Error(MPA_ERR_RUNTIME, "demag: unsupported FFT line length 2^" + std::to_string(l2));
}

}  // namespace

int demag_fft_length(int n) {
  if (n <= 1) return 1;
  int k = 8;  // the line transforms need 1 or >= 8 points
  while (k < 2 * n - 1) k <<= 1;
  return k;
}

DemagPlan::DemagPlan(const DemagGeometry& g) : g_(g) {
  for (int a = 0; a < 3; ++a) {
    if (g.n[a] < 1) throw Error(MPA_ERR_INVALID, "demag: grid size must be positive");
    K_[a] = demag_fft_length(g.n[a]);
    log2K_[a] = ilog2(K_[a]);
    if (log2K_[a] > MAXL2) throw Error(MPA_ERR_INVALID, "demag: grid axis too large for the FFT passes (max 1024 cells)");
  }
  N_ = g.n[0] * g.n[1] * g.n[2];
  Hx_ = K_[0] == 1 ? 1 : K_[0] / 2 + 1;
  Hy_ = K_[1] / 2 + 1;
  Hz_ = K_[2] / 2 + 1;
  for (int a = 0; a < 3; ++a) make_twiddles(tw_[a], log2K_[a]);
  build_kernel(nullptr);
}

void DemagPlan::build_kernel(hipStream_t s) {
  const int nx = g_.n[0], ny = g_.n[1], nz = g_.n[2];
  DeviceArray<double> oct((size_t)6 * N_);
  k_newell_octant<<<grid_for(N_), BLOCK, 0, s>>>(oct.get(), nx, ny, nz, g_.cs[0], g_.cs[1], g_.cs[2],
                                                 g_.mastergrid[0], g_.mastergrid[1], g_.mastergrid[2], g_.pbc[0],
                                                 g_.pbc[1], g_.pbc[2]);
  checkLaunch("k_newell_octant");
  DeviceArray<double> ct[3], st[3];
  for (int a = 0; a < 3; ++a) make_trig(ct[a], st[a], K_[a]);
  // odd components per axis: x: xy(3), xz(4); y: xy(3), yz(5); z: xz(4), yz(5)
  const int oddx = (1 << 3) | (1 << 4), oddy = (1 << 3) | (1 << 5), oddz = (1 << 4) | (1 << 5);
  DeviceArray<double> t1((size_t)6 * nz * ny * Hx_);
  k_trig_axis<<<grid_for((long long)6 * nz * ny * Hx_), BLOCK, 0, s>>>(oct.get(), t1.get(), nz * ny, nx, 1, K_[0],
                                                                       Hx_, oddx, ct[0].get(), st[0].get());
  checkLaunch("k_trig_axis x");
  oct.release();
  DeviceArray<double> t2((size_t)6 * nz * Hy_ * Hx_);
  k_trig_axis<<<grid_for((long long)6 * nz * Hy_ * Hx_), BLOCK, 0, s>>>(t1.get(), t2.get(), nz, ny, Hx_, K_[1], Hy_,
                                                                        oddy, ct[1].get(), st[1].get());
  checkLaunch("k_trig_axis y");
  t1.release();
  DeviceArray<double> t3((size_t)6 * Hz_ * Hy_ * Hx_);
  k_trig_axis<<<grid_for((long long)6 * Hz_ * Hy_ * Hx_), BLOCK, 0, s>>>(t2.get(), t3.get(), 1, nz, Hy_ * Hx_, K_[2],
                                                                        Hz_, oddz, ct[2].get(), st[2].get());
  checkLaunch("k_trig_axis z");
  t2.release();
  T_.allocate((size_t)6 * Hx_ * Hy_ * Hz_);
  const double pref = -MU0 / ((double)K_[0] * K_[1] * K_[2]);
  k_finalize<<<grid_for((long long)6 * Hx_ * Hy_ * Hz_), BLOCK, 0, s>>>(t3.get(), T_.get(), Hx_, Hy_, Hz_, pref);
  checkLaunch("k_finalize");
  HIP_CHECK(hipStreamSynchronize(s));
}

size_t DemagPlan::a_count() const { return (size_t)3 * g_.n[2] * Hx_ * g_.n[1]; }
size_t DemagPlan::b_count() const { return is2d() ? 0 : (size_t)3 * Hx_ * K_[1] * g_.n[2]; }

size_t DemagPlan::workspace_bytes() const {
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  return up(a_count() * sizeof(float2)) + up(b_count() * sizeof(float2));
}

void DemagPlan::ensure_workspace() {
  if (A_.n == 0) A_.allocate(a_count());
  if (!is2d() && B_.n == 0) B_.allocate(b_count());
}

void DemagPlan::exec(const float* m, PS msat, const uint8_t* geom, float* h, hipStream_t s, Profiler* prof,
                     const float* m2, PS msat2) {
  ensure_workspace();
  run(A_.get(), B_.get(), m, msat, geom, h, s, prof, m2, msat2);
}

// Scratch of the allocation-free exec: a file-scope device array, one copy per
// GPU, held until the process ends.  Host code never passes `scratch` itself
// (on the host it names a stand-in, not device memory): the device address
// comes from hipGetSymbolAddress, looked up once per device on its first use
// with that device current.
namespace {
constexpr size_t SCRATCH_BYTES = (size_t)64 << 20;
__device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];
constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
void* scratch_ptr[MAX_DEVICES] = {};

void* scratch_address() {
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) throw Error(MPA_ERR_RUNTIME, "demag scratch: device index out of range");
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_ptr[dev]) {
    void* p = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Error(MPA_ERR_RUNTIME, "demag scratch: no device address");
    scratch_ptr[dev] = p;
  }
  return scratch_ptr[dev];
}
}  // namespace

size_t DemagPlan::scratch_capacity() { return SCRATCH_BYTES; }

void DemagPlan::exec_noalloc(const float* m, PS msat, const uint8_t* geom, float* h, hipStream_t s) {
  const size_t need = workspace_bytes();
  if (need > SCRATCH_BYTES)
    throw Error(MPA_ERR_INVALID, "demag: this grid needs " + std::to_string(need >> 20) +
                                     " MiB of FFT scratch; the allocation-free exec holds " +
                                     std::to_string(SCRATCH_BYTES >> 20) + " MiB");
  unsigned char* base = static_cast<unsigned char*>(scratch_address());
  float2* A = reinterpret_cast<float2*>(base);
  float2* B = is2d() ? nullptr : reinterpret_cast<float2*>(base + (a_count() * sizeof(float2) + 255) / 256 * 256);
  run(A, B, m, msat, geom, h, s, nullptr, nullptr, PS{nullptr, 0.f});
}

void DemagPlan::exec_combine(const float* x0, const Combo& co, float* x, PS msat, const uint8_t* geom, float* h,
                             hipStream_t s, Profiler* prof) {
  if (co.nk < 0 || co.nk > MAX_STAGES || !x0 || !x || x == x0)
    throw Error(MPA_ERR_RUNTIME, "demag: bad stage combination");
  for (int j = 0; j < co.nk; ++j)
    if (co.k[j] == x) throw Error(MPA_ERR_RUNTIME, "demag: stage combination writes a derivative it reads");
  ensure_workspace();
  const StageIn in{x0, &co, x};
  run(A_.get(), B_.get(), x, msat, geom, h, s, prof, nullptr, PS{nullptr, 0.f}, &in);
}

void DemagPlan::run(float2* A, float2* B, const float* m, PS msat, const uint8_t* geom, float* h, hipStream_t s,
                    Profiler* prof, const float* m2, PS msat2, const StageIn* in) {
  DD d;
  d.nx = g_.n[0]; d.ny = g_.n[1]; d.nz = g_.n[2];
  d.Kx = K_[0]; d.Ky = K_[1]; d.Kz = K_[2];
  d.Hx = Hx_; d.Hy = Hy_; d.Hz = Hz_;
  d.l2x = log2K_[0]; d.l2y = log2K_[1]; d.l2z = log2K_[2];
  d.nlines = 3 * d.ny * d.nz;
  d.cs = (long long)d.nx * d.ny * d.nz;
  if (in) {
    // read x0 and nk derivatives, write the state (12 B per cell each) instead
    // of reading it
    const double n = (double)d.nx * d.ny * d.nz;
    Timed t(prof, "k_xfwd_combine", pass_bytes(0) + 12.0 * n * (in->co->nk + 1), s);
    with_line<RunXfwdCombine>(d.l2x, d, in->x0, *in->co, in->x, msat, geom, A, tw_[0].get(), s);
  } else {
    // the second sublattice's magnetisation is one more read of 12 B per cell
    Timed t(prof, "k_xfwd", pass_bytes(0) + (m2 ? 12.0 * d.nx * d.ny * d.nz : 0.0), s);
    with_line<RunXfwd>(d.l2x, d, m, msat, m2, msat2, geom, A, tw_[0].get(), s);
  }
  if (is2d()) {
    Timed t(prof, "k_ypass2d", pass_bytes(1), s);
    with_line<RunYpass2d>(d.l2y, d, A, T_.get(), tw_[1].get(), s);
  } else {
    {
      Timed t(prof, "k_yfwd", pass_bytes(1), s);
      with_line<RunYfwd>(d.l2y, d, A, B, tw_[1].get(), s);
    }
    {
      Timed t(prof, "k_zpass", pass_bytes(2), s);
      with_line<RunZpass>(d.l2z, d, B, T_.get(), tw_[2].get(), s);
    }
    {
      Timed t(prof, "k_yinv", pass_bytes(3), s);
      with_line<RunYinv>(d.l2y, d, B, A, tw_[1].get(), s);
    }
  }
  {
    Timed t(prof, "k_xinv", pass_bytes(npasses() - 1), s);
    with_line<RunXinv>(d.l2x, d, A, geom, h, tw_[0].get(), s);
  }
}

DemagBatch::DemagBatch(const std::vector<Member>& members, long long cs) {
  const int M = (int)members.size();
  std::vector<DD> dd(M);
  for (int j = 0; j < M; ++j) {
    const DemagPlan& p = *members[j].plan;
    DD& d = dd[j];
    d.nx = p.g_.n[0]; d.ny = p.g_.n[1]; d.nz = p.g_.n[2];
    d.Kx = p.K_[0]; d.Ky = p.K_[1]; d.Kz = p.K_[2];
    d.Hx = p.Hx_; d.Hy = p.Hy_; d.Hz = p.Hz_;
    d.l2x = p.log2K_[0]; d.l2y = p.log2K_[1]; d.l2z = p.log2K_[2];
    d.nlines = 3 * d.ny * d.nz;
    d.cs = cs;
  }
  // pass -> (applies to a member, line length of its axis, plan pass index)
  auto applies = [&](int pass, int j) {
    const bool flat = members[j].plan->is2d();
    return pass == 0 || pass == 5 || (pass == 4 ? flat : !flat);
  };
  auto l2_of = [&](int pass, int j) {
    const DD& d = dd[j];
    return (pass == 0 || pass == 5) ? d.l2x : (pass == 2 ? d.l2z : d.l2y);
  };
  auto plan_pass = [&](int pass, int j) {
    if (members[j].plan->is2d()) return pass == 0 ? 0 : (pass == 4 ? 1 : 2);
    return pass == 5 ? 4 : pass;
  };
  // two independent chains: the members with a z transform (group 0) and the
  // single layers (group 1)
  for (int group = 0; group < 2; ++group)
  for (int pass = 0; pass < 6; ++pass) {
    {
      std::vector<BSeg> segs;
      std::vector<int> b2s;
      double bytes = 0.0;
      size_t shm_max = 0;
      for (int j = 0; j < M; ++j) {
        if ((members[j].plan->is2d() ? 1 : 0) != group || !applies(pass, j)) continue;
        unsigned nb = 0;
        size_t shm = 0;
        with_line<BatchShape>(l2_of(pass, j), pass, dd[j], &nb, &shm);
        shm_max = shm > shm_max ? shm : shm_max;
        members[j].plan->ensure_workspace();
        const DemagPlan& p = *members[j].plan;
        BSeg sg{};
        sg.d = dd[j];
        sg.off = members[j].off;
        sg.msat = members[j].msat;
        sg.geom = members[j].geom;
        sg.A = p.A_.get();
        sg.B = p.B_.get();
        sg.T = p.T_.get();
        for (int a = 0; a < 3; ++a) sg.tw[a] = p.tw_[a].get();
        sg.first = (int)b2s.size();
        b2s.insert(b2s.end(), nb, (int)segs.size());
        segs.push_back(sg);
        bytes += p.pass_bytes(plan_pass(pass, j));
      }
      if (segs.empty()) continue;
      Launch L;
      L.pass = pass;
      L.group = group;
      L.shm = shm_max;
      L.nblocks = (unsigned)b2s.size();
      L.bytes = bytes;
      L.segs.allocate(segs.size() * sizeof(BSeg));
      HIP_CHECK(hipMemcpy(L.segs.get(), segs.data(), segs.size() * sizeof(BSeg), hipMemcpyHostToDevice));
      L.b2s.allocate(b2s.size());
      HIP_CHECK(hipMemcpy(L.b2s.get(), b2s.data(), b2s.size() * sizeof(int), hipMemcpyHostToDevice));
      launches_.push_back(std::move(L));
    }
  }
}

void DemagBatch::exec(const float* X, float* H, hipStream_t s, Profiler* prof, hipStream_t s2, hipEvent_t fork,
                      hipEvent_t join) const {
  static const char* names[] = {"k_xfwd_b", "k_yfwd_b", "k_zpass_b", "k_yinv_b", "k_ypass2d_b", "k_xinv_b"};
  bool has[2] = {false, false};
  for (const Launch& L : launches_) has[L.group] = true;
  const bool split = s2 && fork && join && has[0] && has[1];
  if (split) {
    // the single layers' chain on the second stream, concurrently
    HIP_CHECK(hipEventRecord(fork, s));
    HIP_CHECK(hipStreamWaitEvent(s2, fork, 0));
  }
  for (const Launch& L : launches_) {
    hipStream_t st = split && L.group == 1 ? s2 : s;
    Timed t(prof, names[L.pass], L.bytes, st);
    run_batch_pass(L.pass, L.nblocks, L.shm, reinterpret_cast<const BSeg*>(L.segs.get()), L.b2s.get(), X, H, st);
  }
  if (split) {
    HIP_CHECK(hipEventRecord(join, s2));
    HIP_CHECK(hipStreamWaitEvent(s, join, 0));
  }
}

std::vector<DemagBatch::PassTable> DemagBatch::pass_tables() const {
  std::vector<PassTable> out;
  for (const Launch& L : launches_) out.push_back({L.pass, L.group, L.nblocks, L.shm, L.segs.get(), L.b2s.get()});
  return out;
}

int batch_persist_blocks_per_cu(size_t shm) {
  int nb = 0;
  HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(k_bstep_persist), 256,
                                                         shm));
  return nb;
}

void batch_persist_launch(const bk::PersistArgs* a, int nsteps, unsigned gen0, int grid, size_t shm, hipStream_t s) {
  const bk::PersistArgs* args = a;
  void* params[] = {&args, &nsteps, &gen0};
  HIP_CHECK(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_bstep_persist), dim3((unsigned)grid),
                                       dim3(256), params, (unsigned)shm, s));
}

void DemagBatch::exec_group(const float* X, float* H, int group, hipStream_t s, Profiler* prof) const {
  static const char* names[] = {"k_xfwd_b", "k_yfwd_b", "k_zpass_b", "k_yinv_b", "k_ypass2d_b", "k_xinv_b"};
  for (const Launch& L : launches_) {
    if (L.group != group) continue;
    Timed t(prof, names[L.pass], L.bytes, s);
    run_batch_pass(L.pass, L.nblocks, L.shm, reinterpret_cast<const BSeg*>(L.segs.get()), L.b2s.get(), X, H, s);
  }
}

bool DemagBatch::has_group(int group) const {
  for (const Launch& L : launches_)
    if (L.group == group) return true;
  return false;
}

double DemagBatch::algorithmic_bytes() const {
  double b = 0.0;
  for (const Launch& L : launches_) b += L.bytes;
  return b;
}

const char* DemagPlan::pass_name(int i) const {
  static const char* n3[] = {"k_xfwd", "k_yfwd", "k_zpass", "k_yinv", "k_xinv"};
  static const char* n2[] = {"k_xfwd", "k_ypass2d", "k_xinv"};
  return is2d() ? n2[i] : n3[i];
}

double DemagPlan::pass_bytes(int i) const {
  const double N = (double)g_.n[0] * g_.n[1] * g_.n[2];
  const double ny = g_.n[1], nz = g_.n[2];
  const double xa = 24.0 * Hx_ * ny * nz;    // one pass over A (complex 3-vectors)
  const double yb = 24.0 * Hx_ * K_[1] * nz;  // one pass over B
  if (is2d()) {
    const double t = 16.0 * Hx_ * Hy_;  // xx, yy, zz, xy (xz = yz = 0 in a single layer)
    const double v[3] = {12 * N + xa, 2 * xa + t, xa + 12 * N};
    return v[i];
  }
  const double t = 24.0 * Hx_ * Hy_ * Hz_;
  const double v[5] = {12 * N + xa, xa + yb, 2 * yb + t, yb + xa, xa + 12 * N};
  return v[i];
}

double DemagPlan::algorithmic_bytes() const {
  double b = 0.0;
  for (int i = 0; i < npasses(); ++i) b += pass_bytes(i);
  return b;
}

size_t DemagPlan::device_bytes() const { return A_.bytes() + B_.bytes() + T_.bytes(); }

void DemagPlan::copy_spectrum(float* host) const {
  HIP_CHECK(hipMemcpy(host, T_.get(), T_.bytes(), hipMemcpyDeviceToHost));
}

void reference_kernel_size(const DemagGeometry& g, int ksize[3]) {
  for (int a = 0; a < 3; ++a) {
    int s = 2 * g.n[a] - 1;
    if (s > 5 && (s % 2) == 1) s += 1;
    ksize[a] = s;
  }
}

void newell_grid_device(float* out, const int ksize[3], const int origin[3], const DemagGeometry& g, hipStream_t s) {
  const long long n = (long long)ksize[0] * ksize[1] * ksize[2];
  k_newell_grid<<<grid_for(n), BLOCK, 0, s>>>(out, ksize[0], ksize[1], ksize[2], origin[0], origin[1], origin[2],
                                              g.cs[0], g.cs[1], g.cs[2], g.mastergrid[0], g.mastergrid[1],
                                              g.mastergrid[2], g.pbc[0], g.pbc[1], g.pbc[2]);
  checkLaunch("k_newell_grid");
}

void reference_kernel(const DemagGeometry& g, int ksize[3], float* host_out) {
  reference_kernel_size(g, ksize);
  int origin[3];
  for (int a = 0; a < 3; ++a) origin[a] = -(g.n[a] - 1) - (ksize[a] - (2 * g.n[a] - 1));
  const long long n = (long long)ksize[0] * ksize[1] * ksize[2];
  DeviceArray<float> out((size_t)6 * n);
  k_newell_grid<<<grid_for(n), BLOCK>>>(out.get(), ksize[0], ksize[1], ksize[2], origin[0], origin[1], origin[2],
                                        g.cs[0], g.cs[1], g.cs[2], g.mastergrid[0], g.mastergrid[1], g.mastergrid[2],
                                        g.pbc[0], g.pbc[1], g.pbc[2]);
  checkLaunch("k_newell_grid");
  HIP_CHECK(hipMemcpy(host_out, out.get(), out.bytes(), hipMemcpyDeviceToHost));
}

}  // namespace mpa
