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
// threads x 8 values with two LDS exchanges.
//
// Kernel spectrum: the Newell kernel on offsets -(N-1)..(N-1) placed
// circularly is even (xx, yy, zz) or odd (xy: x,y; xz: x,z; yz: y,z) along
// each axis, so its DFT is real and only the octant kx<Hx, ky<Hy, kz<Hz is
// stored (6 floats per entry), with the sign of odd components flipped for
// ky > Ky/2 and kz > Kz/2.  It is built in FP64 as separable cosine / sine
// sums of the FP64 octant of the real-space kernel, pre-scaled by
// -mu0/(Kx Ky Kz) (the unnormalised inverse), and rounded to FP32 once.
#include "demag.hpp"
#include "fft.hpp"
#include "newell.hpp"

#include <cmath>
#include <vector>

namespace mpa {

namespace {

constexpr double MU0 = 1.25663706212e-6;  // src/physics/constants.hpp:22
constexpr int BLOCK = 256;
constexpr int MAXK = 2048;       // longest FFT line supported per axis
constexpr int MAXTHREADS = 512;  // threads per block of the FFT passes

int ilog2(int v) {
  int l = 0;
  while ((1 << l) < v) ++l;
  return l;
}

struct DD {
  int nx, ny, nz;
  int Kx, Ky, Kz;
  int Hx, Hy, Hz;
  int l2x, l2y, l2z;
  int nlines;  // 3*ny*nz real x-lines
};

__host__ __device__ __forceinline__ int threads_per_line(int K) { return K >= 8 ? K >> 3 : 1; }
__host__ __device__ __forceinline__ int elems_per_thread(int K) { return K >= 8 ? 8 : 1; }

// ---------------------------------------------------------------- kernels

__device__ __forceinline__ float load_M(const float* __restrict__ m, PS msat, const uint8_t* __restrict__ geom,
                                        const DD& d, int line, int x) {
  const int nyz = d.ny * d.nz;
  const int cell = (line % nyz) * d.nx + x;
  if (geom && !geom[cell]) return 0.0f;
  return msat.at(cell) * m[(size_t)line * d.nx + x];
}

// Offset of x-line l (c, z, y) in A[c][z][kx][y] at kx = 0; -1 past the end.
__device__ __forceinline__ long long a_line_base(const DD& d, int l) {
  if (l >= d.nlines) return -1;
  const int yz = l / d.ny;
  return (long long)yz * d.Hx * d.ny + (l - yz * d.ny);
}

// x forward: PB line pairs per block (PB a power of two), T = Kx/8 threads
// per pair.
__global__ void __launch_bounds__(MAXTHREADS) k_xfwd(DD d, const float* __restrict__ m, PS msat,
                                                     const uint8_t* __restrict__ geom, float2* __restrict__ A,
                                                     const float2* __restrict__ tw, int l2PB) {
  extern __shared__ float2 lds[];
  __shared__ long long abase[64];
  const int PB = 1 << l2PB;
  const int T = threads_per_line(d.Kx), E = elems_per_thread(d.Kx);
  const int LP = padded_len(d.Kx);
  const int g = threadIdx.x / T, t = threadIdx.x - g * T;
  const int l0 = 2 * blockIdx.x * PB;
  const int la = l0 + 2 * g, lb = la + 1;
  // blocks can have fewer threads than lines (one thread per line when Kx < 8)
  for (int i = threadIdx.x; i < 2 * PB; i += blockDim.x) abase[i] = a_line_base(d, l0 + i);
  const int nyz = d.ny * d.nz;
  const int cella = (la % nyz) * d.nx, cellb = (lb % nyz) * d.nx;
  float2 v[1][8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int x = t + T * q;
    float2 a = make_float2(0.f, 0.f);
    if (q < E && x < d.nx) {
      if (la < d.nlines && (!geom || geom[cella + x])) a.x = msat.at(cella + x) * m[(size_t)la * d.nx + x];
      if (lb < d.nlines && (!geom || geom[cellb + x])) a.y = msat.at(cellb + x) * m[(size_t)lb * d.nx + x];
    }
    v[0][q] = a;
  }
  if (d.Kx >= 8) line_fft<false, 1>(v, t, d.l2x, lds + g * LP, 0, tw);
#pragma unroll
  for (int q = 0; q < 8; ++q)
    if (q < E) lds[g * LP + lpad(t + T * q)] = v[0][q];
  __syncthreads();
  const int l2nl = l2PB + 1;
  for (int e = threadIdx.x; e < (d.Hx << l2nl); e += blockDim.x) {
    const int kx = e >> l2nl, li = e & ((1 << l2nl) - 1);
    const long long base = abase[li];
    if (base < 0) continue;
    const int p = li >> 1;
    const float2 z1 = lds[p * LP + lpad(kx)];
    const float2 z2 = lds[p * LP + lpad((d.Kx - kx) & (d.Kx - 1))];
    float2 X;
    if (!(li & 1)) X = make_float2(0.5f * (z1.x + z2.x), 0.5f * (z1.y - z2.y));
    else X = make_float2(0.5f * (z1.y + z2.y), -0.5f * (z1.x - z2.x));
    A[base + (long long)kx * d.ny] = X;
  }
}

// Offset of y-line q = (c, kx, z) in B[c][kx][ky][z] at ky = 0; -1 past the end.
__device__ __forceinline__ long long b_line_base(const DD& d, int q) {
  if (q >= 3 * d.Hx * d.nz) return -1;
  const int ckx = q / d.nz, z = q - ckx * d.nz;
  return (long long)ckx * d.Ky * d.nz + z;
}

// y forward (3-D): LB lines (c, kx, z), z fastest, per block (LB a power of two).
__global__ void __launch_bounds__(MAXTHREADS) k_yfwd(DD d, const float2* __restrict__ A, float2* __restrict__ B,
                                                     const float2* __restrict__ tw, int l2LB) {
  extern __shared__ float2 lds[];
  __shared__ long long bbase[64];
  const int LB = 1 << l2LB;
  const int T = threads_per_line(d.Ky), E = elems_per_thread(d.Ky);
  const int LP = padded_len(d.Ky);
  const int nq = 3 * d.Hx * d.nz;
  const int g = threadIdx.x / T, t = threadIdx.x - g * T;
  const int q0 = blockIdx.x * LB;
  for (int i = threadIdx.x; i < LB; i += blockDim.x) bbase[i] = b_line_base(d, q0 + i);
  const int ql = q0 + g;
  float2 v[1][8];
  const float2* src = nullptr;
  if (ql < nq) {
    const int ckx = ql / d.nz, z = ql - ckx * d.nz;
    const int c = ckx / d.Hx, kx = ckx - c * d.Hx;
    src = A + ((size_t)(c * d.nz + z) * d.Hx + kx) * d.ny;
  }
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int y = t + T * q;
    v[0][q] = (src && q < E && y < d.ny) ? src[y] : make_float2(0.f, 0.f);
  }
  if (d.Ky >= 8) line_fft<false, 1>(v, t, d.l2y, lds + g * LP, 0, tw);
#pragma unroll
  for (int q = 0; q < 8; ++q)
    if (q < E) lds[g * LP + lpad(t + T * q)] = v[0][q];
  __syncthreads();
  for (int e = threadIdx.x; e < (d.Ky << l2LB); e += blockDim.x) {
    const int ky = e >> l2LB, li = e & (LB - 1);
    const long long base = bbase[li];
    if (base < 0) continue;
    B[base + (long long)ky * d.nz] = lds[li * LP + lpad(ky)];
  }
}

// h = T . m at one spectral point; T components are xx, yy, zz, xy, xz, yz.
__device__ __forceinline__ void kernel_multiply(const float* __restrict__ T, const DD& d, int kx, int ky, int kz,
                                                float2& a, float2& b, float2& c) {
  int kyp = ky, kzp = kz;
  float sy = 1.f, sz = 1.f;
  if (ky > d.Ky / 2) { kyp = d.Ky - ky; sy = -1.f; }
  if (kz > d.Kz / 2) { kzp = d.Kz - kz; sz = -1.f; }
  const size_t stride = (size_t)d.Hx * d.Hy * d.Hz;
  const size_t base = ((size_t)kx * d.Hy + kyp) * d.Hz + kzp;
  const float nxx = T[base];
  const float nyy = T[stride + base];
  const float nzz = T[2 * stride + base];
  const float nxy = sy * T[3 * stride + base];
  const bool flat = d.Kz == 1;  // single layer: xz and yz vanish
  const float nxz = flat ? 0.f : sz * T[4 * stride + base];
  const float nyz = flat ? 0.f : sy * sz * T[5 * stride + base];
  const float2 ma = a, mb = b, mc = c;
  a = make_float2(nxx * ma.x + nxy * mb.x + nxz * mc.x, nxx * ma.y + nxy * mb.y + nxz * mc.y);
  b = make_float2(nxy * ma.x + nyy * mb.x + nyz * mc.x, nxy * ma.y + nyy * mb.y + nyz * mc.y);
  c = make_float2(nxz * ma.x + nyz * mb.x + nzz * mc.x, nxz * ma.y + nyz * mb.y + nzz * mc.y);
}

// z pass (3-D): PT spectral points (kx, ky) per block; each thread holds
// the three components of its 8 z values, so the kernel multiply between
// the forward and the inverse transform stays in registers.
__global__ void __launch_bounds__(MAXTHREADS) k_zpass(DD d, float2* __restrict__ B, const float* __restrict__ Tk,
                                                      const float2* __restrict__ tw, int PT) {
  extern __shared__ float2 lds[];
  const int T = threads_per_line(d.Kz);
  const int LP = padded_len(d.Kz);
  const int npt = d.Hx * d.Ky;
  const size_t cs = (size_t)npt;  // points per component
  const int g = threadIdx.x / T, t = threadIdx.x - g * T;
  const int pt = blockIdx.x * PT + g;
  const bool valid = pt < npt;
  float2 v[3][8];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int z = t + T * q;
      v[c][q] = (valid && z < d.nz) ? B[((size_t)c * cs + pt) * d.nz + z] : make_float2(0.f, 0.f);
    }
  line_fft<false, 3>(v, t, d.l2z, lds + g * LP, PT * LP, tw);
  if (valid) {
    const int kx = pt >> d.l2y, ky = pt & (d.Ky - 1);
#pragma unroll
    for (int q = 0; q < 8; ++q) kernel_multiply(Tk, d, kx, ky, t + T * q, v[0][q], v[1][q], v[2][q]);
  }
  line_fft<true, 3>(v, t, d.l2z, lds + g * LP, PT * LP, tw);
  if (valid) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int z = t + T * q;
        if (z < d.nz) B[((size_t)c * cs + pt) * d.nz + z] = v[c][q];
      }
  }
}

// y inverse (3-D): B -> A, keeps y < Ny.
__global__ void __launch_bounds__(MAXTHREADS) k_yinv(DD d, const float2* __restrict__ B, float2* __restrict__ A,
                                                     const float2* __restrict__ tw, int l2LB) {
  extern __shared__ float2 lds[];
  __shared__ long long bbase[64];
  const int LB = 1 << l2LB;
  const int T = threads_per_line(d.Ky), E = elems_per_thread(d.Ky);
  const int LP = padded_len(d.Ky);
  const int nq = 3 * d.Hx * d.nz;
  const int q0 = blockIdx.x * LB;
  for (int i = threadIdx.x; i < LB; i += blockDim.x) bbase[i] = b_line_base(d, q0 + i);
  __syncthreads();
  for (int e = threadIdx.x; e < (d.Ky << l2LB); e += blockDim.x) {
    const int ky = e >> l2LB, li = e & (LB - 1);
    const long long base = bbase[li];
    lds[li * LP + lpad(ky)] = base >= 0 ? B[base + (long long)ky * d.nz] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  const int g = threadIdx.x / T, t = threadIdx.x - g * T;
  float2 v[1][8];
#pragma unroll
  for (int q = 0; q < 8; ++q) v[0][q] = q < E ? lds[g * LP + lpad(t + T * q)] : make_float2(0.f, 0.f);
  __syncthreads();
  if (d.Ky >= 8) line_fft<true, 1>(v, t, d.l2y, lds + g * LP, 0, tw);
  const int ql = q0 + g;
  if (ql < nq) {
    const int ckx = ql / d.nz, z = ql - ckx * d.nz;
    const int c = ckx / d.Hx, kx = ckx - c * d.Hx;
    float2* dst = A + ((size_t)(c * d.nz + z) * d.Hx + kx) * d.ny;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int y = t + T * q;
      if (q < E && y < d.ny) dst[y] = v[0][q];
    }
  }
}

// single layer: forward y, kernel multiply, inverse y; PT points kx per block.
__global__ void __launch_bounds__(MAXTHREADS) k_ypass2d(DD d, float2* __restrict__ A, const float* __restrict__ Tk,
                                                        const float2* __restrict__ tw, int PT) {
  extern __shared__ float2 lds[];
  const int T = threads_per_line(d.Ky), E = elems_per_thread(d.Ky);
  const int LP = padded_len(d.Ky);
  const int g = threadIdx.x / T, t = threadIdx.x - g * T;
  const int kx = blockIdx.x * PT + g;
  const bool valid = kx < d.Hx;
  float2 v[3][8];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int y = t + T * q;
      v[c][q] = (valid && q < E && y < d.ny) ? A[((size_t)c * d.Hx + kx) * d.ny + y] : make_float2(0.f, 0.f);
    }
  if (d.Ky >= 8) line_fft<false, 3>(v, t, d.l2y, lds + g * LP, PT * LP, tw);
  if (valid) {
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (q < E) kernel_multiply(Tk, d, kx, t + T * q, 0, v[0][q], v[1][q], v[2][q]);
  }
  if (d.Ky >= 8) line_fft<true, 3>(v, t, d.l2y, lds + g * LP, PT * LP, tw);
  if (valid) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int y = t + T * q;
        if (q < E && y < d.ny) A[((size_t)c * d.Hx + kx) * d.ny + y] = v[c][q];
      }
  }
}

// x inverse: PB line pairs per block; C2R of two lines at a time.
__global__ void __launch_bounds__(MAXTHREADS) k_xinv(DD d, const float2* __restrict__ A,
                                                     const uint8_t* __restrict__ geom, float* __restrict__ h,
                                                     const float2* __restrict__ tw, int l2PB) {
  extern __shared__ float2 lds[];
  __shared__ long long abase[64];
  const int PB = 1 << l2PB;
  const int T = threads_per_line(d.Kx), E = elems_per_thread(d.Kx);
  const int LP = padded_len(d.Kx);
  const int XS = d.Hx + 1;
  const int l2nl = l2PB + 1;
  const int nl = 2 * PB;
  float2* X = lds;
  float2* Z = lds + nl * XS;
  const int l0 = 2 * blockIdx.x * PB;
  for (int i = threadIdx.x; i < nl; i += blockDim.x) abase[i] = a_line_base(d, l0 + i);
  __syncthreads();
  for (int e = threadIdx.x; e < (d.Hx << l2nl); e += blockDim.x) {
    const int kx = e >> l2nl, li = e & (nl - 1);
    const long long base = abase[li];
    X[li * XS + kx] = base >= 0 ? A[base + (long long)kx * d.ny] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  const int g = threadIdx.x / T, t = threadIdx.x - g * T;
  const int half = d.Kx >> 1;
  float2 v[1][8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int k = t + T * q;
    float2 zv = make_float2(0.f, 0.f);
    if (q < E) {
      const bool upper = k > half;
      const int kk = upper ? d.Kx - k : k;
      float2 xa = X[(2 * g) * XS + kk];
      float2 xb = X[(2 * g + 1) * XS + kk];
      if (kk == 0 || kk == half) { xa.y = 0.f; xb.y = 0.f; }
      zv = upper ? make_float2(xa.x + xb.y, xb.x - xa.y) : make_float2(xa.x - xb.y, xa.y + xb.x);
    }
    v[0][q] = zv;
  }
  if (d.Kx >= 8) line_fft<true, 1>(v, t, d.l2x, Z + g * LP, 0, tw);
  const int la = l0 + 2 * g, lb = la + 1;
  const int nyz = d.ny * d.nz;
  const int cella = (la % nyz) * d.nx, cellb = (lb % nyz) * d.nx;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int x = t + T * q;
    if (q >= E || x >= d.nx) continue;
    if (la < d.nlines) h[(size_t)la * d.nx + x] = (geom && !geom[cella + x]) ? 0.f : v[0][q].x;
    if (lb < d.nlines) h[(size_t)lb * d.nx + x] = (geom && !geom[cellb + x]) ? 0.f : v[0][q].y;
  }
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
  // output index ((c*Hx + kx)*Hy + ky)*Hz + kz
  const int kz = (int)(t % Hz);
  const int ky = (int)((t / Hz) % Hy);
  const int kx = (int)((t / ((long long)Hz * Hy)) % Hx);
  const int c = (int)(t / ((long long)Hz * Hy * Hx));
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

void make_twiddles(DeviceArray<float2>& tw, int K) {
  std::vector<float2> h(K);
  for (int q = 0; q < K; ++q) {
    const double a = -2.0 * M_PI * (double)q / (double)K;
    h[q] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  tw.allocate(K);
  HIP_CHECK(hipMemcpy(tw.get(), h.data(), K * sizeof(float2), hipMemcpyHostToDevice));
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

constexpr size_t LDS_LIMIT = 65536;  // dynamic LDS per workgroup used by every pass

void check_launch_shape(size_t lds_bytes, int threads, const char* pass) {
  // 1 KB is left for the static per-line offset tables of the passes
  if (lds_bytes + 1024 > LDS_LIMIT) throw Error(MPA_ERR_RUNTIME, std::string("demag: LDS budget exceeded in ") + pass);
  if (threads < 1 || threads > MAXTHREADS)
    throw Error(MPA_ERR_RUNTIME, std::string("demag: bad block size in ") + pass);
}

int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

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
    if (K_[a] > MAXK) throw Error(MPA_ERR_INVALID, "demag: grid axis too large for the FFT passes (max 1024 cells)");
  }
  N_ = g.n[0] * g.n[1] * g.n[2];
  Hx_ = K_[0] == 1 ? 1 : K_[0] / 2 + 1;
  Hy_ = K_[1] / 2 + 1;
  Hz_ = K_[2] / 2 + 1;
  A_.allocate((size_t)3 * g.n[2] * Hx_ * g.n[1]);
  if (!is2d()) B_.allocate((size_t)3 * Hx_ * K_[1] * g.n[2]);
  for (int a = 0; a < 3; ++a) make_twiddles(tw_[a], K_[a]);
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

void DemagPlan::exec(const float* m, PS msat, const uint8_t* geom, float* h, hipStream_t s, Profiler* prof) {
  DD d;
  d.nx = g_.n[0]; d.ny = g_.n[1]; d.nz = g_.n[2];
  d.Kx = K_[0]; d.Ky = K_[1]; d.Kz = K_[2];
  d.Hx = Hx_; d.Hy = Hy_; d.Hz = Hz_;
  d.l2x = log2K_[0]; d.l2y = log2K_[1]; d.l2z = log2K_[2];
  d.nlines = 3 * d.ny * d.nz;
  const int Tx = threads_per_line(d.Kx), Ty = threads_per_line(d.Ky), Tz = threads_per_line(d.Kz);
  const int LPx = padded_len(d.Kx), LPy = padded_len(d.Ky), LPz = padded_len(d.Kz);
  const int npairs = (d.nlines + 1) / 2;
  {
    const int PB = clampi(MAXTHREADS / Tx, 1, 32);
    const size_t shm = (size_t)PB * LPx * sizeof(float2);
    check_launch_shape(shm, PB * Tx, "k_xfwd");
    Timed t(prof, "k_xfwd", pass_bytes(0), s);
    k_xfwd<<<grid_for(npairs, PB), PB * Tx, shm, s>>>(d, m, msat, geom, A_.get(), tw_[0].get(), ilog2(PB));
    checkLaunch("k_xfwd");
  }
  if (is2d()) {
    const int PT = clampi(BLOCK / Ty, 1, BLOCK);
    const size_t shm = (size_t)3 * PT * LPy * sizeof(float2);
    check_launch_shape(shm, PT * Ty, "k_ypass2d");
    Timed t(prof, "k_ypass2d", pass_bytes(1), s);
    k_ypass2d<<<grid_for(d.Hx, PT), PT * Ty, shm, s>>>(d, A_.get(), T_.get(), tw_[1].get(), PT);
    checkLaunch("k_ypass2d");
  } else {
    const int LB = clampi(MAXTHREADS / Ty, 1, 32);
    const size_t shmy = (size_t)LB * LPy * sizeof(float2);
    check_launch_shape(shmy, LB * Ty, "k_yfwd");
    const long long nq = 3LL * d.Hx * d.nz;
    {
      Timed t(prof, "k_yfwd", pass_bytes(1), s);
      k_yfwd<<<grid_for(nq, LB), LB * Ty, shmy, s>>>(d, A_.get(), B_.get(), tw_[1].get(), ilog2(LB));
      checkLaunch("k_yfwd");
    }
    {
      const int PT = clampi(BLOCK / Tz, 1, BLOCK);
      const size_t shm = (size_t)3 * PT * LPz * sizeof(float2);
      check_launch_shape(shm, PT * Tz, "k_zpass");
      Timed t(prof, "k_zpass", pass_bytes(2), s);
      k_zpass<<<grid_for((long long)d.Hx * d.Ky, PT), PT * Tz, shm, s>>>(d, B_.get(), T_.get(), tw_[2].get(), PT);
      checkLaunch("k_zpass");
    }
    {
      Timed t(prof, "k_yinv", pass_bytes(3), s);
      k_yinv<<<grid_for(nq, LB), LB * Ty, shmy, s>>>(d, B_.get(), A_.get(), tw_[1].get(), ilog2(LB));
      checkLaunch("k_yinv");
    }
  }
  {
    const int PB = clampi(BLOCK / Tx, 1, 32);
    const size_t shm = ((size_t)2 * PB * (d.Hx + 1) + (size_t)PB * LPx) * sizeof(float2);
    check_launch_shape(shm, PB * Tx, "k_xinv");
    Timed t(prof, "k_xinv", pass_bytes(npasses() - 1), s);
    k_xinv<<<grid_for(npairs, PB), PB * Tx, shm, s>>>(d, A_.get(), geom, h, tw_[0].get(), ilog2(PB));
    checkLaunch("k_xinv");
  }
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
