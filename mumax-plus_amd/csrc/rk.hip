// Runge-Kutta stage combination and reductions.
//
// k_combine replaces the per-stage chain of addTo + normalise launches of the
// reference (RungeKuttaStageExecutor::setStageX / setFinalX,
// src/core/rungekutta.cpp:108-134; k_addFields src/core/fieldops.cu:9-22;
// k_normalize :232-247): one read of x0 and of each k_j, one write.
// k_combo_maxnorm replaces getError (rungekutta.cpp:140-151) + k_maxVecNorm
// (src/core/reduce.cu:53-99) without materialising the error field.
// k_sums replaces k_average (reduce.cu:101-127) with a deterministic
// two-level FP64 sum.
#include "rk.hpp"

namespace mpa {

namespace {

constexpr int BLOCK = 256;
constexpr int SUM_BLOCKS = 512;

__global__ void __launch_bounds__(BLOCK) k_combine(float* __restrict__ out, const float* __restrict__ x0, Combo co,
                                                   int ncomp, int N, const uint8_t* __restrict__ geom, int normalize) {
  const int idx = xcd_remap(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
  if (idx >= N) return;
  if (geom && !geom[idx]) {
    for (int c = 0; c < ncomp; ++c) out[(size_t)c * N + idx] = 0.f;
    return;
  }
  float v[3];
  for (int c = 0; c < ncomp; ++c) {
    float a = x0[(size_t)c * N + idx];
    for (int j = 0; j < co.nk; ++j) a += co.c[j] * co.k[j][(size_t)c * N + idx];
    v[c] = a;
  }
  if (normalize) {
    float n2 = 0.f;
    for (int c = 0; c < ncomp; ++c) n2 += v[c] * v[c];
    const float inv = rsqrtf(n2);
    for (int c = 0; c < ncomp; ++c) v[c] *= inv;
  }
  for (int c = 0; c < ncomp; ++c) out[(size_t)c * N + idx] = v[c];
}

// Same as k_combine for 3-vectors, four consecutive cells per thread with
// 16-byte loads and stores (N divisible by 4).
__global__ void __launch_bounds__(BLOCK) k_combine4(float* __restrict__ out, const float* __restrict__ x0, Combo co,
                                                    int N, const uint8_t* __restrict__ geom, int normalize) {
  const int i4 = xcd_remap(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
  const int N4 = N >> 2;
  if (i4 >= N4) return;
  float4 v[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float4 a = reinterpret_cast<const float4*>(x0 + (size_t)c * N)[i4];
    for (int j = 0; j < co.nk; ++j) {
      const float4 kv = reinterpret_cast<const float4*>(co.k[j] + (size_t)c * N)[i4];
      const float cj = co.c[j];
      a.x += cj * kv.x;
      a.y += cj * kv.y;
      a.z += cj * kv.z;
      a.w += cj * kv.w;
    }
    v[c] = a;
  }
  if (normalize) {
    const float i0 = rsqrtf(v[0].x * v[0].x + v[1].x * v[1].x + v[2].x * v[2].x);
    const float i1 = rsqrtf(v[0].y * v[0].y + v[1].y * v[1].y + v[2].y * v[2].y);
    const float i2 = rsqrtf(v[0].z * v[0].z + v[1].z * v[1].z + v[2].z * v[2].z);
    const float i3 = rsqrtf(v[0].w * v[0].w + v[1].w * v[1].w + v[2].w * v[2].w);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      v[c].x *= i0;
      v[c].y *= i1;
      v[c].z *= i2;
      v[c].w *= i3;
    }
  }
  if (geom) {
    const uchar4 g = reinterpret_cast<const uchar4*>(geom)[i4];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if (!g.x) v[c].x = 0.f;
      if (!g.y) v[c].y = 0.f;
      if (!g.z) v[c].z = 0.f;
      if (!g.w) v[c].w = 0.f;
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) reinterpret_cast<float4*>(out + (size_t)c * N)[i4] = v[c];
}

struct CombineBatch {
  CombineJob job[MAX_JOBS];
};

// One job per blockIdx.y; four consecutive cells per thread when N is a
// multiple of 4, one otherwise.  Same arithmetic as k_combine / k_combine4.
// NK >= 0: every job combines NK derivatives, a compile-time count, so the
// loop unrolls and all of a thread's loads are in flight together; NK < 0
// reads each job's count.
template <int NK>
__global__ void __launch_bounds__(BLOCK) k_combine_batch(CombineBatch b) {
  const CombineJob& jb = b.job[blockIdx.y];
  const int N = jb.N;
  const Combo& co = jb.co;
  if ((N & 3) == 0) {
    const int i4 = xcd_remap(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (i4 >= (N >> 2)) return;
    float4 v[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      float4 a = reinterpret_cast<const float4*>(jb.x0 + (size_t)c * N)[i4];
#pragma unroll
      for (int j = 0; j < (NK >= 0 ? NK : MAX_STAGES); ++j) {
        if (NK < 0 && j >= co.nk) break;
        const float4 kv = reinterpret_cast<const float4*>(co.k[j] + (size_t)c * N)[i4];
        const float cj = co.c[j];
        a.x += cj * kv.x;
        a.y += cj * kv.y;
        a.z += cj * kv.z;
        a.w += cj * kv.w;
      }
      v[c] = a;
    }
    if (jb.normalize) {
      const float i0 = rsqrtf(v[0].x * v[0].x + v[1].x * v[1].x + v[2].x * v[2].x);
      const float i1 = rsqrtf(v[0].y * v[0].y + v[1].y * v[1].y + v[2].y * v[2].y);
      const float i2 = rsqrtf(v[0].z * v[0].z + v[1].z * v[1].z + v[2].z * v[2].z);
      const float i3 = rsqrtf(v[0].w * v[0].w + v[1].w * v[1].w + v[2].w * v[2].w);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        v[c].x *= i0;
        v[c].y *= i1;
        v[c].z *= i2;
        v[c].w *= i3;
      }
    }
    if (jb.geom) {
      const uchar4 g = reinterpret_cast<const uchar4*>(jb.geom)[i4];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (!g.x) v[c].x = 0.f;
        if (!g.y) v[c].y = 0.f;
        if (!g.z) v[c].z = 0.f;
        if (!g.w) v[c].w = 0.f;
      }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      reinterpret_cast<float4*>(jb.out + (size_t)c * N)[i4] = v[c];
      if (jb.out2) reinterpret_cast<float4*>(jb.out2 + (size_t)c * N)[i4] = v[c];
    }
    return;
  }
  const int idx = xcd_remap(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
  if (idx >= N) return;
  float v[3];
  const bool in = !jb.geom || jb.geom[idx];
  for (int c = 0; c < 3; ++c) {
    float a = jb.x0[(size_t)c * N + idx];
#pragma unroll
    for (int j = 0; j < (NK >= 0 ? NK : MAX_STAGES); ++j) {
      if (NK < 0 && j >= co.nk) break;
      a += co.c[j] * co.k[j][(size_t)c * N + idx];
    }
    v[c] = a;
  }
  if (jb.normalize) {
    const float inv = rsqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int c = 0; c < 3; ++c) v[c] *= inv;
  }
  for (int c = 0; c < 3; ++c) {
    const float r = in ? v[c] : 0.f;
    jb.out[(size_t)c * N + idx] = r;
    if (jb.out2) jb.out2[(size_t)c * N + idx] = r;
  }
}

__device__ __forceinline__ float block_max(float v) {
  __shared__ float sh[BLOCK / 64];
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_down(v, off, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < BLOCK / 64; ++i) v = fmaxf(v, sh[i]);
  }
  return v;
}

__global__ void __launch_bounds__(BLOCK) k_combo_maxnorm(Combo co, int N, const uint8_t* __restrict__ geom,
                                                         unsigned* __restrict__ result) {
  float best = 0.f;
  for (int idx = blockIdx.x * BLOCK + threadIdx.x; idx < N; idx += gridDim.x * BLOCK) {
    if (geom && !geom[idx]) continue;
    float e[3];
    for (int c = 0; c < 3; ++c) {
      float a = 0.f;
      for (int j = 0; j < co.nk; ++j) a += co.c[j] * co.k[j][(size_t)c * N + idx];
      e[c] = a;
    }
    best = fmaxf(best, sqrtf(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]));
  }
  best = block_max(best);
  if (threadIdx.x == 0) atomicMax(result, __float_as_uint(best));
}

__global__ void k_zero_u32(unsigned* p) { *p = 0u; }

__global__ void __launch_bounds__(BLOCK) k_maxabs(const float* __restrict__ v, int N, const uint8_t* __restrict__ geom,
                                                  unsigned* __restrict__ result) {
  float best = 0.f;
  for (int idx = blockIdx.x * BLOCK + threadIdx.x; idx < N; idx += gridDim.x * BLOCK) {
    if (geom && !geom[idx]) continue;
    best = fmaxf(best, fabsf(v[idx]));
  }
  best = block_max(best);
  if (threadIdx.x == 0) atomicMax(result, __float_as_uint(best));
}

__global__ void __launch_bounds__(BLOCK) k_sums(const float* __restrict__ v, int ncomp, int N,
                                                const uint8_t* __restrict__ geom, double* __restrict__ partials) {
  __shared__ double sh[BLOCK];
  for (int c = 0; c < ncomp; ++c) {
    double acc = 0.0;
    for (int idx = blockIdx.x * BLOCK + threadIdx.x; idx < N; idx += gridDim.x * BLOCK) {
      if (geom && !geom[idx]) continue;
      acc += (double)v[(size_t)c * N + idx];
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = BLOCK / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) partials[(size_t)c * gridDim.x + blockIdx.x] = sh[0];
    __syncthreads();
  }
}

// Per-block FP64 sums of |v|^2 (one component slot).
__global__ void __launch_bounds__(BLOCK) k_sumsq(const float* __restrict__ v, int N, const uint8_t* __restrict__ geom,
                                                 double* __restrict__ partials) {
  __shared__ double sh[BLOCK];
  double acc = 0.0;
  for (int idx = blockIdx.x * BLOCK + threadIdx.x; idx < N; idx += gridDim.x * BLOCK) {
    if (geom && !geom[idx]) continue;
    for (int c = 0; c < 3; ++c) {
      const double x = v[(size_t)c * N + idx];
      acc += x * x;
    }
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = BLOCK / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = sh[0];
}

// The three Barzilai-Borwein products of dm = m1 - m0 and dt = t0 - t1
// (minimizer.cu:76-90, 118-120) in FP64, and max |dm| (minimizer.cu:124).
__global__ void __launch_bounds__(BLOCK) k_bb_sums(const float* __restrict__ m1, const float* __restrict__ m0,
                                                   const float* __restrict__ t1, const float* __restrict__ t0, int N,
                                                   const uint8_t* __restrict__ geom, double* __restrict__ partials,
                                                   unsigned* __restrict__ maxbits) {
  __shared__ double sh[BLOCK];
  double acc[3] = {0.0, 0.0, 0.0};
  float best = 0.f;
  for (int idx = blockIdx.x * BLOCK + threadIdx.x; idx < N; idx += gridDim.x * BLOCK) {
    if (geom && !geom[idx]) continue;
    float dm2 = 0.f;
    for (int c = 0; c < 3; ++c) {
      const size_t i = (size_t)c * N + idx;
      const float dm = m1[i] - m0[i];
      const float dt = t0[i] - t1[i];
      acc[0] += (double)dm * dm;
      acc[1] += (double)dm * dt;
      acc[2] += (double)dt * dt;
      dm2 += dm * dm;
    }
    best = fmaxf(best, sqrtf(dm2));
  }
  for (int c = 0; c < 3; ++c) {
    sh[threadIdx.x] = acc[c];
    __syncthreads();
    for (int s = BLOCK / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) partials[(size_t)c * gridDim.x + blockIdx.x] = sh[0];
    __syncthreads();
  }
  best = block_max(best);
  if (threadIdx.x == 0) atomicMax(maxbits, __float_as_uint(best));
}

// One descent step of every cell (k_step, minimizer.cu:59-74):
// m = ((4 - h^2 |t|^2) m0 + 4 h t) / (4 + h^2 |t|^2).
__global__ void __launch_bounds__(BLOCK) k_descent(float* __restrict__ m, const float* __restrict__ m0,
                                                   const float* __restrict__ t, float h, int N) {
  const int idx = blockIdx.x * BLOCK + threadIdx.x;
  if (idx >= N) return;
  float a[3], b[3];
  for (int c = 0; c < 3; ++c) {
    a[c] = m0[(size_t)c * N + idx];
    b[c] = t[(size_t)c * N + idx];
  }
  const float t2 = h * h * (b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
  for (int c = 0; c < 3; ++c) m[(size_t)c * N + idx] = ((4.f - t2) * a[c] + 4.f * h * b[c]) / (4.f + t2);
}

__global__ void k_sums_final(double* partials, int nparts, int ncomp) {
  __shared__ double sh[BLOCK];
  for (int c = 0; c < ncomp; ++c) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < nparts; i += BLOCK) acc += partials[(size_t)c * nparts + i];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = BLOCK / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
      __syncthreads();
    }
    __syncthreads();
    if (threadIdx.x == 0) partials[(size_t)ncomp * nparts + c] = sh[0];
    __syncthreads();
  }
}

int reduce_blocks(int N) {
  int b = (N + BLOCK - 1) / BLOCK;
  return b > 1024 ? 1024 : (b < 1 ? 1 : b);
}

}  // namespace

void launch_combine(float* out, const float* x0, const Combo& co, int ncomp, int N, const uint8_t* geom,
                    bool normalize, hipStream_t s) {
  if (ncomp == 3 && (N & 3) == 0) {
    const int nblk = (N / 4 + BLOCK - 1) / BLOCK;
    k_combine4<<<nblk, BLOCK, 0, s>>>(out, x0, co, N, geom, normalize ? 1 : 0);
    checkLaunch("k_combine4");
    return;
  }
  const int nblk = (N + BLOCK - 1) / BLOCK;
  k_combine<<<nblk, BLOCK, 0, s>>>(out, x0, co, ncomp, N, geom, normalize ? 1 : 0);
  checkLaunch("k_combine");
}

void launch_combine_batch(const CombineJob* jobs, int njobs, hipStream_t s) {
  if (njobs < 1) return;
  if (njobs > MAX_JOBS) throw Error(MPA_ERR_INVALID, "launch_combine_batch: too many variables");
  CombineBatch b{};
  int nmax = 0;
  for (int i = 0; i < njobs; ++i) {
    b.job[i] = jobs[i];
    const int per = (jobs[i].N & 3) == 0 ? jobs[i].N / 4 : jobs[i].N;
    nmax = per > nmax ? per : nmax;
  }
  const dim3 grid((unsigned)((nmax + BLOCK - 1) / BLOCK), (unsigned)njobs);
  int nk = jobs[0].co.nk;
  for (int i = 1; i < njobs; ++i)
    if (jobs[i].co.nk != nk) nk = -1;
  switch (nk) {
    case 0: k_combine_batch<0><<<grid, BLOCK, 0, s>>>(b); break;
    case 1: k_combine_batch<1><<<grid, BLOCK, 0, s>>>(b); break;
    case 2: k_combine_batch<2><<<grid, BLOCK, 0, s>>>(b); break;
    case 3: k_combine_batch<3><<<grid, BLOCK, 0, s>>>(b); break;
    case 4: k_combine_batch<4><<<grid, BLOCK, 0, s>>>(b); break;
    case 5: k_combine_batch<5><<<grid, BLOCK, 0, s>>>(b); break;
    case 6: k_combine_batch<6><<<grid, BLOCK, 0, s>>>(b); break;
    case 7: k_combine_batch<7><<<grid, BLOCK, 0, s>>>(b); break;
    default: k_combine_batch<-1><<<grid, BLOCK, 0, s>>>(b); break;
  }
  checkLaunch("k_combine_batch");
}

Reducer::Reducer() : maxbits_(1), partials_((size_t)3 * SUM_BLOCKS + 3) {}

float Reducer::max_combo_norm(const Combo& co, int N, const uint8_t* geom, hipStream_t s) {
  k_zero_u32<<<1, 1, 0, s>>>(maxbits_.get());
  k_combo_maxnorm<<<reduce_blocks(N), BLOCK, 0, s>>>(co, N, geom, maxbits_.get());
  checkLaunch("k_combo_maxnorm");
  unsigned bits = 0;
  HIP_CHECK(hipMemcpyAsync(&bits, maxbits_.get(), sizeof(unsigned), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  float r;
  static_assert(sizeof(r) == sizeof(bits), "float size");
  __builtin_memcpy(&r, &bits, sizeof(r));
  return r;
}

float Reducer::max_abs(const float* v, int N, const uint8_t* geom, hipStream_t s) {
  k_zero_u32<<<1, 1, 0, s>>>(maxbits_.get());
  k_maxabs<<<reduce_blocks(N), BLOCK, 0, s>>>(v, N, geom, maxbits_.get());
  checkLaunch("k_maxabs");
  unsigned bits = 0;
  HIP_CHECK(hipMemcpyAsync(&bits, maxbits_.get(), sizeof(unsigned), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  float r;
  __builtin_memcpy(&r, &bits, sizeof(r));
  return r;
}

float Reducer::max_vec_norm(const float* v, int N, const uint8_t* geom, hipStream_t s) {
  Combo co{};
  co.nk = 1;
  co.c[0] = 1.f;
  co.k[0] = v;
  return max_combo_norm(co, N, geom, s);
}

double Reducer::dot_sum(const float* v, int N, const uint8_t* geom, hipStream_t s) {
  int nb = (N + BLOCK - 1) / BLOCK;
  if (nb > SUM_BLOCKS) nb = SUM_BLOCKS;
  if (nb < 1) nb = 1;
  k_sumsq<<<nb, BLOCK, 0, s>>>(v, N, geom, partials_.get());
  checkLaunch("k_sumsq");
  k_sums_final<<<1, BLOCK, 0, s>>>(partials_.get(), nb, 1);
  checkLaunch("k_sums_final");
  double out = 0.0;
  HIP_CHECK(hipMemcpyAsync(&out, partials_.get() + nb, sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  return out;
}

void Reducer::bb_sums(const float* m1, const float* m0, const float* t1, const float* t0, int N,
                      const uint8_t* geom, double out[3], float* max_dm, hipStream_t s) {
  int nb = (N + BLOCK - 1) / BLOCK;
  if (nb > SUM_BLOCKS) nb = SUM_BLOCKS;
  if (nb < 1) nb = 1;
  k_zero_u32<<<1, 1, 0, s>>>(maxbits_.get());
  k_bb_sums<<<nb, BLOCK, 0, s>>>(m1, m0, t1, t0, N, geom, partials_.get(), maxbits_.get());
  checkLaunch("k_bb_sums");
  k_sums_final<<<1, BLOCK, 0, s>>>(partials_.get(), nb, 3);
  checkLaunch("k_sums_final");
  unsigned bits = 0;
  HIP_CHECK(hipMemcpyAsync(out, partials_.get() + (size_t)3 * nb, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipMemcpyAsync(&bits, maxbits_.get(), sizeof(unsigned), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  __builtin_memcpy(max_dm, &bits, sizeof(float));
}

void launch_descent(float* m, const float* m0, const float* t, float h, int N, hipStream_t s) {
  k_descent<<<(N + BLOCK - 1) / BLOCK, BLOCK, 0, s>>>(m, m0, t, h, N);
  checkLaunch("k_descent");
}

void Reducer::component_sums(const float* v, int ncomp, int N, const uint8_t* geom, double* out, hipStream_t s) {
  if (ncomp > 3) throw Error(MPA_ERR_INVALID, "component_sums: at most 3 components");
  int nb = (N + BLOCK - 1) / BLOCK;
  if (nb > SUM_BLOCKS) nb = SUM_BLOCKS;
  if (nb < 1) nb = 1;
  k_sums<<<nb, BLOCK, 0, s>>>(v, ncomp, N, geom, partials_.get());
  checkLaunch("k_sums");
  k_sums_final<<<1, BLOCK, 0, s>>>(partials_.get(), nb, ncomp);
  checkLaunch("k_sums_final");
  HIP_CHECK(hipMemcpyAsync(out, partials_.get() + (size_t)ncomp * nb, ncomp * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace mpa
