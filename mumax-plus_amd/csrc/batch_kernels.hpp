// Device bodies of the batch step's own kernels (stage combination, error
// norms, step controller), shared by the per-pass launches of batch.hip and
// the persistent step kernel.  The one-block / one-thread bodies take their
// block or member index as an argument.
#pragma once

#include <cmath>

#include "batch.hpp"
#include "rk.hpp"

namespace mpa {
namespace bk {

constexpr int CH = 256;  // cells per chunk; members start on chunk boundaries

struct StageArgs {
  const float* k[MAX_STAGES];
  double a[MAX_STAGES];
  int nk;
  int commit;
};

struct ErrArgs {
  const float* k[MAX_STAGES];
  double e[MAX_STAGES];
  int nk;
  const float* tq;
};

struct CtlArgs {
  int order1, order2;
  int prime;
};

__device__ __forceinline__ void axpy4(float4& a, float c, const float4& k) {
  a.x += c * k.x;
  a.y += c * k.y;
  a.z += c * k.z;
  a.w += c * k.w;
}

// One wavefront per chunk (64 threads x 4 consecutive cells), four chunks
// per block; a chunk belongs to one member, so every branch on the member is
// uniform over the wavefront.  X: the state (taken over as x0 when the
// member's last attempt was accepted and sa.commit), XO: the stage state
// (may be X: every element is read before it is written, by its thread).
__device__ __forceinline__ void bstage_block(const StageArgs& sa, const float* X, float* XO, float* __restrict__ X0,
                                             float* __restrict__ K0, const float* __restrict__ KL,
                                             const uint8_t* __restrict__ G, const int* __restrict__ c2m,
                                             const BatchMember* __restrict__ st, int cs, int nch, int blk) {
  const int chunk = blk * 4 + (threadIdx.x >> 6);
  if (chunk >= nch) return;
  const int m = c2m[chunk];
  if (m < 0) return;
  const BatchMember& S = st[m];
  if (!S.active) return;
  const int q = chunk * (CH / 4) + (threadIdx.x & 63);  // float4 index within a component
  const int cs4 = cs >> 2;
  const bool take = sa.commit && S.accept;
  const double dt = S.dt;
  float c[MAX_STAGES];
#pragma unroll
  for (int j = 0; j < MAX_STAGES; ++j) c[j] = j < sa.nk ? (float)(dt * sa.a[j]) : 0.f;
  const float4* X4 = reinterpret_cast<const float4*>(X);
  float4* XO4 = reinterpret_cast<float4*>(XO);
  float4* X04 = reinterpret_cast<float4*>(X0);
  float4* K04 = reinterpret_cast<float4*>(K0);
  const float4* KL4 = reinterpret_cast<const float4*>(KL);
  // every component's loads first (all in flight together), then the
  // take-over stores, then the combination
  float4 x0[3], k0[3];
  const float4* xs = take ? X4 : X04;
  const float4* ks = take ? KL4 : K04;
#pragma unroll
  for (int comp = 0; comp < 3; ++comp) {
    const size_t i = (size_t)comp * cs4 + q;
    x0[comp] = xs[i];
    k0[comp] = ks[i];
  }
  const uchar4 g = reinterpret_cast<const uchar4*>(G)[q];
  if (take) {
#pragma unroll
    for (int comp = 0; comp < 3; ++comp) {
      const size_t i = (size_t)comp * cs4 + q;
      X04[i] = x0[comp];
      K04[i] = k0[comp];
    }
  }
  float4 v[3];
#pragma unroll
  for (int comp = 0; comp < 3; ++comp) {
    const size_t i = (size_t)comp * cs4 + q;
    float4 a = x0[comp];
    axpy4(a, c[0], k0[comp]);
    for (int j = 1; j < sa.nk; ++j) axpy4(a, c[j], reinterpret_cast<const float4*>(sa.k[j])[i]);
    v[comp] = a;
  }
  const float i0 = rsqrtf(v[0].x * v[0].x + v[1].x * v[1].x + v[2].x * v[2].x);
  const float i1 = rsqrtf(v[0].y * v[0].y + v[1].y * v[1].y + v[2].y * v[2].y);
  const float i2 = rsqrtf(v[0].z * v[0].z + v[1].z * v[1].z + v[2].z * v[2].z);
  const float i3 = rsqrtf(v[0].w * v[0].w + v[1].w * v[1].w + v[2].w * v[2].w);
#pragma unroll
  for (int comp = 0; comp < 3; ++comp) {
    float4 o;
    o.x = g.x ? v[comp].x * i0 : 0.f;
    o.y = g.y ? v[comp].y * i1 : 0.f;
    o.z = g.z ? v[comp].z * i2 : 0.f;
    o.w = g.w ? v[comp].w * i3 : 0.f;
    XO4[(size_t)comp * cs4 + q] = o;
  }
}

__device__ __forceinline__ float bblock_max(float v, float* red) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) red[w] = v;
  __syncthreads();
  float r = 0.f;
  if (threadIdx.x == 0)
    for (int i = 0; i < CH / 64; ++i) r = fmaxf(r, red[i]);
  return r;
}

// red: 2 x CH/64 floats of LDS; every thread of the block calls this with
// the same chunk (the early returns are uniform over the block).
__device__ __forceinline__ void berr_block(const ErrArgs& ea, const uint8_t* __restrict__ G,
                                           const int* __restrict__ c2m, const BatchMember* __restrict__ st,
                                           unsigned* __restrict__ errbits, unsigned* __restrict__ tqbits, int cs,
                                           int chunk, float (*red)[CH / 64]) {
  const int m = c2m[chunk];
  if (m < 0) return;
  const BatchMember& S = st[m];
  if (!S.active) return;
  const bool want_err = S.adaptive && ea.nk > 0;
  const bool want_tq = ea.tq && (S.mode == BATCH_RELAX || S.accept == 2);
  if (!want_err && !want_tq) return;
  const int g = chunk * CH + threadIdx.x;
  float err = 0.f, tq = 0.f;
  if (G[g]) {
    if (want_err) {
      float c[MAX_STAGES];
#pragma unroll
      for (int j = 0; j < MAX_STAGES; ++j) c[j] = j < ea.nk ? (float)(S.dt * ea.e[j]) : 0.f;
      float e[3];
#pragma unroll
      for (int comp = 0; comp < 3; ++comp) {
        float a = 0.f;
        for (int j = 0; j < ea.nk; ++j) a += c[j] * ea.k[j][(size_t)comp * cs + g];
        e[comp] = a;
      }
      err = fmaxf(0.f, sqrtf(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]));
    }
    if (want_tq) {
      const float t0 = ea.tq[g], t1 = ea.tq[(size_t)cs + g], t2 = ea.tq[2 * (size_t)cs + g];
      tq = fmaxf(0.f, sqrtf(t0 * t0 + t1 * t1 + t2 * t2));
    }
  }
  const float be = bblock_max(err, red[0]);
  const float bt = bblock_max(tq, red[1]);
  if (threadIdx.x == 0) {
    if (want_err) atomicMax(errbits + m, __float_as_uint(be));
    if (want_tq) atomicMax(tqbits + m, __float_as_uint(bt));
  }
}

// Member slot m (one thread each).
__device__ __forceinline__ void bcontrol_member(BatchMember* __restrict__ st, unsigned* __restrict__ errbits,
                                                unsigned* __restrict__ tqbits, int nslots, const CtlArgs& c, int m) {
  if (m >= nslots) return;
  BatchMember S = st[m];
  if (!S.active) return;
  const float err = __uint_as_float(errbits[m]);
  const float tq = __uint_as_float(tqbits[m]);
  errbits[m] = 0u;
  tqbits[m] = 0u;
  if (c.prime) {
    if (S.accept != 2) return;
    // the initial step of TimeSolver.run / steps (timesolver.py:42-47,
    // timesolver.cpp:37-47): the sensible step when it is smaller
    if (S.adaptive && !S.keep_step) {
      const double sens = tq == 0.f ? S.sdefault : S.sfactor / (double)tq;
      if (S.dt == 0.0 || S.dt > sens) S.dt = sens;
    }
    S.last_torque = tq;
    S.accept = 1;  // the first combination takes x0 = x and k_0 = rhs(x)
    bool done = false;
    if (S.mode == BATCH_STEPS) done = S.target <= 0;
    else if (S.mode == BATCH_RELAX) done = tq <= S.threshold;
    else if (!(S.time < S.stop - S.dt)) {
      S.dt_saved = S.dt;
      S.dt = S.stop - S.time;
      S.last = 1;
    }
    if (done) {
      S.active = 0;
      S.done = 1;
    }
    st[m] = S;
    return;
  }
  const double h = S.dt;
  bool ok = true;
  if (S.adaptive) {
    ok = err < S.max_error;
    const double ratio = S.max_error / (double)err;
    double corr = ok ? pow(ratio, 1.0 / (c.order2 + 1)) : pow(ratio, 1.0 / (c.order1 + 1));
    if (isnan(corr)) corr = 1.0;
    corr *= S.headroom;
    corr = corr > S.upper ? S.upper : corr;
    corr = corr < S.lower ? S.lower : corr;
    S.dt = h * corr;
    S.last_error = err;
  }
  S.accept = ok ? 1 : 0;
  if (!ok) {
    ++S.rejected;
    st[m] = S;
    return;
  }
  S.time = S.time + h;
  ++S.steps;
  S.last_torque = tq;
  bool done = false;
  if (S.last) {
    done = true;
    if (!S.adaptive) S.dt = S.dt_saved;
  } else if (S.mode == BATCH_STEPS) {
    done = S.steps >= S.target;
  } else if (S.mode == BATCH_RELAX) {
    done = tq <= S.threshold || S.steps >= S.target;  // the target caps a relaxation that stalls
  } else if (!(S.time < S.stop - S.dt)) {
    S.dt_saved = S.dt;
    S.dt = S.stop - S.time;
    S.last = 1;
  }
  if (done) {
    S.active = 0;
    S.done = 1;
  }
  st[m] = S;
}

// ---------------------------------------------------------------------------
// Persistent step kernel: one launch runs n batch steps.  Each launch of the
// per-pass step becomes a phase; a phase's blocks (its "virtual blocks") are
// shared out over the resident grid in a grid-stride loop, and a grid barrier
// separates phases.

enum PhaseKind : int { PH_STAGE = 0, PH_DEMAG = 1, PH_FIELD = 2, PH_ERR = 3, PH_CONTROL = 4 };

// One batched demag pass of a phase: pass number (0 xfwd, 1 yfwd, 2 zpass,
// 3 yinv, 4 ypass2d, 5 xinv), its block count and its block tables.
struct PhasePass {
  int pass;
  int nblocks;
  const void* segs;
  const int* b2s;
};

struct Phase {
  int kind;
  int nvb;             // virtual blocks
  int npass;           // PH_DEMAG: independent passes run side by side
  PhasePass pass[2];
  const float* x;      // PH_DEMAG, PH_FIELD: the stage state
  float* out;          // PH_FIELD: the stage derivative; PH_STAGE: the stage state
  int has_next;        // PH_FIELD: also the next stage state (next)
  BatchNext next;
  StageArgs sa;        // PH_STAGE
  ErrArgs ea;          // PH_ERR
  CtlArgs ca;          // PH_CONTROL
};

struct PersistArgs {
  const Phase* ph;
  int nph;
  float* X;
  float* X0;
  float* K0;
  const float* KL;
  float* HD;
  const uint8_t* G;
  const int* c2m;
  BatchMember* st;
  unsigned* errbits;
  unsigned* tqbits;
  const void* ftab;  // batched field table and block -> member map
  const int* fb2i;
  int cs, nch, nslots;
  // grid barrier: per-block arrival words, the release word, the error word
  // (0, or 1 + the ordinal within the launch of the barrier that timed out,
  // or -1 when a block gave up waiting for the release), and the spin limit
  // (polls) of the barrier's coordinator; spin 0 makes the first barrier
  // fail at once.
  unsigned* flags;
  unsigned* go;
  int* err;
  unsigned spin;
};

}  // namespace bk

// Resident 256-thread blocks of the persistent step kernel per compute unit
// with shm bytes of dynamic LDS, and its cooperative launch (every block of
// `grid` resident at once, or the launch fails); a: the arguments in device
// memory.
int batch_persist_blocks_per_cu(size_t shm);
// nsteps batch steps; the launch's barriers use the generations gen0,
// gen0 + 1, ... (nsteps x a->nph of them).
void batch_persist_launch(const bk::PersistArgs* a, int nsteps, unsigned gen0, int grid, size_t shm, hipStream_t s);

}  // namespace mpa
