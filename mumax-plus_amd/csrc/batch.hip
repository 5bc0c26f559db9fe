// Batched stepping of many small independent ferromagnets (see batch.hpp).
//
// One batch step, for a tableau whose last stage is the new state
// (first-same-as-last), with S stages:
//   for s = 1 .. S-1:
//     k_bstage   x = normalise(x0 + dt_m sum_j<s a_sj k_j) per member m; at
//                s = 1 a member whose previous attempt was accepted first
//                takes x0 = x and k_0 = k_(S-1) (the reference swaps these
//                buffers, rungekutta.cpp:25-84; here the take-over is fused
//                into the first combination, so no pointer changes between
//                steps and the step can be replayed from a graph)
//     demag passes and field kernel of every member -> k_s
//   k_berr       per-member max |dt sum_j (w1_j - w2_j) k_j| (getError,
//                rungekutta.cpp:140-151) and max |k_(S-1)| (the torque at
//                the new state, for relaxation)
//   k_bcontrol   per-member step controller (rungekutta.cpp:69-74,
//                timesolver.cpp:85-96), accept / reject, time, and the
//                stopping rules (TimeSolver::run, timesolver.cpp:97-111)
#include "batch.hpp"

#include <algorithm>
#include <cmath>

#include "world.hpp"

namespace mpa {

namespace {

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
// uniform over the wavefront.
__global__ void __launch_bounds__(256) k_bstage(StageArgs sa, float* __restrict__ X, float* __restrict__ X0,
                                                float* __restrict__ K0, const float* __restrict__ KL,
                                                const uint8_t* __restrict__ G, const int* __restrict__ c2m,
                                                const BatchMember* __restrict__ st, int cs, int nch) {
  const int chunk = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (chunk >= nch) return;
  const int m = c2m[chunk];
  if (m < 0) return;
  const BatchMember& S = st[m];
  if (!S.active) return;
  const int q = chunk * (CH / 4) + (threadIdx.x & 63);  // float4 index within a component
  const int cs4 = cs >> 2;
  const bool take = sa.commit && S.accept;
  float c[MAX_STAGES];
#pragma unroll
  for (int j = 0; j < MAX_STAGES; ++j) c[j] = j < sa.nk ? (float)(S.dt * sa.a[j]) : 0.f;
  float4* X4 = reinterpret_cast<float4*>(X);
  float4* X04 = reinterpret_cast<float4*>(X0);
  float4* K04 = reinterpret_cast<float4*>(K0);
  const float4* KL4 = reinterpret_cast<const float4*>(KL);
  float4 v[3];
#pragma unroll
  for (int comp = 0; comp < 3; ++comp) {
    const size_t i = (size_t)comp * cs4 + q;
    float4 x0, k0;
    if (take) {
      x0 = X4[i];
      k0 = KL4[i];
      X04[i] = x0;
      K04[i] = k0;
    } else {
      x0 = X04[i];
      k0 = K04[i];
    }
    float4 a = x0;
    axpy4(a, c[0], k0);
    for (int j = 1; j < sa.nk; ++j) axpy4(a, c[j], reinterpret_cast<const float4*>(sa.k[j])[i]);
    v[comp] = a;
  }
  const float i0 = rsqrtf(v[0].x * v[0].x + v[1].x * v[1].x + v[2].x * v[2].x);
  const float i1 = rsqrtf(v[0].y * v[0].y + v[1].y * v[1].y + v[2].y * v[2].y);
  const float i2 = rsqrtf(v[0].z * v[0].z + v[1].z * v[1].z + v[2].z * v[2].z);
  const float i3 = rsqrtf(v[0].w * v[0].w + v[1].w * v[1].w + v[2].w * v[2].w);
  const uchar4 g = reinterpret_cast<const uchar4*>(G)[q];
#pragma unroll
  for (int comp = 0; comp < 3; ++comp) {
    float4 o;
    o.x = g.x ? v[comp].x * i0 : 0.f;
    o.y = g.y ? v[comp].y * i1 : 0.f;
    o.z = g.z ? v[comp].z * i2 : 0.f;
    o.w = g.w ? v[comp].w * i3 : 0.f;
    X4[(size_t)comp * cs4 + q] = o;
  }
}

__device__ __forceinline__ float block_max(float v, float* red) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) red[w] = v;
  __syncthreads();
  float r = 0.f;
  if (threadIdx.x == 0)
    for (int i = 0; i < CH / 64; ++i) r = fmaxf(r, red[i]);
  return r;
}

__global__ void __launch_bounds__(CH) k_berr(ErrArgs ea, const uint8_t* __restrict__ G, const int* __restrict__ c2m,
                                             const BatchMember* __restrict__ st, unsigned* __restrict__ errbits,
                                             unsigned* __restrict__ tqbits, int cs) {
  __shared__ float red[2][CH / 64];
  const int chunk = blockIdx.x;
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
  const float be = block_max(err, red[0]);
  const float bt = block_max(tq, red[1]);
  if (threadIdx.x == 0) {
    if (want_err) atomicMax(errbits + m, __float_as_uint(be));
    if (want_tq) atomicMax(tqbits + m, __float_as_uint(bt));
  }
}

// One thread per member slot.
__global__ void k_bcontrol(BatchMember* __restrict__ st, unsigned* __restrict__ errbits,
                           unsigned* __restrict__ tqbits, int nslots, CtlArgs c) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
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

long long round_chunk(long long n) { return (n + CH - 1) / CH * CH; }

}  // namespace

Batch::Batch(long long capacity, const std::string& method) : method_(method) {
  const Tableau& tb = tableau_by_name(method);
  if (!tb.fsal)
    throw Error(MPA_ERR_INVALID, "batch: the method must be DormandPrince or BogackiShampine (last stage = new state)");
  if (capacity < 1) throw Error(MPA_ERR_INVALID, "batch: capacity must be positive");
  cs_ = round_chunk(capacity);
  if (3 * cs_ >= (1LL << 31)) throw Error(MPA_ERR_INVALID, "batch: capacity above 715 million cells");
  HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  HIP_CHECK(hipStreamCreateWithFlags(&stream2, hipStreamNonBlocking));
  HIP_CHECK(hipEventCreateWithFlags(&fork_, hipEventDisableTiming));
  HIP_CHECK(hipEventCreateWithFlags(&join_, hipEventDisableTiming));
  const size_t n3 = (size_t)3 * cs_;
  X_.allocate(n3);
  X0_.allocate(n3);
  HD_.allocate(n3);
  const int S = (int)tb.nodes.size();
  for (int s = 0; s < S; ++s) K_.emplace_back(n3);
  G_.allocate((size_t)cs_);
  for (DeviceArray<float>* a : {&X_, &X0_, &HD_})
    HIP_CHECK(hipMemsetAsync(a->get(), 0, a->bytes(), stream));
  for (auto& k : K_) HIP_CHECK(hipMemsetAsync(k.get(), 0, k.bytes(), stream));
  HIP_CHECK(hipMemsetAsync(G_.get(), 0, G_.bytes(), stream));
  c2m_host_.assign((size_t)(cs_ / CH), -1);
  c2m_.allocate(c2m_host_.size());
  free_.push_back({0, cs_});
  HIP_CHECK(hipStreamSynchronize(stream));
}

Batch::~Batch() {
  drop_graph();
  if (fork_) (void)hipEventDestroy(fork_);
  if (join_) (void)hipEventDestroy(join_);
  if (stream2) (void)hipStreamDestroy(stream2);
  if (stream) (void)hipStreamDestroy(stream);
}

void Batch::drop_graph() {
  if (exec_) (void)hipGraphExecDestroy(exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  exec_ = nullptr;
  graph_ = nullptr;
}

long long Batch::alloc_range(long long span) {
  for (size_t i = 0; i < free_.size(); ++i) {
    if (free_[i].second < span) continue;
    const long long off = free_[i].first;
    free_[i].first += span;
    free_[i].second -= span;
    if (free_[i].second == 0) free_.erase(free_.begin() + (long)i);
    return off;
  }
  return -1;
}

void Batch::free_range(long long off, long long span) {
  free_.push_back({off, span});
  std::sort(free_.begin(), free_.end());
  std::vector<std::pair<long long, long long>> merged;
  for (auto& r : free_) {
    if (!merged.empty() && merged.back().first + merged.back().second == r.first) merged.back().second += r.second;
    else merged.push_back(r);
  }
  free_.swap(merged);
}

long long Batch::free_cells() const {
  long long n = 0;
  for (auto& r : free_) n += r.second;
  return n;
}

int Batch::nmembers() const {
  int n = 0;
  for (auto& s : slots_) n += s.used ? 1 : 0;
  return n;
}

int Batch::add(Ferromagnet* fm, int mode_flags, double arg) {
  if (!fm) throw Error(MPA_ERR_INVALID, "batch: magnet must not be NULL");
  const int mode = mode_flags & 0xff;
  const bool keep_step = (mode_flags & BATCH_KEEP_STEP) != 0;
  if ((mode_flags & ~(0xff | BATCH_KEEP_STEP)) != 0) throw Error(MPA_ERR_INVALID, "batch: unknown mode flags");
  if (fm->host) throw Error(MPA_ERR_INVALID, "batch: sublattices are not batched");
  World* w = fm->world;
  if (w->magnets.size() != 1 || !w->afms.empty())
    throw Error(MPA_ERR_INVALID, "batch: a member's world must hold only that magnet");
  if (fm->el.enabled) throw Error(MPA_ERR_INVALID, "batch: magnets with elastodynamics are not batched");
  if (mode < BATCH_STEPS || mode > BATCH_RELAX) throw Error(MPA_ERR_INVALID, "batch: unknown mode");
  for (auto& s : slots_)
    if (s.used && s.fm == fm) throw Error(MPA_ERR_INVALID, "batch: magnet is already a member");
  // a member's parameters are read once, on admission
  for (const auto& kv : fm->params)
    if (kv.second.is_dynamic())
      throw Error(MPA_ERR_INVALID, "batch: parameter '" + kv.first + "' has time-dependent terms");
  const unsigned terms = fm->active_terms();
  const FMDesc d = fm->desc(terms);
  if (!field_batchable(d)) throw Error(MPA_ERR_INVALID, "batch: every material parameter must be uniform");
  const TimeSolver& ts = w->solver;
  if (mode == BATCH_RUN && !(arg > 0)) throw Error(MPA_ERR_INVALID, "batch: duration must be positive");
  if (!ts.adaptive && !(ts.dt > 0)) throw Error(MPA_ERR_INVALID, "batch: a fixed step must be positive");
  const long long span = round_chunk(fm->N);
  const long long off = alloc_range(span);
  if (off < 0) return -1;
  int id = -1;
  for (size_t i = 0; i < slots_.size(); ++i)
    if (!slots_[i].used) {
      id = (int)i;
      break;
    }
  if (id < 0) {
    id = (int)slots_.size();
    slots_.emplace_back();
  }
  Slot& sl = slots_[id];
  sl.fm = fm;
  sl.off = off;
  sl.ncell = fm->N;
  sl.span = span;
  sl.mode = mode;
  sl.demag = (terms & TERM_DEMAG) != 0;
  sl.desc = d;
  sl.used = true;
  HIP_CHECK(hipStreamSynchronize(w->stream));
  // state and mask into the packed arrays (the span's tail stays masked out)
  HIP_CHECK(hipMemcpy2DAsync(X_.get() + off, (size_t)cs_ * sizeof(float), fm->m.get(), (size_t)fm->N * sizeof(float),
                             (size_t)fm->N * sizeof(float), 3, hipMemcpyDeviceToDevice, stream));
  HIP_CHECK(hipMemsetAsync(G_.get() + off, 0, (size_t)span, stream));
  if (fm->geom_ptr())
    HIP_CHECK(hipMemcpyAsync(G_.get() + off, fm->geom_ptr(), (size_t)fm->N, hipMemcpyDeviceToDevice, stream));
  else
    HIP_CHECK(hipMemsetAsync(G_.get() + off, 1, (size_t)fm->N, stream));
  for (long long c = off / CH; c < (off + span) / CH; ++c) c2m_host_[(size_t)c] = id;
  // member state
  if ((int)st_.n < (int)slots_.size()) {
    DeviceArray<BatchMember> ns(slots_.size() + 64);
    DeviceArray<unsigned> ne(ns.n), nt(ns.n);
    HIP_CHECK(hipMemsetAsync(ns.get(), 0, ns.bytes(), stream));
    HIP_CHECK(hipMemsetAsync(ne.get(), 0, ne.bytes(), stream));
    HIP_CHECK(hipMemsetAsync(nt.get(), 0, nt.bytes(), stream));
    if (st_.n) {
      HIP_CHECK(hipMemcpyAsync(ns.get(), st_.get(), st_.bytes(), hipMemcpyDeviceToDevice, stream));
      HIP_CHECK(hipMemcpyAsync(ne.get(), errbits_.get(), errbits_.bytes(), hipMemcpyDeviceToDevice, stream));
      HIP_CHECK(hipMemcpyAsync(nt.get(), tqbits_.get(), tqbits_.bytes(), hipMemcpyDeviceToDevice, stream));
    }
    HIP_CHECK(hipStreamSynchronize(stream));
    st_ = std::move(ns);
    errbits_ = std::move(ne);
    tqbits_ = std::move(nt);
  }
  BatchMember S{};
  S.time = ts.time;
  S.dt = ts.dt;
  S.max_error = ts.max_error;
  S.headroom = ts.headroom;
  S.lower = ts.lower_bound;
  S.upper = ts.upper_bound;
  S.sfactor = ts.sensible_factor;
  S.sdefault = ts.sensible_default;
  S.adaptive = ts.adaptive ? 1 : 0;
  S.keep_step = keep_step ? 1 : 0;
  if (keep_step && S.adaptive && !(S.dt > 0)) throw Error(MPA_ERR_INVALID, "batch: no step to keep");
  S.mode = mode;
  if (mode == BATCH_STEPS) S.target = (long long)arg;
  else if (mode == BATCH_RUN) S.stop = ts.time + arg;
  else {
    S.threshold = (float)arg;
    S.target = BATCH_RELAX_MAX_STEPS;
  }
  S.active = 1;
  S.accept = 2;  // waiting for its first right-hand side
  HIP_CHECK(hipMemcpyAsync(st_.get() + id, &S, sizeof(S), hipMemcpyHostToDevice, stream));
  HIP_CHECK(hipStreamSynchronize(stream));
  dirty_ = true;
  return id;
}

void Batch::remove(int id) {
  if (id < 0 || id >= (int)slots_.size() || !slots_[id].used) throw Error(MPA_ERR_INDEX, "batch: no such member");
  Slot& sl = slots_[id];
  BatchMember S{};
  HIP_CHECK(hipMemcpyAsync(&S, st_.get() + id, sizeof(S), hipMemcpyDeviceToHost, stream));
  Ferromagnet* fm = sl.fm;
  HIP_CHECK(hipMemcpy2DAsync(fm->m.get(), (size_t)fm->N * sizeof(float), X_.get() + sl.off,
                             (size_t)cs_ * sizeof(float), (size_t)fm->N * sizeof(float), 3, hipMemcpyDeviceToDevice,
                             stream));
  BatchMember Z{};
  HIP_CHECK(hipMemcpyAsync(st_.get() + id, &Z, sizeof(Z), hipMemcpyHostToDevice, stream));
  HIP_CHECK(hipMemsetAsync(G_.get() + sl.off, 0, (size_t)sl.span, stream));
  HIP_CHECK(hipStreamSynchronize(stream));
  TimeSolver& ts = fm->world->solver;
  ts.time = S.time;
  // a member taken off while the shortened final step of a run is pending
  // gets the step it had before that piece, not the piece
  ts.dt = (S.last && !S.done) ? S.dt_saved : S.dt;
  ts.nsteps += S.steps;
  ts.nrejected += S.rejected;
  fm->world->touch();
  for (long long c = sl.off / CH; c < (sl.off + sl.span) / CH; ++c) c2m_host_[(size_t)c] = -1;
  free_range(sl.off, sl.span);
  sl = Slot{};
  dirty_ = true;
}

void Batch::rebuild() {
  drop_graph();
  std::vector<DemagBatch::Member> dm;
  std::vector<FieldBatchMember> fb;
  for (const Slot& sl : slots_) {
    if (!sl.used) continue;
    // the parameters read on admission: a member's result never depends on
    // when other members join or leave
    Ferromagnet* fm = sl.fm;
    const FMDesc& d = sl.desc;
    if (sl.demag) dm.push_back({&fm->demag(), sl.off, PS{nullptr, d.msat.v}, fm->geom_ptr()});
    fb.push_back({d, (int)sl.off, sl.mode == BATCH_RELAX ? MODE_DAMPING : MODE_LLG});
  }
  demag_ = dm.empty() ? nullptr : std::make_unique<DemagBatch>(dm, cs_);
  field_ = std::make_unique<FieldBatch>(fb, (int)cs_);
  HIP_CHECK(hipMemcpyAsync(c2m_.get(), c2m_host_.data(), c2m_host_.size() * sizeof(int), hipMemcpyHostToDevice,
                           stream));
  HIP_CHECK(hipStreamSynchronize(stream));
  dirty_ = false;
}

void Batch::eval_rhs(float* out) {
  if (demag_) demag_->exec(X_.get(), HD_.get(), stream, &prof, stream2, fork_, join_);
  double bytes = 0.0;
  for (const Slot& sl : slots_)
    if (sl.used) bytes += 12.0 * sl.ncell * (sl.demag ? 3 : 2);
  Timed t(&prof, "k_field_uniform_b", bytes, stream);
  field_->launch(X_.get(), HD_.get(), out, stream);
}

void Batch::prime(const std::vector<int>& fresh) {
  if (fresh.empty()) return;
  const Tableau& tb = tableau_by_name(method_);
  const int S = (int)tb.nodes.size();
  // k_(S-1) = rhs(x) for every member: unchanged for members whose last
  // attempt was accepted (x is that attempt's last stage state), unused by
  // the others, and the fresh members' first derivative.
  eval_rhs(K_[S - 1].get());
  ErrArgs ea{};
  ea.nk = 0;
  ea.tq = K_[S - 1].get();
  const unsigned nch = (unsigned)(cs_ / CH);
  k_berr<<<nch, CH, 0, stream>>>(ea, G_.get(), c2m_.get(), st_.get(), errbits_.get(), tqbits_.get(), (int)cs_);
  checkLaunch("k_berr");
  CtlArgs c{tb.order1, tb.order2, 1};
  const int ns = (int)slots_.size();
  k_bcontrol<<<(ns + 63) / 64, 64, 0, stream>>>(st_.get(), errbits_.get(), tqbits_.get(), ns, c);
  checkLaunch("k_bcontrol");
}

void Batch::record_step() {
  const Tableau& tb = tableau_by_name(method_);
  const int S = (int)tb.nodes.size();
  const unsigned nch = (unsigned)(cs_ / CH);
  double used = 0.0;
  for (const Slot& sl : slots_)
    if (sl.used) used += (double)sl.ncell;
  for (int s = 1; s < S; ++s) {
    StageArgs sa{};
    sa.nk = s;
    sa.commit = s == 1;
    for (int j = 0; j < s; ++j) {
      sa.k[j] = K_[j].get();
      sa.a[j] = tb.A[s][j];
    }
    {
      Timed t(&prof, "k_bstage", 12.0 * used * (2 + s) + (s == 1 ? 24.0 * used : 0.0), stream);
      k_bstage<<<(nch + 3) / 4, 256, 0, stream>>>(sa, X_.get(), X0_.get(), K_[0].get(), K_[S - 1].get(), G_.get(),
                                                  c2m_.get(), st_.get(), (int)cs_, (int)nch);
      checkLaunch("k_bstage");
    }
    eval_rhs(K_[s].get());
  }
  ErrArgs ea{};
  ea.nk = S;
  for (int j = 0; j < S; ++j) {
    ea.k[j] = K_[j].get();
    ea.e[j] = tb.w1[j] - tb.w2[j];
  }
  ea.tq = K_[S - 1].get();
  {
    Timed t(&prof, "k_berr", 0.0, stream);
    k_berr<<<nch, CH, 0, stream>>>(ea, G_.get(), c2m_.get(), st_.get(), errbits_.get(), tqbits_.get(), (int)cs_);
    checkLaunch("k_berr");
  }
  CtlArgs c{tb.order1, tb.order2, 0};
  const int ns = (int)slots_.size();
  k_bcontrol<<<(ns + 63) / 64, 64, 0, stream>>>(st_.get(), errbits_.get(), tqbits_.get(), ns, c);
  checkLaunch("k_bcontrol");
}

void Batch::advance(long long n) {
  if (n <= 0) return;
  std::vector<int> fresh;
  if (dirty_) {
    rebuild();
    std::vector<BatchMember> host(slots_.size());
    HIP_CHECK(hipMemcpy(host.data(), st_.get(), host.size() * sizeof(BatchMember), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < slots_.size(); ++i)
      if (slots_[i].used && host[i].active && host[i].accept == 2) fresh.push_back((int)i);
  }
  prime(fresh);
  if (nmembers() == 0) return;
  const bool graph = use_graph && !prof.enabled;
  if (graph && !exec_) {
    HIP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    try {
      record_step();
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(stream, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HIP_CHECK(hipStreamEndCapture(stream, &graph_));
    HIP_CHECK(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
  }
  for (long long i = 0; i < n; ++i) {
    if (graph) HIP_CHECK(hipGraphLaunch(exec_, stream));
    else record_step();
  }
  HIP_CHECK(hipStreamSynchronize(stream));
  prof.flush_sync(stream);
}

std::vector<int> Batch::finished() {
  std::vector<BatchMember> host(slots_.size());
  if (!host.empty())
    HIP_CHECK(hipMemcpyAsync(host.data(), st_.get(), host.size() * sizeof(BatchMember), hipMemcpyDeviceToHost, stream));
  HIP_CHECK(hipStreamSynchronize(stream));
  std::vector<int> out;
  for (size_t i = 0; i < slots_.size(); ++i)
    if (slots_[i].used && host[i].done) out.push_back((int)i);
  return out;
}

const BatchMember& Batch::member_state(int id) {
  static thread_local BatchMember S;
  if (id < 0 || id >= (int)slots_.size() || !slots_[id].used) throw Error(MPA_ERR_INDEX, "batch: no such member");
  HIP_CHECK(hipMemcpyAsync(&S, st_.get() + id, sizeof(S), hipMemcpyDeviceToHost, stream));
  HIP_CHECK(hipStreamSynchronize(stream));
  return S;
}

double Batch::step_bytes() {
  if (dirty_) rebuild();
  const Tableau& tb = tableau_by_name(method_);
  const int S = (int)tb.nodes.size();
  double used = 0.0;
  for (const Slot& sl : slots_)
    if (sl.used) used += (double)sl.ncell;
  double b = 0.0;
  for (int s = 1; s < S; ++s) b += 12.0 * used * (2 + s);
  b += 24.0 * used;                            // the take-over at the first stage
  for (const Slot& sl : slots_)
    if (sl.used) b += (S - 1) * 12.0 * sl.ncell * (sl.demag ? 3 : 2);  // field kernel: m, h_demag, k
  if (demag_) b += (S - 1) * demag_->algorithmic_bytes();
  return b;
}

int Batch::launches_per_step() {
  if (dirty_) rebuild();
  const Tableau& tb = tableau_by_name(method_);
  const int S = (int)tb.nodes.size();
  return (S - 1) * (2 + (demag_ ? demag_->nlaunches() : 0)) + 2;
}

}  // namespace mpa
