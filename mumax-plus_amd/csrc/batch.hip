// Batched stepping of many small independent ferromagnets (see batch.hpp).
//
// One batch step, for a tableau whose last stage is the new state
// (first-same-as-last), with S stages:
//   k_bstage     x_1 = normalise(x0 + dt_m a_10 k_0) per member m, after a
//                member whose previous attempt was accepted takes x0 = x and
//                k_0 = k_(S-1) (the reference swaps these buffers,
//                rungekutta.cpp:25-84; here the take-over is fused into the
//                first combination, so no pointer changes between steps and
//                the step can be replayed from a graph)
//   for s = 1 .. S-1:
//     demag passes and field kernel of every member at x_s -> k_s; for
//                s < S-1 the field kernel also forms
//                x_(s+1) = normalise(x0 + dt_m sum_j<=s a_(s+1)j k_j)
//                (the stage states alternate between two buffers); at
//                s = S-1 it forms the per-member max
//                |dt sum_j (w1_j - w2_j) k_j| (getError,
//                rungekutta.cpp:140-151) and max |k_(S-1)| (the torque at
//                the new state, for relaxation); k_berr does the same from
//                memory when a member is primed
//   k_bcontrol   per-member step controller (rungekutta.cpp:69-74,
//                timesolver.cpp:85-96), accept / reject, time, and the
//                stopping rules (TimeSolver::run, timesolver.cpp:97-111)
#include "batch.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>

#include "batch_kernels.hpp"
#include "world.hpp"

namespace mpa {

namespace {

using namespace bk;

// One wavefront per chunk (64 threads x 4 consecutive cells), four chunks
// per block.
__global__ void __launch_bounds__(256) k_bstage(StageArgs sa, const float* X, float* XO,
                                                float* __restrict__ X0, float* __restrict__ K0,
                                                const float* __restrict__ KL, const uint8_t* __restrict__ G,
                                                const int* __restrict__ c2m, const BatchMember* __restrict__ st,
                                                int cs, int nch) {
  bstage_block(sa, X, XO, X0, K0, KL, G, c2m, st, cs, nch, blockIdx.x);
}

__global__ void __launch_bounds__(CH) k_berr(ErrArgs ea, const uint8_t* __restrict__ G, const int* __restrict__ c2m,
                                             const BatchMember* __restrict__ st, unsigned* __restrict__ errbits,
                                             unsigned* __restrict__ tqbits, int cs) {
  __shared__ float red[2][CH / 64];
  berr_block(ea, G, c2m, st, errbits, tqbits, cs, blockIdx.x, red);
}

// One thread per member slot.
__global__ void k_bcontrol(BatchMember* __restrict__ st, unsigned* __restrict__ errbits,
                           unsigned* __restrict__ tqbits, int nslots, CtlArgs c) {
  bcontrol_member(st, errbits, tqbits, nslots, c, blockIdx.x * blockDim.x + threadIdx.x);
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
  XO_.allocate(n3);
  X0_.allocate(n3);
  HD_.allocate(n3);
  const int S = (int)tb.nodes.size();
  for (int s = 0; s < S; ++s) K_.emplace_back(n3);
  G_.allocate((size_t)cs_);
  for (DeviceArray<float>* a : {&X_, &XO_, &X0_, &HD_})
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
  HIP_CHECK(hipStreamSynchronize(stream));
  Ferromagnet* fm = sl.fm;
  // after a rejected attempt the state is the attempt's start, X0 (X holds
  // the rejected attempt's last stage)
  const float* src = S.accept == 0 ? X0_.get() : X_.get();
  HIP_CHECK(hipMemcpy2DAsync(fm->m.get(), (size_t)fm->N * sizeof(float), src + sl.off,
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
  std::vector<FieldBatchMember> fb, fb0, fb1;
  fbytes0_ = fbytes1_ = 0.0;
  for (size_t id = 0; id < slots_.size(); ++id) {
    const Slot& sl = slots_[id];
    if (!sl.used) continue;
    // the parameters read on admission: a member's result never depends on
    // when other members join or leave
    Ferromagnet* fm = sl.fm;
    const FMDesc& d = sl.desc;
    if (sl.demag) dm.push_back({&fm->demag(), sl.off, PS{nullptr, d.msat.v}, fm->geom_ptr()});
    fb.push_back({d, (int)sl.off, sl.mode == BATCH_RELAX ? MODE_DAMPING : MODE_LLG, (int)id});
    // single layers with demag form the demag batch's second group
    const bool g1 = sl.demag && d.nz == 1;
    (g1 ? fb1 : fb0).push_back(fb.back());
    (g1 ? fbytes1_ : fbytes0_) += 12.0 * sl.ncell * (sl.demag ? 3 : 2);
  }
  demag_ = dm.empty() ? nullptr : std::make_unique<DemagBatch>(dm, cs_);
  field_ = std::make_unique<FieldBatch>(fb, (int)cs_);
  field0_ = std::make_unique<FieldBatch>(fb0, (int)cs_);
  field1_ = std::make_unique<FieldBatch>(fb1, (int)cs_);
  HIP_CHECK(hipMemcpyAsync(c2m_.get(), c2m_host_.data(), c2m_host_.size() * sizeof(int), hipMemcpyHostToDevice,
                           stream));
  HIP_CHECK(hipStreamSynchronize(stream));
  dirty_ = false;
  pdirty_ = true;
}

bool Batch::two_chains() const {
  return demag_ && demag_->has_group(1) && field0_->nblocks() > 0 && field1_->nblocks() > 0;
}

void Batch::eval_rhs(const float* x, float* out, const BatchNext* next, const BatchErr* err) {
  if (demag_) demag_->exec(x, HD_.get(), stream, &prof, stream2, fork_, join_);
  double bytes = 0.0;
  for (const Slot& sl : slots_)
    if (sl.used) bytes += 12.0 * sl.ncell * ((sl.demag ? 3 : 2) + (next ? next->nk + 2 : 0));
  Timed t(&prof, "k_field_uniform_b", bytes, stream);
  if (err) field_->launch_err(x, HD_.get(), out, stream, *err);
  else field_->launch(x, HD_.get(), out, stream, next);
}

BatchErr Batch::step_error() {
  // the error norms of the step, formed in the last stage's field launch
  const Tableau& tb = tableau_by_name(method_);
  const int S = (int)tb.nodes.size();
  BatchErr e{};
  e.nk = S - 1;
  for (int j = 0; j < S; ++j) {
    if (j < S - 1) e.k[j] = K_[j].get();
    e.e[j] = tb.w1[j] - tb.w2[j];
  }
  e.errbits = errbits_.get();
  e.tqbits = tqbits_.get();
  e.st = st_.get();
  return e;
}

float* Batch::stage_state(int s) {
  // the last stage's state is the new state, in X_; the stages alternate
  // between X_ and XO_ backwards from there, so a field launch reads one and
  // writes the next stage state into the other
  const int S = (int)tableau_by_name(method_).nodes.size();
  return ((S - 1 - s) % 2 == 0) ? X_.get() : XO_.get();
}

BatchNext Batch::next_stage(int s) {
  // the state of stage s + 1, formed in the field launch of stage s
  const Tableau& tb = tableau_by_name(method_);
  BatchNext n{};
  n.x0 = X0_.get();
  n.nk = s;
  for (int j = 0; j <= s; ++j) {
    if (j < s) n.k[j] = K_[j].get();
    n.a[j] = tb.A[s + 1][j];
  }
  n.xout = stage_state(s + 1);
  n.st = st_.get();
  return n;
}

void Batch::prime(const std::vector<int>& fresh) {
  if (fresh.empty()) return;
  const Tableau& tb = tableau_by_name(method_);
  const int S = (int)tb.nodes.size();
  // k_(S-1) = rhs(x) for every member: unchanged for members whose last
  // attempt was accepted (x is that attempt's last stage state), unused by
  // the others, and the fresh members' first derivative.
  eval_rhs(X_.get(), K_[S - 1].get(), nullptr, nullptr);
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
  // stage 1: take over the accepted state and form the first stage state;
  // every later stage state is formed in the previous stage's field launch
  {
    StageArgs sa{};
    sa.nk = 1;
    sa.commit = 1;
    sa.k[0] = K_[0].get();
    sa.a[0] = tb.A[1][0];
    Timed t(&prof, "k_bstage", 12.0 * used * 3 + 24.0 * used, stream);
    k_bstage<<<(nch + 3) / 4, 256, 0, stream>>>(sa, X_.get(), stage_state(1), X0_.get(), K_[0].get(),
                                                K_[S - 1].get(), G_.get(), c2m_.get(), st_.get(), (int)cs_, (int)nch);
    checkLaunch("k_bstage");
  }
  const BatchErr err = step_error();
  if (!two_chains()) {
    for (int s = 1; s < S; ++s) {
      if (s + 1 < S) {
        const BatchNext n = next_stage(s);
        eval_rhs(stage_state(s), K_[s].get(), &n, nullptr);
      } else {
        eval_rhs(stage_state(s), K_[s].get(), nullptr, &err);
      }
    }
  } else {
    // The single layers with demag depend on the other members nowhere in
    // the stages (their cells, spectra and stage states are their own), so
    // their whole chain of stages runs on the second stream, forked after
    // the stage combination and joined before the error norms.
    HIP_CHECK(hipEventRecord(fork_, stream));
    HIP_CHECK(hipStreamWaitEvent(stream2, fork_, 0));
    for (int s = 1; s < S; ++s) {
      const bool has_next = s + 1 < S;
      const BatchNext n = has_next ? next_stage(s) : BatchNext{};
      const double nb = has_next ? 12.0 * (n.nk + 2) : 0.0;
      const float* x = stage_state(s);
      for (int g = 0; g < 2; ++g) {
        hipStream_t st = g == 0 ? stream : stream2;
        const FieldBatch& fbat = g == 0 ? *field0_ : *field1_;
        demag_->exec_group(x, HD_.get(), g, st, &prof);
        double cells = 0.0;
        for (const Slot& sl : slots_)
          if (sl.used && ((sl.demag && sl.desc.nz == 1) == (g == 1))) cells += (double)sl.ncell;
        Timed t(&prof, "k_field_uniform_b", (g == 0 ? fbytes0_ : fbytes1_) + nb * cells, st);
        if (has_next) fbat.launch(x, HD_.get(), K_[s].get(), st, &n);
        else fbat.launch_err(x, HD_.get(), K_[s].get(), st, err);
      }
    }
    HIP_CHECK(hipEventRecord(join_, stream2));
    HIP_CHECK(hipStreamWaitEvent(stream, join_, 0));
  }
  CtlArgs c{tb.order1, tb.order2, 0};
  const int ns = (int)slots_.size();
  k_bcontrol<<<(ns + 63) / 64, 64, 0, stream>>>(st_.get(), errbits_.get(), tqbits_.get(), ns, c);
  checkLaunch("k_bcontrol");
}

namespace {

// Autotuner of the batch step path: per size class, every candidate path is
// timed on real steps (wall clock per step over an advance, the advances
// that capture a graph or build a plan left out) until each has
// TUNE_SAMPLES samples; the fastest by its best sample is then used for
// every later advance of that class in the process.
constexpr int TUNE_SAMPLES = 2;
constexpr int PERSIST_MAX_GRID = 4096;
// Candidates: the per-pass launches, and the persistent kernel with as many
// blocks per compute unit as fit, with one per compute unit, and with 64
// blocks (a grid barrier costs more the more blocks take part, a phase's
// work runs faster the more blocks share it).
constexpr int TUNE_CANDIDATES = 4;
const char* const TUNE_NAMES[TUNE_CANDIDATES] = {"passes", "persistent", "persistent_1pcu", "persistent_64"};
const int TUNE_BPC[TUNE_CANDIDATES] = {0, 0, 1, 1};
const int TUNE_CAP[TUNE_CANDIDATES] = {0, 0, 0, 64};

struct TuneEntry {
  int count[TUNE_CANDIDATES] = {0, 0, 0, 0};
  double best[TUNE_CANDIDATES] = {0.0, 0.0, 0.0, 0.0};  // ms per step
  bool usable[TUNE_CANDIDATES] = {true, true, true, true};
  int choice = -1;
};

std::mutex tune_mu;
std::map<std::string, TuneEntry>& tune_table() {
  static std::map<std::string, TuneEntry> t;
  return t;
}

}  // namespace

// The record as text, one size class a line: the class, then tab-separated
// the chosen candidate (-1: undecided) and, per candidate, "usable count
// best_ms".
void batch_tune_save(const std::string& path) {
  std::lock_guard<std::mutex> lk(tune_mu);
  const std::string tmp = path + ".tmp";
  {
    std::ofstream f(tmp);
    if (!f) throw Error(MPA_ERR_RUNTIME, "batch tuning: cannot write " + tmp);
    f.precision(17);
    for (const auto& kv : tune_table()) {
      const TuneEntry& e = kv.second;
      f << kv.first << '\t' << e.choice;
      for (int c = 0; c < TUNE_CANDIDATES; ++c) f << '\t' << (e.usable[c] ? 1 : 0) << ' ' << e.count[c] << ' ' << e.best[c];
      f << '\n';
    }
    if (!f) throw Error(MPA_ERR_RUNTIME, "batch tuning: cannot write " + tmp);
  }
  if (std::rename(tmp.c_str(), path.c_str()) != 0) throw Error(MPA_ERR_RUNTIME, "batch tuning: cannot write " + path);
}

void batch_tune_load(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw Error(MPA_ERR_RUNTIME, "batch tuning: cannot read " + path);
  std::map<std::string, TuneEntry> read;
  std::string line;
  while (std::getline(f, line)) {
    if (line.empty()) continue;
    const size_t tab = line.find('\t');
    if (tab == std::string::npos) throw Error(MPA_ERR_INVALID, "batch tuning: malformed line in " + path);
    std::istringstream in(line.substr(tab + 1));
    TuneEntry e;
    in >> e.choice;
    for (int c = 0; c < TUNE_CANDIDATES; ++c) {
      int u = 0;
      in >> u >> e.count[c] >> e.best[c];
      e.usable[c] = u != 0;
    }
    if (!in || e.choice < -1 || e.choice >= TUNE_CANDIDATES)
      throw Error(MPA_ERR_INVALID, "batch tuning: malformed line in " + path);
    read[line.substr(0, tab)] = e;
  }
  std::lock_guard<std::mutex> lk(tune_mu);
  for (auto& kv : read) tune_table()[kv.first] = kv.second;
}

void batch_tune_clear() {
  std::lock_guard<std::mutex> lk(tune_mu);
  tune_table().clear();
}

std::string Batch::size_class() {
  long long cells = 0;
  int members = 0;
  bool has3d = false;
  int maxline = 0;
  for (const Slot& sl : slots_) {
    if (!sl.used) continue;
    ++members;
    cells += sl.ncell;
    has3d = has3d || sl.desc.nz > 1;
    maxline = std::max({maxline, sl.desc.nx, sl.desc.ny, sl.desc.nz});
  }
  auto lg = [](long long v) {
    int k = 0;
    while ((1LL << (k + 1)) <= v) ++k;
    return k;
  };
  std::ostringstream o;
  o << "cells<2^" << lg(cells) + 1 << " members<2^" << lg(members) + 1 << " line<=2^" << lg(maxline) + (maxline & (maxline - 1) ? 1 : 0)
    << (has3d ? " 3d" : " 2d") << " " << method_;
  return o.str();
}

std::string Batch::tune_report() const {
  std::lock_guard<std::mutex> lk(tune_mu);
  std::ostringstream o;
  for (const auto& kv : tune_table()) {
    const TuneEntry& e = kv.second;
    o << kv.first << " | " << (e.choice >= 0 ? TUNE_NAMES[e.choice] : "undecided");
    for (int c = 0; c < TUNE_CANDIDATES; ++c) {
      o << " | " << TUNE_NAMES[c] << " ";
      if (!e.usable[c]) o << "n/a";
      else if (e.count[c] == 0) o << "-";
      else o << e.best[c] << " ms";
    }
    o << "\n";
  }
  return o.str();
}

bool Batch::persist_supported() {
  if (dirty_) rebuild();
  for (const Slot& sl : slots_) {
    if (!sl.used || !sl.demag) continue;
    for (int a = 0; a < 3; ++a) {
      const int n = a == 0 ? sl.desc.nx : (a == 1 ? sl.desc.ny : sl.desc.nz);
      if (demag_fft_length(n) > 256) return false;
    }
  }
  return nmembers() > 0;
}

int Batch::persist_grid() {
  if (!persist_supported()) return 0;
  if (pdirty_) build_persist();
  return grid_for(persist_blocks_per_cu, persist_max_grid);
}

int Batch::grid_for(int blocks_per_cu, int cap) const {
  const int per = blocks_per_cu > 0 ? std::min(blocks_per_cu, pfit_) : pfit_;
  const int g = std::min({per * pcus_, pmaxnvb_, PERSIST_MAX_GRID});
  return cap > 0 ? std::min(g, cap) : g;
}

void Batch::build_persist() {
  const Tableau& tb = tableau_by_name(method_);
  const int S = (int)tb.nodes.size();
  const int nch = (int)(cs_ / CH);
  const int ns = (int)slots_.size();
  std::vector<bk::Phase> ph;
  // the demag passes as phases: the two groups' chains side by side
  // (members with a z transform: xfwd, yfwd, zpass, yinv, xinv; single
  // layers: xfwd, ypass2d, xinv, whose x inverse shares the z pass's phase)
  bk::Phase dph[5];
  for (auto& d : dph) {
    d = bk::Phase{};
    d.kind = bk::PH_DEMAG;
  }
  size_t shm = 0;
  if (demag_) {
    for (const DemagBatch::PassTable& t : demag_->pass_tables()) {
      static const int slot0[6] = {0, 1, 2, 3, -1, 4};
      static const int slot1[6] = {0, -1, -1, -1, 1, 2};
      const int slot = t.group == 0 ? slot0[t.pass] : slot1[t.pass];
      if (slot < 0) throw Error(MPA_ERR_RUNTIME, "batch: unexpected demag pass");
      bk::Phase& d = dph[slot];
      if (d.npass >= 2) throw Error(MPA_ERR_RUNTIME, "batch: two passes of one group in a phase");
      d.pass[d.npass++] = bk::PhasePass{t.pass, (int)t.nblocks, t.segs, t.b2s};
      d.nvb += (int)t.nblocks;
      shm = std::max(shm, t.shm);
    }
  }
  // as the per-pass step: the first stage state by the combination, every
  // later one by the previous stage's field phase
  bk::Phase st{};
  st.kind = bk::PH_STAGE;
  st.nvb = (nch + 3) / 4;
  st.sa.nk = 1;
  st.sa.commit = 1;
  st.sa.k[0] = K_[0].get();
  st.sa.a[0] = tb.A[1][0];
  st.out = stage_state(1);
  ph.push_back(st);
  for (int s = 1; s < S; ++s) {
    for (bk::Phase d : dph)
      if (d.npass > 0) {
        d.x = stage_state(s);
        ph.push_back(d);
      }
    bk::Phase f{};
    f.kind = bk::PH_FIELD;
    f.nvb = (int)field_->nblocks();
    f.x = stage_state(s);
    f.out = K_[s].get();
    f.has_next = s + 1 < S;
    if (f.has_next) f.next = next_stage(s);
    ph.push_back(f);
  }
  bk::Phase e{};
  e.kind = bk::PH_ERR;
  e.nvb = nch;
  e.ea.nk = S;
  for (int j = 0; j < S; ++j) {
    e.ea.k[j] = K_[j].get();
    e.ea.e[j] = tb.w1[j] - tb.w2[j];
  }
  e.ea.tq = K_[S - 1].get();
  ph.push_back(e);
  bk::Phase c{};
  c.kind = bk::PH_CONTROL;
  c.nvb = (ns + 255) / 256;
  c.ca = bk::CtlArgs{tb.order1, tb.order2, 0};
  ph.push_back(c);
  int maxnvb = 1;
  for (const bk::Phase& p : ph) maxnvb = std::max(maxnvb, p.nvb);
  phases_.allocate(ph.size() * sizeof(bk::Phase));
  HIP_CHECK(hipMemcpy(phases_.get(), ph.data(), ph.size() * sizeof(bk::Phase), hipMemcpyHostToDevice));
  nph_ = (int)ph.size();
  pshm_ = shm;
  // grid: every block resident at once
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  HIP_CHECK(hipDeviceGetAttribute(&pcus_, hipDeviceAttributeMultiprocessorCount, dev));
  pfit_ = batch_persist_blocks_per_cu(shm);
  if (pfit_ < 1) throw Error(MPA_ERR_RUNTIME, "batch: the persistent step kernel does not fit a compute unit");
  pmaxnvb_ = maxnvb;
  if (pflags_.n == 0) {
    pflags_.allocate(PERSIST_MAX_GRID + 1);
    perr_.allocate(1);
    HIP_CHECK(hipMemset(pflags_.get(), 0, pflags_.bytes()));
    HIP_CHECK(hipMemset(perr_.get(), 0, perr_.bytes()));
    pgen_ = 1;
  }
  bk::PersistArgs a{};
  a.ph = reinterpret_cast<const bk::Phase*>(phases_.get());
  a.nph = nph_;
  a.X = X_.get();
  a.X0 = X0_.get();
  a.K0 = K_[0].get();
  a.KL = K_[S - 1].get();
  a.HD = HD_.get();
  a.G = G_.get();
  a.c2m = c2m_.get();
  a.st = st_.get();
  a.errbits = errbits_.get();
  a.tqbits = tqbits_.get();
  a.ftab = field_->table();
  a.fb2i = field_->block_members();
  a.cs = (int)cs_;
  a.nch = nch;
  a.nslots = ns;
  a.flags = pflags_.get();
  a.go = pflags_.get() + PERSIST_MAX_GRID;
  a.err = perr_.get();
  a.spin = persist_spin;
  pargs_.allocate(sizeof(a));
  HIP_CHECK(hipMemcpy(pargs_.get(), &a, sizeof(a), hipMemcpyHostToDevice));
  pdirty_ = false;
}

void Batch::persist_recover(int err) {
  // the barrier words back to their initial state
  HIP_CHECK(hipMemsetAsync(pflags_.get(), 0, pflags_.bytes(), stream));
  HIP_CHECK(hipMemsetAsync(perr_.get(), 0, perr_.bytes(), stream));
  HIP_CHECK(hipStreamSynchronize(stream));
  pgen_ = 1;
  if (err < 0) {
    broken_ = true;
    throw Error(MPA_ERR_RUNTIME, "batch: a block of the persistent step kernel gave up at a grid barrier; "
                                 "the members' state is lost");
  }
  // Barrier `ordinal` failed: every block finished that phase and none
  // started the next.  After the controller phase (the last of a step) the
  // state is that of a completed step; after any other phase the step start
  // is intact in X0 and K0 (the first phase takes it over), so the step is
  // retried from there: accept = 0, error norms cleared.
  const int phase = (err - 1) % nph_;
  if (phase != nph_ - 1) {
    std::vector<BatchMember> host(slots_.size());
    HIP_CHECK(hipMemcpy(host.data(), st_.get(), host.size() * sizeof(BatchMember), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < host.size(); ++i)
      if (slots_[i].used && host[i].active && host[i].accept == 1) host[i].accept = 0;
    HIP_CHECK(hipMemcpy(st_.get(), host.data(), host.size() * sizeof(BatchMember), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemset(errbits_.get(), 0, errbits_.bytes()));
    HIP_CHECK(hipMemset(tqbits_.get(), 0, tqbits_.bytes()));
  }
  throw Error(MPA_ERR_RUNTIME, "batch: a grid barrier of the persistent step kernel timed out at phase " +
                                   std::to_string(phase) + "; the batch is back at the start of that step");
}

void Batch::advance_persistent(long long n, int blocks_per_cu, int cap) {
  if (pdirty_) build_persist();
  const int grid = grid_for(blocks_per_cu, cap);
  const bk::PersistArgs* pa = reinterpret_cast<const bk::PersistArgs*>(pargs_.get());
  while (n > 0) {
    const int k = (int)std::min<long long>(n, 1 << 16);
    const unsigned long long need = (unsigned long long)k * (unsigned)nph_;
    if (pgen_ + need >= 0xFFFFFF00ull) {
      HIP_CHECK(hipMemsetAsync(pflags_.get(), 0, pflags_.bytes(), stream));
      pgen_ = 1;
    }
    batch_persist_launch(pa, k, pgen_, grid, pshm_, stream);
    pgen_ += (unsigned)need;
    int err = 0;
    HIP_CHECK(hipMemcpyAsync(&err, perr_.get(), sizeof(int), hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    if (err != 0) persist_recover(err);
    n -= k;
  }
}

void Batch::advance_passes(long long n) {
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

void Batch::advance(long long n) {
  if (n <= 0) return;
  if (broken_) throw Error(MPA_ERR_RUNTIME, "batch: the members' state was lost at a failed grid barrier");
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
  // the per-kernel profile times the per-pass launches
  int mode = prof.enabled ? STEP_PASSES : step_mode;
  if (mode == STEP_PERSISTENT && !persist_supported())
    throw Error(MPA_ERR_INVALID, "batch: the persistent step kernel takes FFT lines of up to 256 points");
  if (mode == STEP_PERSISTENT) {
    last_path_ = 1;
    advance_persistent(n, persist_blocks_per_cu, persist_max_grid);
    return;
  }
  if (mode == STEP_PASSES) {
    last_path_ = 0;
    advance_passes(n);
    return;
  }
  // STEP_AUTO
  const std::string key = size_class();
  int cand = -1;
  const bool ps = persist_supported();
  if (ps && pdirty_) build_persist();
  // one block per compute unit is a candidate of its own only when it
  // gives a smaller grid
  bool usable[TUNE_CANDIDATES] = {true, ps, ps && grid_for(1, 0) < grid_for(0, 0), ps && 64 < grid_for(1, 0)};
  {
    std::lock_guard<std::mutex> lk(tune_mu);
    TuneEntry& e = tune_table()[key];
    for (int c = 0; c < TUNE_CANDIDATES; ++c) e.usable[c] = e.usable[c] && usable[c];
    if (e.choice >= 0 && e.usable[e.choice]) {
      cand = e.choice;
    } else {
      e.choice = -1;
      for (int c = 0; c < TUNE_CANDIDATES; ++c)
        if (e.usable[c] && (cand < 0 || e.count[c] < e.count[cand])) cand = c;
    }
  }
  HIP_CHECK(hipStreamSynchronize(stream));
  const bool cold = cand == 0 && use_graph && !exec_;
  const auto t0 = std::chrono::steady_clock::now();
  last_path_ = cand == 0 ? 0 : 1;
  if (cand == 0) advance_passes(n);
  else advance_persistent(n, TUNE_BPC[cand], TUNE_CAP[cand]);
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / (double)n;
  if (cold) return;
  std::lock_guard<std::mutex> lk(tune_mu);
  TuneEntry& e = tune_table()[key];
  if (e.choice >= 0) return;
  e.best[cand] = e.count[cand] == 0 ? ms : std::min(e.best[cand], ms);
  ++e.count[cand];
  bool done = true;
  int best = -1;
  for (int c = 0; c < TUNE_CANDIDATES; ++c) {
    if (!e.usable[c]) continue;
    if (e.count[c] < TUNE_SAMPLES) done = false;
    else if (best < 0 || e.best[c] < e.best[best]) best = c;
  }
  if (done && best >= 0) e.choice = best;
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
  // stage state 1 by the combination (x0 and k0 read, the state written;
  // the take-over copies x and k into x0 and k0), stage state q >= 2 in the
  // field launch of stage q - 1 (x0 and k_0 .. k_(q-2) read, the state
  // written; k_(q-1) is in registers)
  double b = 12.0 * used * 3 + 24.0 * used;
  for (int q = 2; q < S; ++q) b += 12.0 * used * (q + 1);
  for (const Slot& sl : slots_)
    if (sl.used) b += (S - 1) * 12.0 * sl.ncell * (sl.demag ? 3 : 2);  // field kernel: m, h_demag, k
  if (demag_) b += (S - 1) * demag_->algorithmic_bytes();
  return b;
}

int Batch::launches_per_step() {
  if (dirty_) rebuild();
  const Tableau& tb = tableau_by_name(method_);
  const int S = (int)tb.nodes.size();
  return (S - 1) * (1 + (demag_ ? demag_->nlaunches() : 0)) + 2;
}

}  // namespace mpa
