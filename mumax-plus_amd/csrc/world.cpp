// Host object model and the Runge-Kutta time stepper.
#include "world.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace mpa {

// ------------------------------------------------------------ tableaus
// Coefficients of src/core/butchertableau.cpp:89-224 (published rationals).
namespace {
Tableau make_tab(std::vector<double> c, std::vector<std::vector<double>> A, std::vector<double> w1,
                 std::vector<double> w2, int o1, int o2) {
  Tableau t{c, A, w1, w2, o1, o2, false};
  const int S = (int)c.size();
  bool fsal = S > 1 && c[S - 1] == 1.0 && w1[S - 1] == 0.0;
  for (int j = 0; fsal && j < S - 1; ++j) fsal = A[S - 1][j] == w1[j];
  t.fsal = fsal;
  return t;
}

const std::map<std::string, Tableau>& tableaus() {
  static const std::map<std::string, Tableau> tabs = {
      {"Heun", make_tab({0., 1.}, {{}, {1.}}, {1. / 2., 1. / 2.}, {1., 0.}, 2, 1)},
      {"BogackiShampine",
       make_tab({0., 1. / 2., 3. / 4., 1.}, {{}, {1. / 2.}, {0., 3. / 4.}, {2. / 9., 1. / 3, 4. / 9.}},
                {2. / 9., 1. / 3., 4. / 9., 0.}, {7. / 24, 1. / 4., 1. / 3., 1. / 8.}, 3, 2)},
      {"Fehlberg",
       make_tab({0., 1. / 4., 3. / 8., 12. / 13., 1., 1. / 2.},
                {{},
                 {1. / 4.},
                 {3. / 32., 9. / 32.},
                 {1932. / 2197., -7200. / 2197., 7296. / 2197.},
                 {439. / 216., -8., 3680. / 513., -845. / 4104.},
                 {-8. / 27., 2., -3544. / 2565., 1859. / 4104., -11. / 40.}},
                {16. / 135., 0., 6656. / 12825., 28561. / 56430., -9. / 50., 2. / 55.},
                {25. / 216., 0., 1408. / 2565., 2197. / 4104., -1 / 5., 0.}, 5, 4)},
      {"CashKarp",
       make_tab({0., 1. / 5., 3. / 10., 3. / 5., 1., 7. / 8.},
                {{},
                 {1. / 5.},
                 {3. / 40., 9. / 40.},
                 {3. / 10., -9. / 10., 6. / 5.},
                 {-11. / 54., 5. / 2., -70. / 27., 35. / 27.},
                 {1631. / 55296., 175. / 512., 575. / 13824., 44275. / 110592., 253. / 4096.}},
                {37. / 378., 0., 250. / 621., 125. / 594., 0., 512. / 1771.},
                {2825. / 27648., 0, 18575. / 48384., 13525. / 55296., 277. / 14336., 1. / 4.}, 5, 4)},
      {"DormandPrince",
       make_tab({0., 1. / 5., 3. / 10., 4. / 5., 8. / 9., 1., 1.},
                {{},
                 {1. / 5.},
                 {3. / 40., 9. / 40.},
                 {44. / 45., -56. / 15., 32. / 9.},
                 {19372. / 6561., -25360. / 2187., 64448. / 6561., -212. / 729.},
                 {9017. / 3168., -355. / 33., 46732. / 5247., 49. / 176., -5103. / 18656.},
                 {35. / 384., 0., 500. / 1113., 125. / 192., -2187. / 6784., 11. / 84.}},
                {35. / 384., 0., 500. / 1113., 125. / 192., -2187. / 6784., 11. / 84., 0.},
                {5179. / 57600., 0., 7571. / 16695., 393. / 640., -92097. / 339200., 187. / 2100., 1. / 40.}, 5,
                4)},
  };
  return tabs;
}
}  // namespace

const Tableau& tableau_by_name(const std::string& name) {
  auto it = tableaus().find(name);
  if (it == tableaus().end()) throw Error(MPA_ERR_INVALID, "'" + name + "' is not a valid Runge Kutta method name");
  return it->second;
}

// ------------------------------------------------------------ Ferromagnet

namespace {
struct ParamSpec {
  const char* key;
  int ncomp;
  float def[3];
  const char* unit;
};
// Defaults of the reference Ferromagnet constructor (ferromagnet.cpp:24-61).
const ParamSpec kParams[] = {
    {"msat", 1, {1.f, 0, 0}, "A/m"},          {"aex", 1, {0, 0, 0}, "J/m"},
    {"alpha", 1, {0, 0, 0}, ""},              {"ku1", 1, {0, 0, 0}, "J/m3"},
    {"ku2", 1, {0, 0, 0}, "J/m3"},            {"kc1", 1, {0, 0, 0}, "J/m3"},
    {"kc2", 1, {0, 0, 0}, "J/m3"},            {"kc3", 1, {0, 0, 0}, "J/m3"},
    {"frozen_spins", 1, {0, 0, 0}, ""},       {"anisU", 3, {0, 0, 0}, ""},
    {"anisC1", 3, {0, 0, 0}, ""},             {"anisC2", 3, {0, 0, 0}, ""},
    {"bias_magnetic_field", 3, {0, 0, 0}, "T"}, {"dmi_xxy", 1, {0, 0, 0}, "J/m2"},
    {"dmi_xyz", 1, {0, 0, 0}, "J/m2"},        {"dmi_xxz", 1, {0, 0, 0}, "J/m2"},
    {"dmi_yxy", 1, {0, 0, 0}, "J/m2"},        {"dmi_yyz", 1, {0, 0, 0}, "J/m2"},
    {"dmi_yxz", 1, {0, 0, 0}, "J/m2"},        {"dmi_zxy", 1, {0, 0, 0}, "J/m2"},
    {"dmi_zyz", 1, {0, 0, 0}, "J/m2"},        {"dmi_zxz", 1, {0, 0, 0}, "J/m2"},
};
const char* kDmiKeys[D_COUNT] = {"dmi_xxy", "dmi_xyz", "dmi_xxz", "dmi_yxy", "dmi_yyz",
                                 "dmi_yxz", "dmi_zxy", "dmi_zyz", "dmi_zxz"};

// Magnetisation at construction: normalised N(0,1) vectors from a fixed-seed
// generator (the reference seeds from the clock, ferromagnet.cpp:62-76).
void random_unit_vectors(std::vector<float>& out, int N, unsigned long long seed) {
  unsigned long long s = seed;
  auto next = [&s]() {  // splitmix64
    s += 0x9E3779B97F4A7C15ull;
    unsigned long long z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  auto uniform = [&]() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); };
  out.resize((size_t)3 * N);
  for (int i = 0; i < N; ++i) {
    double v[3];
    for (int c = 0; c < 3; ++c) {  // Box-Muller
      const double u1 = uniform(), u2 = uniform();
      v[c] = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
    const double nrm = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int c = 0; c < 3; ++c) out[(size_t)c * N + i] = (float)(v[c] / nrm);
  }
}
}  // namespace

Ferromagnet::Ferromagnet(World* w, const int size[3], const int org[3], const uint8_t* geometry, std::string nm)
    : world(w), name(std::move(nm)) {
  for (int a = 0; a < 3; ++a) {
    if (size[a] < 1) throw Error(MPA_ERR_INVALID, "grid size should be positive");
    n[a] = size[a];
    origin[a] = org[a];
  }
  const long long NN = (long long)n[0] * n[1] * n[2];
  if (NN > (1LL << 30)) throw Error(MPA_ERR_INVALID, "grid too large");
  N = (int)NN;
  cellsingeo = N;
  if (geometry) {
    geom_host.assign(geometry, geometry + N);
    cellsingeo = 0;
    for (auto g : geom_host) cellsingeo += g ? 1 : 0;
    for (auto& g : geom_host) g = g ? 1 : 0;
    geom.allocate(N);
    HIP_CHECK(hipMemcpy(geom.get(), geom_host.data(), N, hipMemcpyHostToDevice));
  }
  for (const auto& sp : kParams) {
    Param p;
    p.name = name + ":" + sp.key;
    p.unit = sp.unit;
    p.ncomp = sp.ncomp;
    for (int c = 0; c < 3; ++c) p.uniform[c] = sp.def[c];
    params.emplace(sp.key, std::move(p));
  }
  m.allocate((size_t)3 * N);
  hdemag.allocate((size_t)3 * N);
  scratch.allocate((size_t)3 * N);
  std::vector<float> init;
  random_unit_vectors(init, N, 0x5eedULL + (unsigned long long)w->magnets.size());
  if (!geom_host.empty())
    for (int i = 0; i < N; ++i)
      if (!geom_host[i]) init[i] = init[(size_t)N + i] = init[(size_t)2 * N + i] = 0.f;
  HIP_CHECK(hipMemcpy(m.get(), init.data(), init.size() * sizeof(float), hipMemcpyHostToDevice));
}

Param& Ferromagnet::param(const std::string& key) {
  auto it = params.find(key);
  if (it == params.end()) throw Error(MPA_ERR_INVALID, "unknown parameter '" + key + "'");
  return it->second;
}

unsigned Ferromagnet::active_terms() const {
  auto az = [&](const char* k) { return params.at(k).assured_zero(); };
  const bool msat0 = az("msat");
  unsigned t = 0;
  // exchange.cu:14-18 (no region-pair exchange in this build)
  if (!(az("aex") || msat0)) t |= TERM_EXCH;
  // anisotropy.cu:9-23 and the uniaxial-else-cubic choice of :123-139
  const bool uni0 = (az("ku1") && az("ku2")) || az("anisU") || msat0;
  const bool cub0 = (az("kc1") && az("kc2") && az("kc3")) || (az("anisC1") && az("anisC2")) || msat0;
  if (!uni0) t |= TERM_UNI;
  else if (!cub0) t |= TERM_CUBIC;
  // zeeman.cu:38-51 (stray fields of other magnets are not coupled here)
  const bool wb0 = world->bias[0] == 0 && world->bias[1] == 0 && world->bias[2] == 0;
  if (!(wb0 && az("bias_magnetic_field"))) t |= TERM_EXT;
  // dmi.cu:80-85
  bool dmi0 = true;
  for (auto k : kDmiKeys) dmi0 = dmi0 && az(k);
  if (!(msat0 || dmi0)) t |= TERM_DMI;
  // demag.cpp:9-11
  if (enable_demag && !msat0) t |= TERM_DEMAG;
  return t;
}

FMDesc Ferromagnet::desc(unsigned terms) const {
  FMDesc d{};
  d.nx = n[0]; d.ny = n[1]; d.nz = n[2]; d.ncell = N;
  d.ox = origin[0]; d.oy = origin[1]; d.oz = origin[2];
  d.mgx = world->mastergrid_size[0]; d.mgy = world->mastergrid_size[1]; d.mgz = world->mastergrid_size[2];
  d.mox = world->mastergrid_origin[0]; d.moy = world->mastergrid_origin[1]; d.moz = world->mastergrid_origin[2];
  d.cx = (float)world->cellsize[0]; d.cy = (float)world->cellsize[1]; d.cz = (float)world->cellsize[2];
  d.geom = geom.get();
  d.msat = params.at("msat").ps();
  d.aex = params.at("aex").ps();
  d.alpha = params.at("alpha").ps();
  d.ku1 = params.at("ku1").ps();
  d.ku2 = params.at("ku2").ps();
  d.kc1 = params.at("kc1").ps();
  d.kc2 = params.at("kc2").ps();
  d.kc3 = params.at("kc3").ps();
  d.frozen = params.at("frozen_spins").ps();
  d.anisU = params.at("anisU").pv();
  d.anisC1 = params.at("anisC1").pv();
  d.anisC2 = params.at("anisC2").pv();
  d.bias = params.at("bias_magnetic_field").pv();
  for (int k = 0; k < D_COUNT; ++k) d.dmi[k] = params.at(kDmiKeys[k]).ps();
  for (int c = 0; c < 3; ++c) d.wb[c] = (float)world->bias[c];
  d.openbc = openbc ? 1 : 0;
  d.terms = terms;
  return d;
}

DemagPlan& Ferromagnet::demag() {
  if (!demag_) {
    DemagGeometry g;
    for (int a = 0; a < 3; ++a) {
      g.n[a] = n[a];
      g.cs[a] = world->cellsize[a];
      g.mastergrid[a] = world->mastergrid_size[a];
      g.pbc[a] = world->pbc[a];
    }
    demag_ = std::make_unique<DemagPlan>(g);
  }
  return *demag_;
}

void Ferromagnet::ensure_stage_buffers(int S) {
  if (x0.n == 0) x0.allocate((size_t)3 * N);
  while ((int)k.size() < S) k.emplace_back((size_t)3 * N);
}

// evalExternalField (zeeman.cu:70-88) adds the stray field of every other
// magnet that is a source (unless its msat is assured zero,
// strayfield.cpp:88-95) to a destination magnet.  That coupling is not part
// of this build, so an evaluation it would change is refused rather than
// computed without it.
void Ferromagnet::check_stray_coupling() const {
  if (!stray_destination) return;
  for (const auto& o : world->magnets) {
    if (o.get() == this || !o->stray_source || o->params.at("msat").assured_zero()) continue;
    throw Error(MPA_ERR_RUNTIME, "the stray field of magnet '" + o->name + "' on magnet '" + name +
                                     "' is not implemented; set enable_as_stray_field_source or "
                                     "enable_as_stray_field_destination to False");
  }
}

void Ferromagnet::eval_rhs(float* out, hipStream_t s, bool check_stray) {
  if (check_stray) check_stray_coupling();
  const unsigned terms = active_terms();
  Profiler* prof = &world->prof;
  if (terms & TERM_DEMAG) demag().exec(m.get(), params.at("msat").ps(), geom.get(), hdemag.get(), s, prof);
  // algorithmic bytes: read m, read h_demag (when present), write the torque
  Timed t(prof, "k_field", 12.0 * N * ((terms & TERM_DEMAG) ? 3 : 2), s);
  launch_field(desc(terms), m.get(), hdemag.get(), out, MODE_LLG, s);
}

namespace {
struct QInfo {
  const char* name;
  int ncomp;
  unsigned terms;  // 0xffffffff: all active terms
  int mode;        // FieldMode, or -1 for the magnetisation itself
};
constexpr unsigned ALL = 0xffffffffu;
const QInfo kQuantities[] = {
    {"magnetization", 3, 0, -1},
    {"exchange_field", 3, TERM_EXCH, MODE_FIELD},
    {"anisotropy_field", 3, TERM_UNI | TERM_CUBIC, MODE_FIELD},
    {"external_field", 3, TERM_EXT, MODE_FIELD},
    {"dmi_field", 3, TERM_DMI, MODE_FIELD},
    {"demag_field", 3, TERM_DEMAG, MODE_FIELD},
    {"effective_field", 3, ALL, MODE_FIELD},
    {"torque", 3, ALL, MODE_LLG},
    {"llg_torque", 3, ALL, MODE_LLG},
    {"damping_torque", 3, ALL, MODE_DAMPING},
    {"exchange_energy_density", 1, TERM_EXCH, MODE_EDENS},
    {"anisotropy_energy_density", 1, TERM_UNI | TERM_CUBIC, MODE_EDENS},
    {"zeeman_energy_density", 1, TERM_EXT, MODE_EDENS},
    {"dmi_energy_density", 1, TERM_DMI, MODE_EDENS},
    {"demag_energy_density", 1, TERM_DEMAG, MODE_EDENS},
    {"total_energy_density", 1, ALL, MODE_EDENS},
};
const QInfo& qinfo(const std::string& q) {
  for (const auto& qi : kQuantities)
    if (q == qi.name) return qi;
  throw Error(MPA_ERR_INVALID, "unknown quantity '" + q + "'");
}
}  // namespace

int Ferromagnet::quantity_ncomp(const std::string& q) const { return qinfo(q).ncomp; }

const float* Ferromagnet::eval_device(const std::string& q, hipStream_t s, int* ncomp) {
  const QInfo& qi = qinfo(q);
  *ncomp = qi.ncomp;
  if (qi.mode < 0) return m.get();
  if (qi.terms & TERM_EXT) check_stray_coupling();
  const unsigned terms = active_terms() & qi.terms;
  if (terms & TERM_DEMAG) demag().exec(m.get(), params.at("msat").ps(), geom.get(), hdemag.get(), s);
  launch_field(desc(terms), m.get(), hdemag.get(), scratch.get(), (FieldMode)qi.mode, s);
  return scratch.get();
}

int Ferromagnet::eval_quantity(const std::string& q, float* host, hipStream_t s) {
  int nc = 0;
  const float* dev = eval_device(q, s, &nc);
  HIP_CHECK(hipMemcpyAsync(host, dev, (size_t)nc * N * sizeof(float), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  return nc;
}

void Ferromagnet::average_quantity(const std::string& q, double* out, hipStream_t s) {
  int nc = 0;
  const float* dev = eval_device(q, s, &nc);
  world->reducer.component_sums(dev, nc, N, geom.get(), out, s);
  for (int c = 0; c < nc; ++c) out[c] /= (cellsingeo > 0 ? cellsingeo : 1);
}

double Ferromagnet::scalar_quantity(const std::string& q, hipStream_t s) {
  if (q == "max_torque") {
    eval_rhs(scratch.get(), s);
    return world->reducer.max_vec_norm(scratch.get(), N, geom.get(), s);
  }
  // energies: E = ncells * <e> * V (exchange.cu:247-256, energy.cu:78-92)
  static const std::map<std::string, std::string> dens = {
      {"exchange_energy", "exchange_energy_density"}, {"anisotropy_energy", "anisotropy_energy_density"},
      {"zeeman_energy", "zeeman_energy_density"},     {"dmi_energy", "dmi_energy_density"},
      {"demag_energy", "demag_energy_density"},       {"total_energy", "total_energy_density"}};
  auto it = dens.find(q);
  if (it == dens.end()) throw Error(MPA_ERR_INVALID, "unknown scalar quantity '" + q + "'");
  double avg = 0.0;
  average_quantity(it->second, &avg, s);
  const double V = world->cellsize[0] * world->cellsize[1] * world->cellsize[2];
  return (double)N * avg * V;
}

void Ferromagnet::set_magnetization(const float* host, long long count, hipStream_t s) {
  std::vector<float> buf((size_t)3 * N);
  if (count == 3) {
    for (int c = 0; c < 3; ++c)
      for (int i = 0; i < N; ++i) buf[(size_t)c * N + i] = host[c];
  } else if (count == 3LL * N) {
    std::memcpy(buf.data(), host, buf.size() * sizeof(float));
  } else {
    throw Error(MPA_ERR_INVALID, "magnetization: expected 3 or 3*ncells values");
  }
  // normalise inside the geometry, zero outside (NormalizedVariable::set)
  for (int i = 0; i < N; ++i) {
    const bool in = geom_host.empty() || geom_host[i];
    float x = buf[i], y = buf[(size_t)N + i], z = buf[(size_t)2 * N + i];
    if (in) {
      const double nr = std::sqrt((double)x * x + (double)y * y + (double)z * z);
      x = (float)(x / nr); y = (float)(y / nr); z = (float)(z / nr);
    } else {
      x = y = z = 0.f;
    }
    buf[i] = x; buf[(size_t)N + i] = y; buf[(size_t)2 * N + i] = z;
  }
  HIP_CHECK(hipMemcpyAsync(m.get(), buf.data(), buf.size() * sizeof(float), hipMemcpyHostToDevice, s));
  HIP_CHECK(hipStreamSynchronize(s));
  world->touch();
}

void Ferromagnet::get_magnetization(float* host, hipStream_t s) {
  HIP_CHECK(hipMemcpyAsync(host, m.get(), (size_t)3 * N * sizeof(float), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
}

// ------------------------------------------------------------ TimeSolver

TimeSolver::TimeSolver(World* w) : world_(w) {}

void TimeSolver::set_method(const std::string& nm) {
  tableau_by_name(nm);
  method = nm;
  fsal_valid_ = false;
  if (adaptive) dt = sensible_timestep();
}

void TimeSolver::equations_changed() {
  fsal_valid_ = false;
  if (adaptive) dt = sensible_timestep();
}

// TimeSolver::sensibleTimeStep (src/core/timesolver.cpp:37-47)
double TimeSolver::sensible_timestep() {
  if (world_->magnets.empty()) return 0.0;
  double mx = 0.0;
  for (auto& mag : world_->magnets) {
    // The reference re-estimates dt whenever a magnet is added
    // (timesolver.cpp:49-52), before the caller can switch stray coupling
    // off; the estimate runs without the refusal, the steps keep it.
    mag->eval_rhs(mag->scratch.get(), world_->stream, false);
    ++nrhs;
    mx = std::max(mx, (double)world_->reducer.max_vec_norm(mag->scratch.get(), mag->N, mag->geom_ptr(), world_->stream));
  }
  if (mx == 0.0) return sensible_default;
  return sensible_factor / mx;
}

// TimeSolver::adaptTimeStep (src/core/timesolver.cpp:60-75)
void TimeSolver::adapt(double corr) {
  if (!adaptive) return;
  if (std::isnan(corr)) corr = 1.0;
  corr *= headroom;
  if (lower_bound >= upper_bound) throw Error(MPA_ERR_RUNTIME, "The lower bound should be lower than the upper bound.");
  corr = corr > upper_bound ? upper_bound : corr;
  corr = corr < lower_bound ? lower_bound : corr;
  dt *= corr;
}

// RungeKuttaStepper::step (src/core/rungekutta.cpp:25-84).  The stage states
// are x0 + dt * sum_j a_sj k_j, normalised; the final state uses weights1;
// the error is the max-norm of dt * sum_j (w1_j - w2_j) k_j.  For tableaus
// whose last stage is the final state (DormandPrince, BogackiShampine) the
// last stage's k is the next step's first k (first-same-as-last), and the
// final combination is the last stage state itself.
void TimeSolver::step() {
  if (dt <= 0)
    throw Error(MPA_ERR_RUNTIME,
                "Timesolver can not make a step because the timestep is smaller than or equal to zero.");
  auto& mags = world_->magnets;
  if (mags.empty()) {
    time += dt;
    return;
  }
  hipStream_t s = world_->stream;
  const Tableau& tb = tableau_by_name(method);
  const int S = (int)tb.nodes.size();
  Profiler* prof = &world_->prof;
  for (auto& mag : mags) {
    mag->ensure_stage_buffers(S);
    Timed t(prof, "copy_x0", 24.0 * mag->N, s);
    HIP_CHECK(hipMemcpyAsync(mag->x0.get(), mag->m.get(), (size_t)3 * mag->N * sizeof(float),
                             hipMemcpyDeviceToDevice, s));
  }
  const double t0 = time;
  bool k0_valid = fsal_valid_ && fsal_version_ == world_->version && fsal_time_ == t0;
  while (true) {
    const double h = dt;
    for (int st = 0; st < S; ++st) {
      time = t0 + h * tb.nodes[st];
      if (st > 0) {
        for (auto& mag : mags) {
          Combo co{};
          co.nk = st;
          for (int j = 0; j < st; ++j) {
            co.k[j] = mag->k[j].get();
            co.c[j] = (float)(h * tb.A[st][j]);
          }
          Timed t(prof, "k_combine", 12.0 * mag->N * (st + 2), s);
          launch_combine(mag->m.get(), mag->x0.get(), co, 3, mag->N, mag->geom_ptr(), true, s);
        }
      }
      if (st == 0 && k0_valid) continue;
      for (auto& mag : mags) {
        mag->eval_rhs(mag->k[st].get(), s);
        ++nrhs;
      }
    }
    time = t0 + h;
    if (!tb.fsal) {
      for (auto& mag : mags) {
        Combo co{};
        co.nk = S;
        for (int j = 0; j < S; ++j) {
          co.k[j] = mag->k[j].get();
          co.c[j] = (float)(h * tb.w1[j]);
        }
        Timed t(prof, "k_combine", 12.0 * mag->N * (S + 2), s);
        launch_combine(mag->m.get(), mag->x0.get(), co, 3, mag->N, mag->geom_ptr(), true, s);
      }
    }
    ++nsteps;
    if (!adaptive) break;
    double err = 0.0;
    for (auto& mag : mags) {
      Combo co{};
      co.nk = S;
      for (int j = 0; j < S; ++j) {
        co.k[j] = mag->k[j].get();
        co.c[j] = (float)(h * (tb.w1[j] - tb.w2[j]));
      }
      Timed t(prof, "k_combo_maxnorm", 12.0 * mag->N * S, s);
      err = std::max(err, (double)world_->reducer.max_combo_norm(co, mag->N, mag->geom_ptr(), s));
    }
    const bool success = err < max_error;
    const double ratio = max_error / err;  // err == 0 gives +inf, clamped by adapt()
    const double corr = success ? std::pow(ratio, 1.0 / (tb.order2 + 1)) : std::pow(ratio, 1.0 / (tb.order1 + 1));
    adapt(corr);
    if (success) break;
    ++nrejected;
    for (auto& mag : mags)
      HIP_CHECK(hipMemcpyAsync(mag->m.get(), mag->x0.get(), (size_t)3 * mag->N * sizeof(float),
                               hipMemcpyDeviceToDevice, s));
    time = t0;
    k0_valid = true;  // k[0] = rhs(x0, t0) was computed in this attempt
  }
  if (tb.fsal) {
    for (auto& mag : mags) std::swap(mag->k[0], mag->k[S - 1]);
    fsal_valid_ = true;
    fsal_time_ = time;
    fsal_version_ = world_->version;
  } else {
    fsal_valid_ = false;
  }
}

void TimeSolver::steps(long long n) {
  for (long long i = 0; i < n; ++i) step();
  HIP_CHECK(hipStreamSynchronize(world_->stream));
  world_->prof.flush_sync(world_->stream);
}

// TimeSolver::run (src/core/timesolver.cpp:97-111)
void TimeSolver::run(double duration) {
  if (duration <= 0) return;
  const double stop = time + duration;
  while (time < stop - dt) step();
  const double old = dt;
  dt = stop - time;
  step();
  if (!adaptive) dt = old;
  HIP_CHECK(hipStreamSynchronize(world_->stream));
}

// ------------------------------------------------------------ World

World::World(const double cs[3], const int mg_size[3], const int mg_origin[3], const int reps[3]) : solver(this) {
  for (int a = 0; a < 3; ++a) {
    if (!(cs[a] > 0)) throw Error(MPA_ERR_INVALID, "cellsize should be positive");
    // the reference keeps the cell size as FP32 `real3` (world.hpp:73,
    // datatypes.hpp:15): geometry callables, stencils and the Newell kernel
    // all see the rounded value
    cellsize[a] = (double)(float)cs[a];
    mastergrid_size[a] = mg_size[a];
    mastergrid_origin[a] = mg_origin[a];
    pbc[a] = reps[a];
  }
  check_pbc(mastergrid_size, pbc);
  HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
}

World::~World() {
  magnets.clear();
  if (stream) (void)hipStreamDestroy(stream);
}

// World::checkPbcRepetitions / checkPbcCompatibility (src/core/world.cpp:62-76)
void World::check_pbc(const int mg_size[3], const int reps[3]) const {
  for (int a = 0; a < 3; ++a)
    if (reps[a] < 0) throw Error(MPA_ERR_INVALID, "Number of pbcRepetitions should not be negative.");
  for (int a = 0; a < 3; ++a)
    if ((mg_size[a] == 0) != (reps[a] == 0))
      throw Error(MPA_ERR_INVALID, "0 in size of mastergrid should match 0 in pbcRepetitions.");
}

void World::check_in_mastergrid(const int size[3], const int org[3], const int mg_size[3],
                                const int mg_origin[3]) const {
  for (int a = 0; a < 3; ++a) {
    if (mg_size[a] == 0) continue;
    if (org[a] < mg_origin[a] || org[a] + size[a] > mg_origin[a] + mg_size[a])
      throw Error(MPA_ERR_INVALID, "The grid of the magnet does not fit inside the mastergrid.");
  }
}

Ferromagnet* World::add_ferromagnet(const int size[3], const int org[3], const uint8_t* geometry, const char* nm) {
  static int unnamed = 1;  // ferromagnet_<n> naming of mumaxworld.cpp:61-64
  std::string name = nm ? nm : "";
  if (name.empty()) name = "ferromagnet_" + std::to_string(unnamed++);
  if (find(name)) throw Error(MPA_ERR_INVALID, "A magnet with the name '" + name + "' already exists");
  check_in_mastergrid(size, org, mastergrid_size, mastergrid_origin);
  magnets.push_back(std::make_unique<Ferromagnet>(this, size, org, geometry, name));
  touch();
  solver.equations_changed();
  return magnets.back().get();
}

Ferromagnet* World::find(const std::string& name) {
  for (auto& m : magnets)
    if (m->name == name) return m.get();
  return nullptr;
}

void World::bounding_grid(int size[3], int org[3]) const {
  if (magnets.empty()) {
    for (int a = 0; a < 3; ++a) size[a] = org[a] = 0;
    return;
  }
  int lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = magnets[0]->origin[a];
    hi[a] = lo[a] + magnets[0]->n[a];
  }
  for (auto& m : magnets)
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], m->origin[a]);
      hi[a] = std::max(hi[a], m->origin[a] + m->n[a]);
    }
  for (int a = 0; a < 3; ++a) {
    size[a] = hi[a] - lo[a];
    org[a] = lo[a];
  }
}

void World::pbc_changed() {
  for (auto& m : magnets) m->invalidate_demag();
  touch();
}

// MumaxWorld::setPBC(mastergrid, reps) (mumaxworld.cpp:311-320)
void World::set_pbc(const int mg_size[3], const int mg_origin[3], const int reps[3]) {
  check_pbc(mg_size, reps);
  for (auto& m : magnets) check_in_mastergrid(m->n, m->origin, mg_size, mg_origin);
  for (int a = 0; a < 3; ++a) {
    mastergrid_size[a] = mg_size[a];
    mastergrid_origin[a] = mg_origin[a];
    pbc[a] = reps[a];
  }
  pbc_changed();
}

// MumaxWorld::setPBC(reps): mastergrid = bounding grid, open where reps == 0 (:288-309)
void World::set_pbc_auto(const int reps[3]) {
  for (int a = 0; a < 3; ++a)
    if (reps[a] < 0) throw Error(MPA_ERR_INVALID, "Number of pbcRepetitions should not be negative.");
  int size[3], org[3];
  bounding_grid(size, org);
  for (int a = 0; a < 3; ++a)
    if (reps[a] == 0) size[a] = 0;
  set_pbc(size, org, reps);
}

void World::unset_pbc() {
  for (int a = 0; a < 3; ++a) mastergrid_size[a] = mastergrid_origin[a] = pbc[a] = 0;
  pbc_changed();
}

}  // namespace mpa
