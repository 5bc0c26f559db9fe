// Host object model and the Runge-Kutta time stepper.
#include "world.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <deque>

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

// ------------------------------------------------------------ parameters

namespace {
struct ParamSpec {
  const char* key;
  int ncomp;
  float def[3];
  const char* unit;
  const char* label;  // name suffix when it is not ":" + key
};
// Defaults of the reference Ferromagnet constructor (ferromagnet.cpp:24-61).
const ParamSpec kFmParams[] = {
    {"msat", 1, {1.f, 0, 0}, "A/m"},          {"aex", 1, {0, 0, 0}, "J/m"},
    {"alpha", 1, {0, 0, 0}, ""},              {"ku1", 1, {0, 0, 0}, "J/m3"},
    {"ku2", 1, {0, 0, 0}, "J/m3", "ku2"},     {"kc1", 1, {0, 0, 0}, "J/m3"},
    {"kc2", 1, {0, 0, 0}, "J/m3"},            {"kc3", 1, {0, 0, 0}, "J/m3", "kc3"},
    {"frozen_spins", 1, {0, 0, 0}, ""},       {"anisU", 3, {0, 0, 0}, ""},
    {"anisC1", 3, {0, 0, 0}, ""},             {"anisC2", 3, {0, 0, 0}, ""},
    {"bias_magnetic_field", 3, {0, 0, 0}, "T"}, {"dmi_xxy", 1, {0, 0, 0}, "J/m2"},
    {"dmi_xyz", 1, {0, 0, 0}, "J/m2"},        {"dmi_xxz", 1, {0, 0, 0}, "J/m2"},
    {"dmi_yxy", 1, {0, 0, 0}, "J/m2"},        {"dmi_yyz", 1, {0, 0, 0}, "J/m2"},
    {"dmi_yxz", 1, {0, 0, 0}, "J/m2"},        {"dmi_zxy", 1, {0, 0, 0}, "J/m2"},
    {"dmi_zyz", 1, {0, 0, 0}, "J/m2"},        {"dmi_zxz", 1, {0, 0, 0}, "J/m2"},
    {"B1", 1, {0, 0, 0}, "J/m3"},             {"B2", 1, {0, 0, 0}, "J/m3", ":B1"},
    // spin-transfer torque and temperature (ferromagnet.cpp:34-56); the
    // free layer thickness defaults to the grid's thickness (set per magnet)
    {"temperature", 1, {0, 0, 0}, "K"},       {"xi", 1, {0, 0, 0}, ""},
    {"Lambda", 1, {0, 0, 0}, "", ":lambda"},  {"free_layer_thickness", 1, {0, 0, 0}, "m"},
    {"epsilon_prime", 1, {0, 0, 0}, ""},      {"fixed_layer", 3, {0, 0, 0}, ""},
    {"pol", 1, {0, 0, 0}, ""},                {"jcur", 3, {0, 0, 0}, "A/m2"},
    // electrical potential (ferromagnet.cpp:53-55): no applied potential is NaN
    {"applied_potential", 1, {__builtin_nanf(""), 0, 0}, "V"},
    {"conductivity", 1, {0, 0, 0}, "S/m"},    {"amr_ratio", 1, {0, 0, 0}, ""},
};
// The labels of ku2, kc3 (no colon) and B2 (":B1") are the reference's own
// parameter names (ferromagnet.cpp:29,32,61), exposed as Parameter.name.
// Elastic parameters of every magnet (magnet.cpp:26-34).
const ParamSpec kElasticParams[] = {
    {"C11", 1, {0, 0, 0}, "N/m2"},  {"C12", 1, {0, 0, 0}, "N/m2"},
    {"C44", 1, {0, 0, 0}, "N/m2"},  {"eta", 1, {0, 0, 0}, "kg/m3s"},
    {"rho", 1, {1.f, 0, 0}, "kg/m3"}, {"external_body_force", 3, {0, 0, 0}, "N/m3"},
    {"rigid_norm_strain", 3, {0, 0, 0}, ""}, {"rigid_shear_strain", 3, {0, 0, 0}, ""},
};
// Host parameters of an antiferromagnet (antiferromagnet.cpp:15-25) and its
// inter-sublattice DMI tensor.
const ParamSpec kAfmParams[] = {
    {"afmex_cell", 1, {0, 0, 0}, "J/m"}, {"afmex_nn", 1, {0, 0, 0}, "J/m"},
    {"latcon", 1, {0.35e-9f, 0, 0}, "m"}, {"dmi_xxy", 1, {0, 0, 0}, "J/m2"},
    {"dmi_xyz", 1, {0, 0, 0}, "J/m2"},   {"dmi_xxz", 1, {0, 0, 0}, "J/m2"},
    {"dmi_yxy", 1, {0, 0, 0}, "J/m2"},   {"dmi_yyz", 1, {0, 0, 0}, "J/m2"},
    {"dmi_yxz", 1, {0, 0, 0}, "J/m2"},   {"dmi_zxy", 1, {0, 0, 0}, "J/m2"},
    {"dmi_zyz", 1, {0, 0, 0}, "J/m2"},   {"dmi_zxz", 1, {0, 0, 0}, "J/m2"},
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

// ------------------------------------------------------------ MagnetCore

MagnetCore::MagnetCore(World* w, const int size[3], const int org[3], const uint8_t* geometry, std::string nm)
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
}

void MagnetCore::add_params(const void* specs, int count) {
  const ParamSpec* sp = static_cast<const ParamSpec*>(specs);
  for (int i = 0; i < count; ++i) {
    Param p;
    p.name = sp[i].label ? name + sp[i].label : name + ":" + sp[i].key;
    p.unit = sp[i].unit;
    p.ncomp = sp[i].ncomp;
    for (int c = 0; c < 3; ++c) p.uniform[c] = sp[i].def[c];
    p.N = N;
    p.world = world;
    params.emplace(sp[i].key, std::move(p));
  }
}

// DynamicParameter::evalTimeDependentTerms (dynamic_parameter.cpp:8-29) at
// the solver's time, added to the static value (parameter.cu:65-110).
void Param::refresh() const {
  const double t = world->solver.time;
  if (fresh_ && at_ == t) return;
  double u[3] = {uniform[0], uniform[1], uniform[2]};
  std::vector<std::pair<const float*, std::array<float, 3>>> masked;
  for (const auto& term : terms) {
    double v[3] = {0.0, 0.0, 0.0};
    if (term.fn(t, term.ctx, v) != 0)
      throw Error(MPA_ERR_RUNTIME, "the time-dependent term of parameter '" + name + "' failed");
    if (term.mask.n) {
      masked.push_back({term.mask.get(), {(float)v[0], (float)v[1], (float)v[2]}});
    } else {
      for (int c = 0; c < ncomp; ++c) u[c] += v[c];
    }
  }
  for (int c = 0; c < 3; ++c) eff_[c] = c < ncomp ? (float)u[c] : 0.f;
  dyn_cells_ = cells.n != 0 || !masked.empty();
  if (dyn_cells_) {
    const size_t n = (size_t)ncomp * N;
    if (dyn_.n != n) dyn_.allocate(n);
    // the first launch starts from the static cells, later ones (more than
    // ParamFill::MAX masks) accumulate onto the first's output
    size_t k = 0;
    bool first = true;
    do {
      ParamFill f{};
      f.nmask = (int)std::min<size_t>(ParamFill::MAX, masked.size() - k);
      for (int j = 0; j < f.nmask; ++j) {
        f.mask[j] = masked[k + j].first;
        for (int c = 0; c < 3; ++c) f.coef[j][c] = masked[k + j].second[c];
      }
      for (int c = 0; c < 3; ++c) f.add[c] = first ? eff_[c] : 0.f;
      launch_param_fill(dyn_.get(), first ? cells.get() : dyn_.get(), f, ncomp, N, world->stream);
      k += f.nmask;
      first = false;
    } while (k < masked.size());
  }
  at_ = t;
  fresh_ = true;
}

Param& MagnetCore::param(const std::string& key) {
  auto it = params.find(key);
  if (it == params.end()) throw Error(MPA_ERR_INVALID, "unknown parameter '" + key + "'");
  return it->second;
}

const Param& MagnetCore::param(const std::string& key) const {
  auto it = params.find(key);
  if (it == params.end()) throw Error(MPA_ERR_INVALID, "unknown parameter '" + key + "'");
  return it->second;
}

void MagnetCore::add_inter(const std::string& key, float value, const char* unit) {
  InterParam p;
  p.name = name + ":" + key;
  p.unit = unit;
  p.uniform = value;
  inter_params.emplace(key, std::move(p));
}

InterParam& MagnetCore::inter(const std::string& key) {
  auto it = inter_params.find(key);
  if (it == inter_params.end()) throw Error(MPA_ERR_INVALID, "unknown inter-region parameter '" + key + "'");
  return it->second;
}

const InterParam& MagnetCore::inter(const std::string& key) const {
  auto it = inter_params.find(key);
  if (it == inter_params.end()) throw Error(MPA_ERR_INVALID, "unknown inter-region parameter '" + key + "'");
  return it->second;
}

void MagnetCore::set_regions(const uint32_t* r) {
  regions_host.clear();
  unique_regions.clear();
  regions.release();
  if (r) {
    regions_host.assign(r, r + N);
    unique_regions = regions_host;
    std::sort(unique_regions.begin(), unique_regions.end());
    unique_regions.erase(std::unique(unique_regions.begin(), unique_regions.end()), unique_regions.end());
    regions.allocate(N);
    HIP_CHECK(hipMemcpy(regions.get(), regions_host.data(), (size_t)N * sizeof(uint32_t), hipMemcpyHostToDevice));
  }
  // inter_parameter.cu:7-20: N = largest index + 1, at least 1
  const size_t nr = unique_regions.empty() ? 1 : (size_t)unique_regions.back() + 1;
  for (auto& kv : inter_params) {
    kv.second.limit = nr * (nr - 1) / 2;
    kv.second.host.clear();
    kv.second.lut.release();
  }
  world->touch();
}

void MagnetCore::check_region(unsigned idx) const {
  if (!std::binary_search(unique_regions.begin(), unique_regions.end(), (uint32_t)idx))
    throw Error(MPA_ERR_INVALID, "The region index " + std::to_string(idx) + " is not defined.");
}

void MagnetCore::param_set_in_region(const std::string& key, unsigned idx, const double* value) {
  check_region(idx);
  Param& p = param(key);
  const size_t n = (size_t)p.ncomp * N;
  std::vector<float> h(n);
  if (p.cells.n == 0) {
    bool same = true;
    for (int c = 0; c < p.ncomp; ++c) same = same && (float)value[c] == p.uniform[c];
    if (same) return;
    for (int c = 0; c < p.ncomp; ++c) std::fill(h.begin() + (size_t)c * N, h.begin() + (size_t)(c + 1) * N, p.uniform[c]);
    p.cells.allocate(n);
    // the per-cell values now hold the static value; a time term adds to them
    for (int c = 0; c < 3; ++c) p.uniform[c] = 0.f;
  } else {
    HIP_CHECK(hipMemcpy(h.data(), p.cells.get(), n * sizeof(float), hipMemcpyDeviceToHost));
  }
  for (int i = 0; i < N; ++i)
    if (regions_host[i] == idx)
      for (int c = 0; c < p.ncomp; ++c) h[(size_t)c * N + i] = (float)value[c];
  HIP_CHECK(hipMemcpy(p.cells.get(), h.data(), n * sizeof(float), hipMemcpyHostToDevice));
  p.changed();
  world->touch();
}

void MagnetCore::inter_set(const std::string& key, double value) {
  InterParam& p = inter(key);
  p.uniform = (float)value;
  p.host.clear();
  p.lut.release();
  world->touch();
}

void MagnetCore::inter_set_between(const std::string& key, unsigned i, unsigned j, double value) {
  check_region(i);
  check_region(j);
  InterParam& p = inter(key);
  if (i == j)
    throw Error(MPA_ERR_INVALID,
                "Can not set " + p.name + ", region indexes must be different: " + std::to_string(i) + ".");
  if (p.is_uniform()) {
    if ((float)value == p.uniform) return;
    if (p.limit > ((size_t)1 << 28))
      throw Error(MPA_ERR_INVALID, "Can not set " + p.name + ": region indices up to " +
                                       std::to_string(unique_regions.back()) + " need too large a table.");
    p.host.assign(p.limit, p.uniform);
    p.lut.allocate(p.limit);
  }
  const size_t k = IP::index(i, j);
  p.host[k] = (float)value;
  HIP_CHECK(hipMemcpy(p.lut.get(), p.host.data(), p.limit * sizeof(float), hipMemcpyHostToDevice));
  world->touch();
}

double MagnetCore::inter_get_between(const std::string& key, unsigned i, unsigned j) const {
  check_region(i);
  check_region(j);
  const InterParam& p = inter(key);
  if (i == j)
    throw Error(MPA_ERR_INVALID,
                "Can not get " + p.name + ", region indexes must be different: " + std::to_string(i) + ".");
  return p.is_uniform() ? p.uniform : p.host[IP::index(i, j)];
}

// Magnet::setEnableElastodynamics (magnet.cpp:177-213): u and v start at zero.
void MagnetCore::set_elastodynamics(bool on) {
  if (on == el.enabled) return;
  el.enabled = on;
  if (on) {
    el.u.allocate((size_t)3 * N);
    el.v.allocate((size_t)3 * N);
    el.strain.allocate((size_t)6 * N);
    el.stress.allocate((size_t)6 * N);
    HIP_CHECK(hipMemset(el.u.get(), 0, el.u.bytes()));
    HIP_CHECK(hipMemset(el.v.get(), 0, el.v.bytes()));
  } else {
    el = ElasticState{};
  }
  world->touch();
  world->solver.equations_changed();
}

bool MagnetCore::elasticity_assured_zero() const {
  return !el.enabled ||
         (param("C11").assured_zero() && param("C12").assured_zero() && param("C44").assured_zero());
}

bool MagnetCore::rigid_strain_applied() const {
  return !(param("rigid_norm_strain").assured_zero() && param("rigid_shear_strain").assured_zero());
}

ElDesc MagnetCore::eldesc() const {
  ElDesc d{};
  d.nx = n[0]; d.ny = n[1]; d.nz = n[2]; d.ncell = N;
  d.ox = origin[0]; d.oy = origin[1]; d.oz = origin[2];
  d.mgx = world->mastergrid_size[0]; d.mgy = world->mastergrid_size[1]; d.mgz = world->mastergrid_size[2];
  d.mox = world->mastergrid_origin[0]; d.moy = world->mastergrid_origin[1]; d.moz = world->mastergrid_origin[2];
  d.cx = (float)world->cellsize[0]; d.cy = (float)world->cellsize[1]; d.cz = (float)world->cellsize[2];
  d.geom = geom.get();
  d.C11 = param("C11").ps();
  d.C12 = param("C12").ps();
  d.C44 = param("C44").ps();
  d.eta = param("eta").ps();
  d.rho = param("rho").ps();
  d.fext = param("external_body_force").pv();
  d.elastic = elasticity_assured_zero() ? 0 : 1;
  d.ext = param("external_body_force").assured_zero() ? 0 : 1;
  d.damping = (el.enabled && !param("eta").assured_zero()) ? 1 : 0;
  d.divrho = param("rho").assured_zero() ? 0 : 1;
  return d;
}

MEForce MagnetCore::me_force_of(int i) {
  MEForce f{};
  Ferromagnet* fm = fm_at(i);
  if (fm->magnetoelastic_assured_zero()) return f;
  f.m = fm->m.get();
  f.B1 = fm->param("B1").ps();
  f.B2 = fm->param("B2").ps();
  return f;
}

void MagnetCore::compute_strain(hipStream_t s) {
  if (!el.enabled) throw Error(MPA_ERR_RUNTIME, "elastodynamics is not enabled for '" + name + "'");
  Timed t(&world->prof, "k_strain", 12.0 * N + 48.0 * N, s);
  launch_strain(eldesc(), el.u.get(), el.strain.get(), el.stress.get(), s);
}

void MagnetCore::ensure_elastic_stage_buffers(int S) {
  if (el.u0.n == 0) el.u0.allocate((size_t)3 * N);
  if (el.v0.n == 0) el.v0.allocate((size_t)3 * N);
  while ((int)el.ku.size() < S) el.ku.emplace_back((size_t)3 * N);
  while ((int)el.kv.size() < S) el.kv.emplace_back((size_t)3 * N);
}

// Variable::set / get of elastic_displacement and elastic_velocity
// (wrap_variable.cpp:8-16): 3 values (uniform) or 3*ncells; zero outside the
// geometry.
void MagnetCore::set_variable(const std::string& which, const float* host, long long count, hipStream_t s) {
  if (!el.enabled)
    throw Error(MPA_ERR_RUNTIME, which + " Variable does not exist yet. Enable elastodynamics first.");
  DeviceArray<float>* dst = which == "elastic_displacement" ? &el.u : (which == "elastic_velocity" ? &el.v : nullptr);
  if (!dst) throw Error(MPA_ERR_INVALID, "unknown variable '" + which + "'");
  std::vector<float> buf((size_t)3 * N);
  if (count == 3) {
    for (int c = 0; c < 3; ++c)
      for (int i = 0; i < N; ++i) buf[(size_t)c * N + i] = host[c];
  } else if (count == 3LL * N) {
    std::memcpy(buf.data(), host, buf.size() * sizeof(float));
  } else {
    throw Error(MPA_ERR_INVALID, which + ": expected 3 or 3*ncells values");
  }
  if (!geom_host.empty())
    for (int i = 0; i < N; ++i)
      if (!geom_host[i]) buf[i] = buf[(size_t)N + i] = buf[(size_t)2 * N + i] = 0.f;
  HIP_CHECK(hipMemcpyAsync(dst->get(), buf.data(), buf.size() * sizeof(float), hipMemcpyHostToDevice, s));
  HIP_CHECK(hipStreamSynchronize(s));
  world->touch();
}

void MagnetCore::get_variable(const std::string& which, float* host, hipStream_t s) {
  if (!el.enabled)
    throw Error(MPA_ERR_RUNTIME, which + " Variable does not exist yet. Enable elastodynamics first.");
  const DeviceArray<float>* src =
      which == "elastic_displacement" ? &el.u : (which == "elastic_velocity" ? &el.v : nullptr);
  if (!src) throw Error(MPA_ERR_INVALID, "unknown variable '" + which + "'");
  HIP_CHECK(hipMemcpyAsync(host, src->get(), (size_t)3 * N * sizeof(float), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
}

namespace {
enum MagnetQMode { MQ_STRAIN = -1, MQ_STRESS = -2, MQ_VELOCITY = -3, MQ_DISPLACEMENT = -4, MQ_KINETIC = -5,
                   MQ_ELASTIC_EDENS = -6, MQ_POYNTING = -7 };
struct MagnetQ {
  const char* name;
  int ncomp;
  int mode;  // ElMode, or a MagnetQMode
};
const MagnetQ kMagnetQuantities[] = {
    {"strain_tensor", 6, MQ_STRAIN},
    {"stress_tensor", 6, MQ_STRESS},
    {"elastic_force", 3, EL_FORCE},
    {"effective_body_force", 3, EL_BODY},
    {"elastic_damping", 3, EL_DAMPING},
    {"elastic_acceleration", 3, EL_ACCEL},
    {"elastic_velocity", 3, MQ_VELOCITY},
    {"elastic_displacement", 3, MQ_DISPLACEMENT},
    {"kinetic_energy_density", 1, MQ_KINETIC},
    {"elastic_energy_density", 1, MQ_ELASTIC_EDENS},
    {"poynting_vector", 3, MQ_POYNTING},
};
}  // namespace

int MagnetCore::magnet_quantity_ncomp(const std::string& q) const {
  for (const auto& mq : kMagnetQuantities)
    if (q == mq.name) return mq.ncomp;
  return 0;
}

// Magnet-level quantities (straintensor.cu:99-116, stresstensor.cu:42-61,
// elasticforce.cu:104-124, elastodynamics.cu:23-96, elasticdamping.cu:30-49,
// elasticenergies.cu:25-101, poyntingvector.cu:32-46) with their assured-zero
// rules: strain needs elastodynamics; stress, elastic force, elastic energy
// and energy flux need some non-zero stiffness too; damping needs eta;
// kinetic energy needs rho.
const float* MagnetCore::magnet_eval_device(const std::string& q, hipStream_t s, int* ncomp) {
  const MagnetQ* mq = nullptr;
  for (const auto& e : kMagnetQuantities)
    if (q == e.name) mq = &e;
  if (!mq) throw Error(MPA_ERR_INVALID, "unknown quantity '" + q + "'");
  *ncomp = mq->ncomp;
  const int mode = mq->mode;
  if (!el.enabled && (mode == MQ_VELOCITY || mode == MQ_DISPLACEMENT))
    throw Error(MPA_ERR_RUNTIME, q + " Variable does not exist yet. Enable elastodynamics first.");
  if (mode == MQ_VELOCITY) return el.v.get();
  if (mode == MQ_DISPLACEMENT) return el.u.get();
  if (qbuf.n == 0) qbuf.allocate((size_t)6 * N);
  float* out = qbuf.get();
  const ElDesc d = eldesc();
  bool zero = false;
  switch (mode) {
    case MQ_STRAIN: zero = !el.enabled; break;
    case MQ_STRESS: case EL_FORCE: case MQ_ELASTIC_EDENS: case MQ_POYNTING: zero = elasticity_assured_zero(); break;
    case EL_DAMPING: zero = !d.damping; break;
    case MQ_KINETIC: zero = !el.enabled || param("rho").assured_zero(); break;
    default: break;
  }
  if (zero) {
    HIP_CHECK(hipMemsetAsync(out, 0, (size_t)mq->ncomp * N * sizeof(float), s));
    return out;
  }
  const bool needs_strain = mode == MQ_STRAIN || mode == MQ_STRESS || mode == MQ_ELASTIC_EDENS ||
                            mode == MQ_POYNTING || ((mode == EL_FORCE || mode == EL_BODY || mode == EL_ACCEL) && d.elastic);
  if (needs_strain) compute_strain(s);
  if (mode == MQ_STRAIN) return el.strain.get();
  if (mode == MQ_STRESS) return el.stress.get();
  if (mode == MQ_KINETIC) {
    launch_elastic_local(d, nullptr, nullptr, el.v.get(), out, EL_KINETIC_EDENS, s);
  } else if (mode == MQ_ELASTIC_EDENS) {
    launch_elastic_local(d, el.stress.get(), el.strain.get(), nullptr, out, EL_ELASTIC_EDENS, s);
  } else if (mode == MQ_POYNTING) {
    launch_elastic_local(d, el.stress.get(), nullptr, el.v.get(), out, EL_POYNTING, s);
  } else {
    launch_elastic(d, el.enabled ? el.stress.get() : nullptr, el.enabled ? el.v.get() : nullptr, me_force_of(0),
                   num_fm() > 1 ? me_force_of(1) : MEForce{}, out, (ElMode)mode, s);
  }
  return out;
}

// kinetic_energy / elastic_energy: ncells * <e> * V (elasticenergies.cu:40-47, :87-95)
double MagnetCore::magnet_scalar(const std::string& q, hipStream_t s) {
  const char* dens = q == "kinetic_energy" ? "kinetic_energy_density"
                                           : (q == "elastic_energy" ? "elastic_energy_density" : nullptr);
  if (!dens) throw Error(MPA_ERR_INVALID, "unknown scalar quantity '" + q + "'");
  int nc = 0;
  const float* dev = magnet_eval_device(dens, s, &nc);
  double sum = 0.0;
  world->reducer.component_sums(dev, 1, N, geom.get(), &sum, s);
  const double avg = sum / (cellsingeo > 0 ? cellsingeo : 1);
  const double V = world->cellsize[0] * world->cellsize[1] * world->cellsize[2];
  return (double)N * avg * V;
}

namespace {
// out = (a1 Ms1 m1 + a2 Ms2 m2) * scale, with scale = 1 or 1/(Ms1 + Ms2);
// m2 may be nullptr (one magnet).
__global__ void k_afm_combine(const float* __restrict__ m1, PS ms1, const float* __restrict__ m2, PS ms2, float a1,
                              float a2, int normalize, int N, const uint8_t* __restrict__ geom,
                              float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const bool in = !geom || geom[i];
  const float s1 = ms1.at(i), s2 = ms2.at(i);
  float sc = 1.f;
  if (normalize) sc = (s1 + s2) != 0.f ? 1.f / (s1 + s2) : 0.f;
  for (int c = 0; c < 3; ++c)
    out[(size_t)c * N + i] =
        in ? sc * (a1 * s1 * m1[(size_t)c * N + i] + (m2 ? a2 * s2 * m2[(size_t)c * N + i] : 0.f)) : 0.f;
}

// k_angle (afmexchange.cu:260-281): acos(sign(afmex_cell) m1 . m2), 0 outside
// the geometry and where msat1, msat2 or afmex_cell is 0.
__global__ void k_afm_angle(const float* __restrict__ m1, const float* __restrict__ m2, PS afmex, PS ms1, PS ms2, int N,
                            const uint8_t* __restrict__ geom, float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const float a = afmex.at(i);
  if ((geom && !geom[i]) || ms1.at(i) == 0.f || ms2.at(i) == 0.f || a == 0.f) {
    out[i] = 0.f;
    return;
  }
  const float d = m1[i] * m2[i] + m1[N + i] * m2[N + i] + m1[2 * N + i] * m2[2 * N + i];
  out[i] = (float)acos(copysign(1.0, (double)a) * (double)d);
}

// the potential in the geometry (poissonsystem.cu:292-297), 0 outside
__global__ void k_double_to_float(const double* __restrict__ x, const uint8_t* __restrict__ geom, int N,
                                  float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) out[i] = (geom && !geom[i]) ? 0.f : (float)x[i];
}

__global__ void k_add2(const float* __restrict__ a, const float* __restrict__ b, float* __restrict__ out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a[i] + b[i];
}
}  // namespace

// Kinetic and elastic energy densities added to out (energy.cu:64-66).
void MagnetCore::add_elastic_energies(float* out, hipStream_t s) {
  if (!el.enabled) return;
  const ElDesc d = eldesc();
  if (!param("rho").assured_zero()) launch_elastic_local(d, nullptr, nullptr, el.v.get(), out, EL_KINETIC_EDENS, s, true);
  if (!elasticity_assured_zero()) {
    compute_strain(s);
    launch_elastic_local(d, el.stress.get(), el.strain.get(), nullptr, out, EL_ELASTIC_EDENS, s, true);
  }
}

// ------------------------------------------------------------ Ferromagnet

Ferromagnet::Ferromagnet(World* w, const int size[3], const int org[3], const uint8_t* geometry, std::string nm,
                         Antiferromagnet* h)
    : MagnetCore(w, size, org, geometry, std::move(nm)), host(h) {
  add_params(kFmParams, (int)(sizeof(kFmParams) / sizeof(kFmParams[0])));
  // ferromagnet.cpp:26-27
  add_inter("inter_exchange", 0.f, "J/m");
  add_inter("scale_exchange", 1.f, "");
  param("free_layer_thickness").uniform[0] = (float)n[2] * (float)w->cellsize[2];
  noise_seed = 0x9e3779b9u * (unsigned)(1 + w->magnets.size() + 7 * w->afms.size()) +
               (host ? (unsigned)(name.back() - '0') * 0x85ebca6bu : 0u);
  if (!host) add_params(kElasticParams, (int)(sizeof(kElasticParams) / sizeof(kElasticParams[0])));
  m.allocate((size_t)3 * N);
  if (!host) hdemag.allocate((size_t)3 * N);
  scratch.allocate((size_t)3 * N);
  std::vector<float> init;
  random_unit_vectors(init, N, 0x5eedULL + (unsigned long long)w->magnets.size() + 7ULL * w->afms.size() +
                                   (host ? (unsigned long long)(name.back() - '0') * 1000ULL : 0ULL));
  if (!geom_host.empty())
    for (int i = 0; i < N; ++i)
      if (!geom_host[i]) init[i] = init[(size_t)N + i] = init[(size_t)2 * N + i] = 0.f;
  HIP_CHECK(hipMemcpy(m.get(), init.data(), init.size() * sizeof(float), hipMemcpyHostToDevice));
}

Ferromagnet* Ferromagnet::sister() const {
  if (!host) return nullptr;
  return host->sub[0].get() == this ? host->sub[1].get() : host->sub[0].get();
}

MagnetCore* Ferromagnet::owner() { return host ? static_cast<MagnetCore*>(host) : static_cast<MagnetCore*>(this); }

bool Ferromagnet::magnetoelastic_assured_zero() {
  if (param("msat").assured_zero() || (param("B1").assured_zero() && param("B2").assured_zero())) return true;
  MagnetCore* own = owner();
  return !own->el.enabled && !own->rigid_strain_applied();
}

unsigned Ferromagnet::active_terms() {
  auto az = [&](const char* k) { return params.at(k).assured_zero(); };
  const bool msat0 = az("msat");
  unsigned t = 0;
  // exchange.cu:14-18
  if (!((az("aex") && inter("inter_exchange").assured_zero()) || inter("scale_exchange").assured_zero() || msat0))
    t |= TERM_EXCH;
  // anisotropy.cu:9-23 and the uniaxial-else-cubic choice of :123-139
  const bool uni0 = (az("ku1") && az("ku2")) || az("anisU") || msat0;
  const bool cub0 = (az("kc1") && az("kc2") && az("kc3")) || (az("anisC1") && az("anisC2")) || msat0;
  if (!uni0) t |= TERM_UNI;
  else if (!cub0) t |= TERM_CUBIC;
  // zeeman.cu:46-51: world bias, magnet bias and the other magnets' stray fields
  const bool wb0 = world->bias[0] == 0 && world->bias[1] == 0 && world->bias[2] == 0;
  if (!(wb0 && az("bias_magnetic_field")) || owner()->stray_active()) t |= TERM_EXT;
  // dmi.cu:12-17: a sublattice also counts the inter-sublattice tensor
  bool dmi0 = true;
  for (auto k : kDmiKeys) dmi0 = dmi0 && az(k);
  if (host)
    for (auto k : kDmiKeys) dmi0 = dmi0 && host->param(k).assured_zero();
  if (!(msat0 || dmi0)) t |= TERM_DMI;
  // demag.cpp:9-11
  if (enable_demag && !msat0) t |= TERM_DEMAG;
  if (!magnetoelastic_assured_zero()) t |= TERM_ME;
  // stt.cu:9-26: Slonczewski when it is not assured zero, else Zhang-Li
  const bool zl0 = !enable_zhangli || msat0 || az("jcur") || az("pol");
  const bool sl0 = !enable_slonc || msat0 || az("jcur") || az("free_layer_thickness") || az("fixed_layer") ||
                   (az("epsilon_prime") && (az("Lambda") || az("pol")));
  if (!sl0) t |= TERM_SLONC;
  else if (!zl0) t |= TERM_ZHANGLI;
  // thermalnoise.cu:17-19
  if (!az("temperature")) t |= TERM_THERMAL;
  // afmexchange.cu:13-25
  if (host) {
    const bool ms2_0 = sister()->param("msat").assured_zero();
    if (!(host->param("afmex_nn").assured_zero() || ms2_0)) t |= TERM_AFM_NN;
    if (!(host->param("afmex_cell").assured_zero() || ms2_0)) t |= TERM_AFM_SITE;
  }
  return t;
}

FMDesc Ferromagnet::desc(unsigned terms) {
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
  d.tile_zc = world->field_zchunk;
  d.hstray = nullptr;
  if ((terms & TERM_EXT) && owner()->stray_active()) {
    MagnetCore* own = owner();
    if (own->hstray.n == 0) own->hstray.allocate((size_t)3 * own->N);
    d.hstray = own->hstray.get();
  }
  d.B1 = params.at("B1").ps();
  d.B2 = params.at("B2").ps();
  d.pol = params.at("pol").ps();
  d.xi = params.at("xi").ps();
  d.lambda = params.at("Lambda").ps();
  d.epsp = params.at("epsilon_prime").ps();
  d.fthick = params.at("free_layer_thickness").ps();
  d.jcur = params.at("jcur").pv();
  d.fixed = params.at("fixed_layer").pv();
  d.fixed_on_top = fixed_on_top ? 1 : 0;
  d.temperature = params.at("temperature").ps();
  {
    // 2 kB gamma / V in FP32, as the reference forms it (thermalnoise.cu:66-67)
    const float V = d.cx * d.cy * d.cz;
    d.noise_pref = 2.f * 1.38064852e-23f * 1.7595e11f / V;
  }
  d.noise_seed = noise_seed;
  if (terms & TERM_ME) {
    MagnetCore* own = owner();
    d.me_rigid = own->rigid_strain_applied() ? 1 : 0;
    d.rnorm = own->param("rigid_norm_strain").pv();
    d.rshear = own->param("rigid_shear_strain").pv();
    d.strain = own->el.strain.get();
  }
  if (host) {
    Ferromagnet* s2 = sister();
    d.sublattice = 1;
    d.m2 = s2->m.get();
    d.msat2 = s2->param("msat").ps();
    d.afmex_cell = host->param("afmex_cell").ps();
    d.afmex_nn = host->param("afmex_nn").ps();
    d.latcon = host->param("latcon").ps();
    for (int k = 0; k < D_COUNT; ++k) d.hdmi[k] = host->param(kDmiKeys[k]).ps();
  }
  // inter-region exchange: the regions matter when a pair value or scale is
  // not the default (inter 0: the harmonic mean, scale 1)
  auto trivial = [](const InterParam& in, const InterParam& sc) {
    return in.assured_zero() && sc.is_uniform() && sc.uniform == 1.f;
  };
  const InterParam& ie = inter("inter_exchange");
  const InterParam& se = inter("scale_exchange");
  d.inter_ex = ie.ip();
  d.scale_ex = se.ip();
  bool matter = !trivial(ie, se);
  if (host) {
    const InterParam& ia = host->inter("inter_afmex_nn");
    const InterParam& sa = host->inter("scale_afmex_nn");
    d.inter_afm = ia.ip();
    d.scale_afm = sa.ip();
    matter = matter || !trivial(ia, sa);
  } else {
    d.inter_afm = IP{nullptr, 0.f};
    d.scale_afm = IP{nullptr, 1.f};
  }
  d.regions = (matter && regions.n) ? regions.get() : nullptr;
  return d;
}

DemagPlan& Ferromagnet::demag() {
  if (host) return host->demag();
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

const float* Ferromagnet::demag_field() const { return host ? host->hdemag.get() : hdemag.get(); }

void Ferromagnet::ensure_stage_buffers(int S) {
  if (x0.n == 0) x0.allocate((size_t)3 * N);
  while ((int)k.size() < S) k.emplace_back((size_t)3 * N);
}

// Demag field (the host's field of both sublattices for a sublattice) and the
// strain of the magnet holding the displacement, as far as `terms` use them.
void Ferromagnet::prepare(unsigned terms, hipStream_t s, Profiler* prof) {
  if (terms & TERM_DEMAG) {
    if (host) host->compute_demag(s, prof);
    else demag().exec(m.get(), params.at("msat").ps(), geom.get(), hdemag.get(), s, prof);
  }
  if ((terms & TERM_ME) && !owner()->rigid_strain_applied()) owner()->compute_strain(s);
  if (terms & TERM_EXT) owner()->compute_stray(s, prof);
}

// ------------------------------------------------------------ stray fields

bool Ferromagnet::stray_source_of(StraySource& out) {
  if (host || !stray_source || param("msat").assured_zero()) return false;
  out = StraySource{m.get(), param("msat").ps(), nullptr, PS{nullptr, 0.f}, geom.get(),
                    {origin[0], origin[1], origin[2]}, {n[0], n[1], n[2]}};
  return true;
}

bool Antiferromagnet::stray_source_of(StraySource& out) {
  if (!stray_source) return false;
  if (sub[0]->param("msat").assured_zero() && sub[1]->param("msat").assured_zero()) return false;
  out = StraySource{sub[0]->m.get(), sub[0]->param("msat").ps(), sub[1]->m.get(), sub[1]->param("msat").ps(),
                    geom.get(), {origin[0], origin[1], origin[2]}, {n[0], n[1], n[2]}};
  return true;
}

// StrayField(magnet, grid).eval (strayfield.cpp:72-77, wrap_strayfield.cpp:12-24):
// the field of `src` on an arbitrary grid, by the FFT box convolution
// (method 0) or the direct sum (method 1); zero when the source's msat is
// assured zero.  The source switches do not apply here.
void World::stray_field_on_grid(MagnetCore* src, const int size[3], const int org[3], int method, float* host_out) {
  for (int a = 0; a < 3; ++a)
    if (size[a] < 1) throw Error(MPA_ERR_INVALID, "grid size should be positive");
  const size_t nd = (size_t)size[0] * size[1] * size[2];
  StraySource S{};
  bool zero;
  if (auto* fm = dynamic_cast<Ferromagnet*>(src)) {
    if (fm->host) throw Error(MPA_ERR_INVALID, "the stray field source must be a magnet, not a sublattice");
    zero = fm->param("msat").assured_zero();
    S = StraySource{fm->m.get(), fm->param("msat").ps(), nullptr, PS{nullptr, 0.f}, fm->geom.get(),
                    {fm->origin[0], fm->origin[1], fm->origin[2]}, {fm->n[0], fm->n[1], fm->n[2]}};
  } else {
    auto* a = static_cast<Antiferromagnet*>(src);
    zero = a->sub[0]->param("msat").assured_zero() && a->sub[1]->param("msat").assured_zero();
    S = StraySource{a->sub[0]->m.get(), a->sub[0]->param("msat").ps(), a->sub[1]->m.get(),
                    a->sub[1]->param("msat").ps(), a->geom.get(), {a->origin[0], a->origin[1], a->origin[2]},
                    {a->n[0], a->n[1], a->n[2]}};
  }
  if (zero) {
    std::fill(host_out, host_out + 3 * nd, 0.f);
    return;
  }
  DeviceArray<float> out(3 * nd);
  StrayTarget t{{org[0], org[1], org[2]}, {size[0], size[1], size[2]}, nullptr, out.get()};
  if (method == 0) {
    int bo[3], bn[3];
    stray_box(&S, 1, t, bo, bn);
    StrayPlan plan(bo, bn, cellsize, mastergrid_size, pbc);
    plan.exec(&S, 1, t, false, stream, nullptr);
  } else if (method == 1) {
    stray_brute(S, t, cellsize, mastergrid_size, pbc, stream);
  } else {
    throw Error(MPA_ERR_INVALID, "Method should be \"fft\" or \"brute\"");
  }
  HIP_CHECK(hipMemcpyAsync(host_out, out.get(), 3 * nd * sizeof(float), hipMemcpyDeviceToHost, stream));
  HIP_CHECK(hipStreamSynchronize(stream));
}

// evalExternalField (zeeman.cu:70-88): a destination-enabled magnet (a
// sublattice uses its antiferromagnet) receives the stray field of every
// other magnet that is enabled as a source and whose msat is not assured zero
// (strayFieldsAssuredZero, zeeman.cu:13-34).
std::vector<std::pair<MagnetCore*, StraySource>> MagnetCore::stray_sources() {
  std::vector<std::pair<MagnetCore*, StraySource>> v;
  if (!stray_destination) return v;
  StraySource src;
  for (auto& o : world->magnets)
    if (o.get() != this && o->stray_source_of(src)) v.emplace_back(o.get(), src);
  for (auto& a : world->afms)
    if (a.get() != this && a->stray_source_of(src)) v.emplace_back(a.get(), src);
  return v;
}

// One convolution per source, as the reference registers one StrayField per
// (source, destination) pair (mumaxworld.cpp:106-116): the demag convolution
// of the source's magnetisation on the box that holds the source and this
// magnet, read back on this magnet's cells and summed.
void MagnetCore::compute_stray(hipStream_t s, Profiler* prof) {
  const std::vector<std::pair<MagnetCore*, StraySource>> src = stray_sources();
  if (src.empty()) return;
  if (hstray.n == 0) hstray.allocate((size_t)3 * N);
  StrayTarget t{{origin[0], origin[1], origin[2]}, {n[0], n[1], n[2]}, geom.get(), hstray.get()};
  for (size_t k = 0; k < src.size(); ++k) {
    int bo[3], bn[3];
    stray_box(&src[k].second, 1, t, bo, bn);
    StrayPlan* plan = nullptr;
    for (auto& e : stray_plans)
      if (e.first == src[k].first && e.second->matches(bo, bn)) plan = e.second.get();
    if (!plan) {
      for (auto it = stray_plans.begin(); it != stray_plans.end(); ++it)
        if (it->first == src[k].first) {
          stray_plans.erase(it);
          break;
        }
      stray_plans.emplace_back(src[k].first,
                               std::make_unique<StrayPlan>(bo, bn, world->cellsize, world->mastergrid_size, world->pbc));
      plan = stray_plans.back().second.get();
    }
    plan->exec(&src[k].second, 1, t, k > 0, s, prof);
  }
}

// algorithmic bytes of one right-hand side: read m, h_demag (when present),
// the sister's m (for a sublattice) and the strain (dynamic magnetoelastic
// term); write the torque
double Ferromagnet::rhs_bytes(unsigned terms) {
  double bytes = 12.0 * N * ((terms & TERM_DEMAG) ? 3 : 2);
  if (host) bytes += 12.0 * N;
  if ((terms & TERM_ME) && !owner()->rigid_strain_applied()) bytes += 24.0 * N;
  return bytes;
}

FMDesc Ferromagnet::rhs_desc(unsigned terms, double noise_dt) {
  FMDesc d = desc(terms);
  if (noise_dt > 0 && (terms & TERM_THERMAL)) {
    // k += noise / sqrt(dt) with the step's draw (rungekutta.cpp:95-106, in FP32)
    d.noise = 1;
    d.noise_scale = 1.f / sqrtf((float)noise_dt);
    d.noise_purpose = 0;
    d.noise_draw = step_draw;
  }
  return d;
}

void Ferromagnet::eval_rhs(float* out, hipStream_t s, bool prepared, FieldMode mode, double noise_dt,
                           const NextStage* next) {
  const unsigned terms = active_terms();
  Profiler* prof = &world->prof;
  if (!prepared) prepare(terms, s, prof);
  if (next) {
    // plus the combination: read x0 (unless it is the cell's own state) and
    // nk derivatives, write the next stage state
    const double bytes = rhs_bytes(terms) + 12.0 * N * ((next->x0 ? 1 : 0) + next->nk + 1);
    Timed t(prof, "k_field_next", bytes, s);
    launch_field_next(rhs_desc(terms, noise_dt), m.get(), demag_field(), out, mode, *next, s);
    return;
  }
  Timed t(prof, "k_field", rhs_bytes(terms), s);
  launch_field(rhs_desc(terms, noise_dt), m.get(), demag_field(), out, mode, s);
}

bool Ferromagnet::next_fusable() {
  if (host || el.enabled) return false;
  const unsigned terms = active_terms();
  if (terms & TERM_THERMAL) return false;
  return field_next_fusable(desc(terms), MODE_LLG);
}

// Both sublattices' right-hand sides in one launch (everything they read is
// prepared).
void Antiferromagnet::eval_rhs_pair(float* out0, float* out1, hipStream_t s, FieldMode mode, double noise_dt) {
  Ferromagnet* a = sub[0].get();
  Ferromagnet* b = sub[1].get();
  const unsigned ta = a->active_terms(), tb = b->active_terms();
  Timed t(&world->prof, "k_field_pair", a->rhs_bytes(ta) + b->rhs_bytes(tb), s);
  launch_field_pair(a->rhs_desc(ta, noise_dt), a->m.get(), a->demag_field(), out0, b->rhs_desc(tb, noise_dt),
                    b->m.get(), b->demag_field(), out1, mode, s);
}

bool Antiferromagnet::next_fusable() {
  Ferromagnet* a = sub[0].get();
  Ferromagnet* b = sub[1].get();
  const unsigned ta = a->active_terms(), tb = b->active_terms();
  if ((ta | tb) & TERM_THERMAL) return false;
  return field_pair_next_fusable(a->desc(ta), b->desc(tb), MODE_LLG);
}

void Antiferromagnet::eval_rhs_pair_next(float* out0, float* out1, hipStream_t s, const NextStage& n0,
                                         const NextStage& n1) {
  Ferromagnet* a = sub[0].get();
  Ferromagnet* b = sub[1].get();
  const unsigned ta = a->active_terms(), tb = b->active_terms();
  // plus each sublattice's combination: read x0 and nk derivatives, write the next stage state
  const double bytes = a->rhs_bytes(ta) + b->rhs_bytes(tb) + 2 * 12.0 * N * (n0.nk + 2);
  Timed t(&world->prof, "k_field_pair_next", bytes, s);
  launch_field_pair_next(a->rhs_desc(ta, 0.0), a->m.get(), a->demag_field(), out0, n0, b->rhs_desc(tb, 0.0),
                         b->m.get(), b->demag_field(), out1, n1, MODE_LLG, s);
}

bool Antiferromagnet::pair_split(FieldMode mode, double noise_dt) {
  Ferromagnet* a = sub[0].get();
  Ferromagnet* b = sub[1].get();
  return torque_pair_splittable(a->rhs_desc(a->active_terms(), noise_dt), b->rhs_desc(b->active_terms(), noise_dt),
                                mode);
}

void Antiferromagnet::eval_local_pair(hipStream_t s, double noise_dt) {
  Ferromagnet* a = sub[0].get();
  Ferromagnet* b = sub[1].get();
  const unsigned ta = a->active_terms() & ~TERM_DEMAG, tb = b->active_terms() & ~TERM_DEMAG;
  for (auto& h : hloc)
    if (h.n == 0) h.allocate((size_t)3 * N);
  Timed t(&world->prof, "k_field_pair", a->rhs_bytes(ta) + b->rhs_bytes(tb), s);
  launch_field_pair(a->rhs_desc(ta, noise_dt), a->m.get(), nullptr, hloc[0].get(), b->rhs_desc(tb, noise_dt),
                    b->m.get(), nullptr, hloc[1].get(), MODE_FIELD, s);
}

void Antiferromagnet::eval_torque_pair(float* out0, float* out1, hipStream_t s, FieldMode mode, double noise_dt) {
  Ferromagnet* a = sub[0].get();
  Ferromagnet* b = sub[1].get();
  // read m, the field without demag and the demag field; write the torque
  Timed t(&world->prof, "k_torque_pair", 2 * 48.0 * N, s);
  launch_torque_pair(a->rhs_desc(a->active_terms(), noise_dt), a->m.get(), hloc[0].get(), a->demag_field(), out0,
                     b->rhs_desc(b->active_terms(), noise_dt), b->m.get(), hloc[1].get(), b->demag_field(), out1,
                     mode, s);
}

namespace {
struct QInfo {
  const char* name;
  int ncomp;
  unsigned terms;
  int mode;  // FieldMode, or -1 for the magnetisation itself
};
constexpr unsigned ALL = ~0u;
const QInfo kQuantities[] = {
    {"magnetization", 3, 0, -1},
    {"exchange_field", 3, TERM_EXCH, MODE_FIELD},
    {"anisotropy_field", 3, TERM_UNI | TERM_CUBIC, MODE_FIELD},
    {"external_field", 3, TERM_EXT, MODE_FIELD},
    {"dmi_field", 3, TERM_DMI, MODE_FIELD},
    {"demag_field", 3, TERM_DEMAG, MODE_FIELD},
    {"magnetoelastic_field", 3, TERM_ME, MODE_FIELD},
    {"homogeneous_exchange_field", 3, TERM_AFM_SITE, MODE_FIELD},
    {"inhomogeneous_exchange_field", 3, TERM_AFM_NN, MODE_FIELD},
    {"effective_field", 3, ALL, MODE_FIELD},
    {"torque", 3, ALL, MODE_LLG},
    {"llg_torque", 3, ALL & ~TERM_STT, MODE_LLG},
    {"spin_transfer_torque", 3, TERM_STT, MODE_STT},
    {"thermal_noise", 3, TERM_THERMAL, MODE_NOISE},
    {"damping_torque", 3, ALL, MODE_DAMPING},
    {"exchange_energy_density", 1, TERM_EXCH, MODE_EDENS},
    {"anisotropy_energy_density", 1, TERM_UNI | TERM_CUBIC, MODE_EDENS},
    {"zeeman_energy_density", 1, TERM_EXT, MODE_EDENS},
    {"dmi_energy_density", 1, TERM_DMI, MODE_EDENS},
    {"demag_energy_density", 1, TERM_DEMAG, MODE_EDENS},
    {"magnetoelastic_energy_density", 1, TERM_ME, MODE_EDENS},
    {"homogeneous_exchange_energy_density", 1, TERM_AFM_SITE, MODE_EDENS},
    {"inhomogeneous_exchange_energy_density", 1, TERM_AFM_NN, MODE_EDENS},
    {"total_energy_density", 1, ALL, MODE_EDENS},
};
const QInfo* find_qinfo(const std::string& q) {
  for (const auto& qi : kQuantities)
    if (q == qi.name) return &qi;
  return nullptr;
}
}  // namespace

PoissonDesc Ferromagnet::poisson_desc() {
  PoissonDesc d{};
  d.nx = n[0]; d.ny = n[1]; d.nz = n[2];
  d.cx = world->cellsize[0]; d.cy = world->cellsize[1]; d.cz = world->cellsize[2];
  d.geom = geom.get();
  d.sigma0 = param("conductivity").ps();
  d.amr = param("amr_ratio").ps();
  d.m = m.get();
  d.aniso = param("amr_ratio").assured_zero() ? 0 : 1;  // poissonsystem.cu:262-263
  d.pot = param("applied_potential").ps();
  return d;
}

int Ferromagnet::quantity_ncomp(const std::string& q) const {
  if (const QInfo* qi = find_qinfo(q)) return qi->ncomp;
  if (q == "magnetoelastic_force" || q == "full_magnetization") return 3;
  if (q == "conductivity_tensor") return 6;
  if (q == "electrical_potential") return 1;
  if (!host) {
    const int nc = magnet_quantity_ncomp(q);
    if (nc) return nc;
  }
  throw Error(MPA_ERR_INVALID, "unknown quantity '" + q + "'");
}

const float* Ferromagnet::eval_device(const std::string& q, hipStream_t s, int* ncomp) {
  const QInfo* qi = find_qinfo(q);
  if (!qi) {
    if (q == "magnetoelastic_force") {
      // magnetoelasticforce.cu:97-111 (geometry and cell size of the owner)
      *ncomp = 3;
      MEForce f{};
      if (!magnetoelastic_assured_zero()) f = MEForce{m.get(), param("B1").ps(), param("B2").ps()};
      launch_elastic(owner()->eldesc(), nullptr, nullptr, f, MEForce{}, scratch.get(), EL_MEFORCE, s);
      return scratch.get();
    }
    if (q == "full_magnetization") {
      // Ms m (fullmag.cu:9-11)
      *ncomp = 3;
      k_afm_combine<<<(N + 255) / 256, 256, 0, s>>>(m.get(), param("msat").ps(), nullptr, PS{nullptr, 0.f}, 1.f, 0.f,
                                                    0, N, geom.get(), scratch.get());
      checkLaunch("k_afm_combine");
      return scratch.get();
    }
    if (q == "conductivity_tensor") {
      // evalConductivityTensor (conductivitytensor.cu:34-41)
      *ncomp = 6;
      if (qbuf.n < (size_t)6 * N) qbuf.allocate((size_t)6 * N);
      PoissonDesc pd = poisson_desc();
      pd.aniso = 1;
      launch_conductivity_tensor(pd, qbuf.get(), s);
      return qbuf.get();
    }
    if (q == "electrical_potential") {
      // evalElectricalPotential (electricalpotential.cu:9-19): PoissonSystem::solve
      *ncomp = 1;
      poisson.init(poisson_desc(), s);
      poisson.solve(s);
      k_double_to_float<<<(N + 255) / 256, 256, 0, s>>>(poisson.x(), geom.get(), N, scratch.get());
      checkLaunch("k_double_to_float");
      return scratch.get();
    }
    if (!host && magnet_quantity_ncomp(q)) return magnet_eval_device(q, s, ncomp);
    throw Error(MPA_ERR_INVALID, "unknown quantity '" + q + "'");
  }
  *ncomp = qi->ncomp;
  if (qi->mode < 0) return m.get();
  const unsigned terms = active_terms() & qi->terms;
  prepare(terms, s, nullptr);
  FMDesc d = desc(terms);
  if (qi->mode == MODE_NOISE) {
    // a fresh thermal field each evaluation (evalThermalNoise, thermalnoise.cu:46-69)
    d.noise_purpose = 1;
    d.noise_draw = noise_draws[1]++;
  }
  launch_field(d, m.get(), demag_field(), scratch.get(), (FieldMode)qi->mode, s);
  if (q == "total_energy_density" && !host) add_elastic_energies(scratch.get(), s);
  return scratch.get();
}

int Ferromagnet::eval_quantity(const std::string& q, float* host_out, hipStream_t s) {
  int nc = 0;
  const float* dev = eval_device(q, s, &nc);
  HIP_CHECK(hipMemcpyAsync(host_out, dev, (size_t)nc * N * sizeof(float), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  return nc;
}

void Ferromagnet::average_quantity(const std::string& q, double* out, hipStream_t s) {
  int nc = 0;
  const float* dev = eval_device(q, s, &nc);
  for (int c0 = 0; c0 < nc; c0 += 3) {
    const int cnt = std::min(3, nc - c0);
    world->reducer.component_sums(dev + (size_t)c0 * N, cnt, N, geom.get(), out + c0, s);
  }
  for (int c = 0; c < nc; ++c) out[c] /= (cellsingeo > 0 ? cellsingeo : 1);
}

double Ferromagnet::scalar_quantity(const std::string& q, hipStream_t s) {
  if (q == "max_torque") {
    eval_rhs(scratch.get(), s);
    return world->reducer.max_vec_norm(scratch.get(), N, geom.get(), s);
  }
  if (q == "max_angle") {
    // evalMaxAngle (exchange.cu:315-321)
    launch_max_angle(desc(0), m.get(), scratch.get(), s);
    return world->reducer.max_abs(scratch.get(), N, geom.get(), s);
  }
  // energies: E = ncells * <e> * V (exchange.cu:247-256, energy.cu:78-92)
  static const std::map<std::string, std::string> dens = {
      {"exchange_energy", "exchange_energy_density"},
      {"anisotropy_energy", "anisotropy_energy_density"},
      {"zeeman_energy", "zeeman_energy_density"},
      {"dmi_energy", "dmi_energy_density"},
      {"demag_energy", "demag_energy_density"},
      {"magnetoelastic_energy", "magnetoelastic_energy_density"},
      {"homogeneous_exchange_energy", "homogeneous_exchange_energy_density"},
      {"inhomogeneous_exchange_energy", "inhomogeneous_exchange_energy_density"},
      {"total_energy", "total_energy_density"}};
  auto it = dens.find(q);
  if (it == dens.end()) {
    if (!host && (q == "kinetic_energy" || q == "elastic_energy")) return magnet_scalar(q, s);
    throw Error(MPA_ERR_INVALID, "unknown scalar quantity '" + q + "'");
  }
  double avg = 0.0;
  average_quantity(it->second, &avg, s);
  const double V = world->cellsize[0] * world->cellsize[1] * world->cellsize[2];
  return (double)N * avg * V;
}

void Ferromagnet::set_magnetization(const float* host_in, long long count, hipStream_t s) {
  std::vector<float> buf((size_t)3 * N);
  if (count == 3) {
    for (int c = 0; c < 3; ++c)
      for (int i = 0; i < N; ++i) buf[(size_t)c * N + i] = host_in[c];
  } else if (count == 3LL * N) {
    std::memcpy(buf.data(), host_in, buf.size() * sizeof(float));
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

void Ferromagnet::get_magnetization(float* host_out, hipStream_t s) {
  HIP_CHECK(hipMemcpyAsync(host_out, m.get(), (size_t)3 * N * sizeof(float), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
}

// ------------------------------------------------------------ Antiferromagnet

Antiferromagnet::Antiferromagnet(World* w, const int size[3], const int org[3], const uint8_t* geometry,
                                 std::string nm)
    : MagnetCore(w, size, org, geometry, std::move(nm)) {
  add_params(kAfmParams, (int)(sizeof(kAfmParams) / sizeof(kAfmParams[0])));
  add_params(kElasticParams, (int)(sizeof(kElasticParams) / sizeof(kElasticParams[0])));
  // antiferromagnet.cpp:22-23
  add_inter("inter_afmex_nn", 0.f, "J/m");
  add_inter("scale_afmex_nn", 1.f, "");
  const uint8_t* g = geom_host.empty() ? nullptr : geom_host.data();
  sub[0] = std::make_unique<Ferromagnet>(w, size, org, g, name + ":sublattice_1", this);
  sub[1] = std::make_unique<Ferromagnet>(w, size, org, g, name + ":sublattice_2", this);
  hdemag.allocate((size_t)3 * N);
  scratch.allocate((size_t)3 * N);
}

void Antiferromagnet::set_regions(const uint32_t* r) {
  MagnetCore::set_regions(r);
  sub[0]->set_regions(r);
  sub[1]->set_regions(r);
}

DemagPlan& Antiferromagnet::demag() {
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

// StrayFieldFFTExecutor::exec for an antiferromagnet source: the field of
// Ms1 m1 + Ms2 m2 (strayfieldfft.cu:206-213), shared by both sublattices.
void Antiferromagnet::compute_demag(hipStream_t s, Profiler* prof) {
  demag().exec(sub[0]->m.get(), sub[0]->param("msat").ps(), geom.get(), hdemag.get(), s, prof, sub[1]->m.get(),
               sub[1]->param("msat").ps());
}

int Antiferromagnet::quantity_ncomp(const std::string& q) const {
  if (q == "neel_vector" || q == "full_magnetization") return 3;
  if (q == "total_energy_density" || q == "angle_field") return 1;
  const int nc = magnet_quantity_ncomp(q);
  if (nc) return nc;
  throw Error(MPA_ERR_INVALID, "unknown quantity '" + q + "'");
}


// neel_vector (Ms1 m1 - Ms2 m2)/(Ms1 + Ms2) and full_magnetization Ms1 m1 + Ms2 m2
// (antiferromagnet.py:341-356 of the reference package).
const float* Antiferromagnet::eval_device(const std::string& q, hipStream_t s, int* ncomp) {
  if (q == "neel_vector" || q == "full_magnetization") {
    *ncomp = 3;
    const bool neel = q == "neel_vector";
    k_afm_combine<<<(N + 255) / 256, 256, 0, s>>>(sub[0]->m.get(), sub[0]->param("msat").ps(), sub[1]->m.get(),
                                                  sub[1]->param("msat").ps(), 1.f, neel ? -1.f : 1.f, neel ? 1 : 0,
                                                  N, geom.get(), scratch.get());
    checkLaunch("k_afm_combine");
    return scratch.get();
  }
  if (q == "angle_field") {
    // evalAngleField (afmexchange.cu:283-294)
    *ncomp = 1;
    k_afm_angle<<<(N + 255) / 256, 256, 0, s>>>(sub[0]->m.get(), sub[1]->m.get(), param("afmex_cell").ps(),
                                               sub[0]->param("msat").ps(), sub[1]->param("msat").ps(), N, geom.get(),
                                               scratch.get());
    checkLaunch("k_afm_angle");
    return scratch.get();
  }
  if (q == "total_energy_density") {
    // energy.cu:70-76: both sublattices, then the magnet's kinetic and elastic
    *ncomp = 1;
    int nc = 0;
    const float* e1 = sub[0]->eval_device(q, s, &nc);
    const float* e2 = sub[1]->eval_device(q, s, &nc);
    k_add2<<<(N + 255) / 256, 256, 0, s>>>(e1, e2, scratch.get(), N);
    checkLaunch("k_add2");
    add_elastic_energies(scratch.get(), s);
    return scratch.get();
  }
  return magnet_eval_device(q, s, ncomp);
}

double Antiferromagnet::scalar_quantity(const std::string& q, hipStream_t s) {
  if (q == "kinetic_energy" || q == "elastic_energy") return magnet_scalar(q, s);
  if (q == "max_intracell_angle") {
    // evalMaxAngle (afmexchange.cu:296-298)
    int nc = 0;
    const float* a = eval_device("angle_field", s, &nc);
    return world->reducer.max_abs(a, N, geom.get(), s);
  }
  if (q != "total_energy") throw Error(MPA_ERR_INVALID, "unknown scalar quantity '" + q + "'");
  double avg = 0.0;
  average_quantity("total_energy_density", &avg, s);
  const double V = world->cellsize[0] * world->cellsize[1] * world->cellsize[2];
  return (double)N * avg * V;
}

int Antiferromagnet::eval_quantity(const std::string& q, float* host_out, hipStream_t s) {
  int nc = 0;
  const float* dev = eval_device(q, s, &nc);
  HIP_CHECK(hipMemcpyAsync(host_out, dev, (size_t)nc * N * sizeof(float), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  return nc;
}

void Antiferromagnet::average_quantity(const std::string& q, double* out, hipStream_t s) {
  int nc = 0;
  const float* dev = eval_device(q, s, &nc);
  for (int c0 = 0; c0 < nc; c0 += 3) {
    const int cnt = std::min(3, nc - c0);
    world->reducer.component_sums(dev + (size_t)c0 * N, cnt, N, geom.get(), out + c0, s);
  }
  for (int c = 0; c < nc; ++c) out[c] /= (cellsingeo > 0 ? cellsingeo : 1);
}

// ------------------------------------------------------------ equations

std::vector<StateVar> World::state_vars() {
  std::vector<StateVar> v;
  auto add_mag = [&](Ferromagnet* fm) {
    v.push_back(
        StateVar{StateVar::MAG, fm, nullptr, fm->m.get(), &fm->m, &fm->x0, &fm->k, fm->N, fm->geom_ptr(), true});
  };
  if (relaxing_) {
    for (Ferromagnet* fm : relax_fms_) add_mag(fm);
  } else {
    for (auto& fm : magnets) add_mag(fm.get());
    for (auto& a : afms) {
      add_mag(a->sub[0].get());
      add_mag(a->sub[1].get());
    }
  }
  if (relaxing_ && !relax_elastic_) return v;
  auto add_elastic = [&](MagnetCore* own) {
    if (!own->el.enabled) return;
    v.push_back(StateVar{StateVar::DISP, nullptr, own, own->el.u.get(), &own->el.u, &own->el.u0, &own->el.ku, own->N,
                         own->geom_ptr(), false});
    v.push_back(StateVar{StateVar::VEL, nullptr, own, own->el.v.get(), &own->el.v, &own->el.v0, &own->el.kv, own->N,
                         own->geom_ptr(), false});
  };
  for (auto& fm : magnets) add_elastic(fm.get());
  for (auto& a : afms) add_elastic(a.get());
  return v;
}

// All right-hand sides at the current (stage) state: strain of every magnet
// with elastodynamics and the stray field of every antiferromagnet once,
// then the torques, velocities and accelerations.
void World::eval_rhs_all(const std::vector<StateVar>& vars, int stage, hipStream_t s, bool disp_ready,
                         double noise_dt, const NextStage* next, const StageIn* stage_in, const PairNext* pair_next) {
  if (next && (vars.size() != 1 || vars[0].kind != StateVar::MAG || relaxing_))
    throw Error(MPA_ERR_RUNTIME, "a fused stage combination needs a single magnetisation variable");
  if (stage_in && (next || vars.size() != 1 || vars[0].kind != StateVar::MAG || vars[0].fm->host ||
                   !(vars[0].fm->active_terms() & TERM_DEMAG) || vars[0].x != vars[0].fm->m.get()))
    throw Error(MPA_ERR_RUNTIME, "a stage state formed in the demag pass needs one magnetisation with demag");
  // The strains and accelerations depend on the displacements, velocities
  // and magnetisations of the stage state only, so with elastodynamics they
  // run on stream2 beside the demag passes: the field kernels (the
  // magnetoelastic field reads the strain) wait for the strains, and the
  // evaluation ends when both streams are done.
  // Likewise the sublattice fields without demag of an antiferromagnet
  // (after the strains, on stream2), whose torques a launch after the demag
  // passes completes; the accelerations then go to stream3.
  bool elastic = false;
  for (auto& fm : magnets) elastic = elastic || fm->el.enabled;
  for (auto& a : afms) elastic = elastic || a->el.enabled;
  const FieldMode fmode = relaxing_ ? MODE_DAMPING : MODE_LLG;
  std::vector<Antiferromagnet*> dsplit;
  if (split_demag && !next && !stage_in)
    for (auto& a : afms)
      if (a->pair_split(fmode, noise_dt)) dsplit.push_back(a.get());
  auto in_split = [&](const Antiferromagnet* a) {
    return std::find(dsplit.begin(), dsplit.end(), a) != dsplit.end();
  };
  const bool split = elastic && split_elastic;
  const bool fork = split || !dsplit.empty();
  // the accelerations follow the strains on stream2, or go to stream3 when
  // stream2 also computes the sublattice fields (a third stream costs more
  // than it overlaps otherwise: 0.63 against 0.49 ms per 512 x 512 step)
  const bool three = split && !dsplit.empty();
  hipStream_t se = fork ? stream2 : s;                        // the strains
  hipStream_t sa = split ? (three ? stream3 : stream2) : s;  // the accelerations
  if (fork) {
    HIP_CHECK(hipEventRecord(ev_fork, s));
    HIP_CHECK(hipStreamWaitEvent(stream2, ev_fork, 0));
    if (three) HIP_CHECK(hipStreamWaitEvent(stream3, ev_fork, 0));
  }
  for (auto& fm : magnets)
    if (fm->el.enabled) fm->compute_strain(se);
  for (auto& a : afms)
    if (a->el.enabled) a->compute_strain(se);
  if (fork) HIP_CHECK(hipEventRecord(ev_strain, stream2));
  if (three) HIP_CHECK(hipStreamWaitEvent(stream3, ev_strain, 0));
  for (Antiferromagnet* a : dsplit) a->eval_local_pair(stream2, noise_dt);
  if (!dsplit.empty()) HIP_CHECK(hipEventRecord(ev_local, stream2));
  for (auto& fm : magnets) fm->compute_stray(s, &prof);
  for (auto& a : afms) {
    if ((a->sub[0]->active_terms() | a->sub[1]->active_terms()) & TERM_DEMAG) a->compute_demag(s, &prof);
    a->compute_stray(s, &prof);
  }
  if (fork) HIP_CHECK(hipStreamWaitEvent(s, dsplit.empty() ? ev_strain : ev_local, 0));
  for (size_t vi = 0; vi < vars.size(); ++vi) {
    const StateVar& sv = vars[vi];
    float* out = (*sv.k)[stage].get();
    if (sv.kind == StateVar::MAG) {
      Ferromagnet* fm = sv.fm;
      const FieldMode mode = relaxing_ ? MODE_DAMPING : MODE_LLG;
      if (fm->host) {
        // the two sublattices follow each other: one launch for both
        if (fm == fm->host->sub[0].get() && vi + 1 < vars.size() && vars[vi + 1].kind == StateVar::MAG &&
            vars[vi + 1].fm == fm->host->sub[1].get()) {
          if (in_split(fm->host)) fm->host->eval_torque_pair(out, (*vars[vi + 1].k)[stage].get(), s, mode, noise_dt);
          else if (pair_next && pair_next->afm == fm->host)
            fm->host->eval_rhs_pair_next(out, (*vars[vi + 1].k)[stage].get(), s, pair_next->m[0], pair_next->m[1]);
          else fm->host->eval_rhs_pair(out, (*vars[vi + 1].k)[stage].get(), s, mode, noise_dt);
          ++vi;
          continue;
        }
        fm->eval_rhs(out, s, true, mode, noise_dt);
      } else {
        const unsigned t = fm->active_terms();
        if (stage_in)
          fm->demag().exec_combine(stage_in->x0, stage_in->co, fm->m.get(), fm->param("msat").ps(), fm->geom_ptr(),
                                   fm->hdemag.get(), s, &prof);
        else if (t & TERM_DEMAG)
          fm->demag().exec(fm->m.get(), fm->param("msat").ps(), fm->geom_ptr(), fm->hdemag.get(), s, &prof);
        fm->eval_rhs(out, s, true, mode, noise_dt, next);
      }
    } else if (sv.kind == StateVar::DISP) {
      // du/dt = v (elastodynamics.cu:93-96); a stage's combination may have
      // written it already
      if (disp_ready) continue;
      Timed t(&prof, "copy_v", 24.0 * sv.N, s);
      HIP_CHECK(hipMemcpyAsync(out, sv.owner->el.v.get(), (size_t)3 * sv.N * sizeof(float),
                               hipMemcpyDeviceToDevice, s));
    } else {
      MagnetCore* own = sv.owner;
      if (pair_next && pair_next->has_el && pair_next->afm == own) {
        // plus the (u, v) combinations: read u0, v0, nk accelerations and nk + 1
        // velocities; write u, v and the next stage velocity
        const int nk = pair_next->el.nk;
        Timed t(&prof, "k_elastic_next", (48.0 + 12.0 * own->num_fm() + 12.0 * (2 * nk + 6)) * sv.N, sa);
        launch_elastic_next(own->eldesc(), own->el.stress.get(), own->el.v.get(), own->me_force_of(0),
                            own->num_fm() > 1 ? own->me_force_of(1) : MEForce{}, out, pair_next->el, sa);
        continue;
      }
      // read stress (24 B), v (12 B) and each coupled magnetisation (12 B); write a (12 B)
      Timed t(&prof, "k_elastic", (48.0 + 12.0 * own->num_fm()) * sv.N, sa);
      launch_elastic(own->eldesc(), own->el.stress.get(), own->el.v.get(), own->me_force_of(0),
                     own->num_fm() > 1 ? own->me_force_of(1) : MEForce{}, out, EL_ACCEL, sa);
    }
  }
  if (fork) {
    HIP_CHECK(hipEventRecord(ev_join, stream2));
    HIP_CHECK(hipStreamWaitEvent(s, ev_join, 0));
  }
  if (three) {
    HIP_CHECK(hipEventRecord(ev_join3, stream3));
    HIP_CHECK(hipStreamWaitEvent(s, ev_join3, 0));
  }
}

// ------------------------------------------------------ relax / minimize

double World::total_energy(MagnetCore* m) {
  if (auto* fm = dynamic_cast<Ferromagnet*>(m)) return fm->scalar_quantity("total_energy", stream);
  return static_cast<Antiferromagnet*>(m)->scalar_quantity("total_energy", stream);
}

// Relaxer::exec (relaxer.cpp:87-174): damping-only dynamics with Fehlberg
// and an adaptive step, first while the total energy decreases (checked
// every 3 steps), then with the error tolerance lowered by sqrt(2) at a time
// down to `tol`, each level run while the summed squared torque decreases (or,
// with positive thresholds, until every max torque is below its threshold).
// The solver's method, error, step mode, time and step are restored; with an
// adaptive step the step is the sensible one of the restored equations.
void World::relax(MagnetCore* target, double tol) {
  std::vector<MagnetCore*> targets;
  if (target) {
    targets.push_back(target);
  } else {
    for (auto& m : magnets) targets.push_back(m.get());
    for (auto& a : afms) targets.push_back(a.get());
  }
  std::vector<Ferromagnet*> fms;
  std::vector<double> thr;
  for (MagnetCore* t : targets) {
    if (auto* fm = dynamic_cast<Ferromagnet*>(t)) {
      fms.push_back(fm);
      thr.push_back(target ? fm->relax_threshold : relax_threshold);
    } else {
      auto* a = static_cast<Antiferromagnet*>(t);
      double t1 = a->sub[0]->relax_threshold, t2 = a->sub[1]->relax_threshold;
      if (!target) t1 = t2 = relax_threshold;
      // antiferromagnet.cpp:57-68: one user-set threshold serves both sublattices
      if (t1 > 0 && t2 <= 0) t2 = t1;
      else if (t1 <= 0 && t2 > 0) t1 = t2;
      fms.push_back(a->sub[0].get());
      fms.push_back(a->sub[1].get());
      thr.push_back(t1);
      thr.push_back(t2);
    }
  }
  TimeSolver& ts = solver;
  const double time0 = ts.time, dt0 = ts.dt, maxerr0 = ts.max_error;
  const bool adaptive0 = ts.adaptive;
  const std::string method0 = ts.method;
  auto restore = [&]() {
    relaxing_ = false;
    relax_elastic_ = false;
    relax_fms_.clear();
    ts.method = method0;
    ts.max_error = maxerr0;
    ts.adaptive = adaptive0;
    ts.time = time0;
    ts.dt = dt0;
    ts.equations_changed();  // FSAL state dropped; sensible step when adaptive
  };
  try {
    ts.adaptive = true;
    ts.method = "Fehlberg";
    relaxing_ = true;
    relax_fms_ = fms;
    relax_elastic_ = target == nullptr;  // a world relax keeps the elastic equations
    ts.equations_changed();
    constexpr int NS = 3;  // energy every 3 steps
    auto energy = [&]() {
      double e = 0.0;
      for (MagnetCore* t : targets) e += total_energy(t);
      return e;
    };
    auto torque_of = [&](Ferromagnet* fm) {
      int nc = 0;
      return fm->eval_device("damping_torque", stream, &nc);
    };
    auto torque_sum = [&]() {
      double t = 0.0;
      for (Ferromagnet* fm : fms) t += reducer.dot_sum(torque_of(fm), fm->N, fm->geom_ptr(), stream);
      return t;
    };
    double E0 = energy();
    ts.steps(NS);
    double E1 = energy();
    while (E1 < E0) {
      ts.steps(NS);
      E0 = E1;
      E1 = energy();
    }
    const bool all_negative = std::all_of(thr.begin(), thr.end(), [](double t) { return t < 0; });
    if (all_negative) {
      double t1 = torque_sum(), t0 = 0.0;
      double err = ts.max_error;
      while (err > tol) {
        err /= std::sqrt(2.0);
        ts.max_error = err;
        ts.steps(NS);
        t0 = t1;
        t1 = torque_sum();
        while (t1 < t0) {
          ts.steps(NS);
          t0 = t1;
          t1 = torque_sum();
        }
      }
    } else if (std::find(thr.begin(), thr.end(), 0.0) != thr.end()) {
      throw Error(MPA_ERR_INVALID, "The relax threshold should not be zero.");
    } else {
      double err = ts.max_error;
      while (err > tol) {
        bool converged = true;
        for (size_t i = 0; i < fms.size() && converged; ++i)
          if (reducer.max_vec_norm(torque_of(fms[i]), fms[i]->N, fms[i]->geom_ptr(), stream) > thr[i])
            converged = false;
        if (converged) {
          err /= std::sqrt(2.0);
          ts.max_error = err;
        }
        ts.steps(NS);
      }
    }
  } catch (...) {
    restore();
    throw;
  }
  restore();
}

// Minimizer (minimizer.cu:68-160): steepest descent along the damping-only
// torque with Barzilai-Borwein step sizes (starting at 1e-14), all magnets
// stepped, then normalised, then their torques re-evaluated; converged when
// the last nsamples (times the number of ferromagnets for a world) maximum
// magnetisation changes are all <= tol.
void World::minimize(MagnetCore* target, double tol, int nsamples) {
  if (nsamples < 1) throw Error(MPA_ERR_INVALID, "nsamples should be positive");
  std::vector<Ferromagnet*> fms;
  auto add = [&](MagnetCore* t) {
    if (auto* fm = dynamic_cast<Ferromagnet*>(t)) {
      fms.push_back(fm);
    } else {
      auto* a = static_cast<Antiferromagnet*>(t);
      fms.push_back(a->sub[0].get());
      fms.push_back(a->sub[1].get());
    }
  };
  size_t nsamp = (size_t)nsamples;
  if (target) {
    add(target);
  } else {
    for (auto& m : magnets) add(m.get());
    for (auto& a : afms) add(a.get());
    nsamp *= fms.size();
  }
  const size_t F = fms.size();
  std::vector<DeviceArray<float>> m0(F), m1(F), t0(F), t1(F);
  for (size_t i = 0; i < F; ++i) {
    const size_t n3 = (size_t)3 * fms[i]->N;
    m0[i].allocate(n3);
    m1[i].allocate(n3);
    t0[i].allocate(n3);
    t1[i].allocate(n3);
  }
  std::vector<double> h(F, 1e-14);
  std::deque<float> last;
  auto torque_into = [&](size_t i, DeviceArray<float>& dst) {
    int nc = 0;
    const float* t = fms[i]->eval_device("damping_torque", stream, &nc);
    HIP_CHECK(hipMemcpyAsync(dst.get(), t, dst.bytes(), hipMemcpyDeviceToDevice, stream));
  };
  long long nsteps = 0;
  auto converged = [&]() {
    if (last.size() < nsamp) return false;
    for (float d : last)
      if (d > tol) return false;
    return true;
  };
  while (!converged()) {
    for (size_t i = 0; i < F; ++i) {
      Ferromagnet* fm = fms[i];
      HIP_CHECK(hipMemcpyAsync(m0[i].get(), fm->m.get(), m0[i].bytes(), hipMemcpyDeviceToDevice, stream));
      if (nsteps == 0) torque_into(i, t0[i]);
      else std::swap(t0[i], t1[i]);
      launch_descent(m1[i].get(), m0[i].get(), t0[i].get(), (float)h[i], fm->N, stream);
    }
    for (size_t i = 0; i < F; ++i) {
      // NormalizedVariable::set: normalised inside the geometry, zero outside
      Combo co{};
      co.nk = 0;
      launch_combine(fms[i]->m.get(), m1[i].get(), co, 3, fms[i]->N, fms[i]->geom_ptr(), true, stream);
    }
    touch();
    for (size_t i = 0; i < F; ++i) torque_into(i, t1[i]);
    for (size_t i = 0; i < F; ++i) {
      double d[3];
      float maxdm = 0.f;
      reducer.bb_sums(m1[i].get(), m0[i].get(), t1[i].get(), t0[i].get(), fms[i]->N, fms[i]->geom_ptr(), d, &maxdm,
                      stream);
      // BarzilianBorweinStepSize (minimizer.cu:76-90): alternating forms
      const double nom = nsteps % 2 == 0 ? d[0] : d[1];
      const double div = nsteps % 2 == 0 ? d[1] : d[2];
      h[i] = div == 0.0 ? 1e-14 : (double)(float)(nom / div);
      last.push_back(maxdm);
      if (last.size() > nsamp) last.pop_front();
    }
    ++nsteps;
  }
  HIP_CHECK(hipStreamSynchronize(stream));
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

// TimeSolver::sensibleTimeStep (src/core/timesolver.cpp:37-47).  The
// reference re-estimates dt whenever a magnet is added (timesolver.cpp:49-52),
// before the caller can switch stray coupling off; the estimate runs without
// the refusal, the steps keep it.
double TimeSolver::sensible_timestep() {
  std::vector<StateVar> vars = world_->state_vars();
  if (vars.empty()) return 0.0;
  for (auto& sv : vars)
    while ((int)sv.k->size() < 1) sv.k->emplace_back((size_t)3 * sv.N);
  world_->eval_rhs_all(vars, 0, world_->stream);
  nrhs += (long long)vars.size();
  double mx = 0.0;
  for (auto& sv : vars)
    mx = std::max(mx, (double)world_->reducer.max_vec_norm((*sv.k)[0].get(), sv.N, sv.geom, world_->stream));
  fsal_valid_ = false;  // k[0] now holds the rhs at the current state, not the FSAL value
  if (mx == 0.0) return sensible_default;
  return sensible_factor / mx;
}

void TimeSolver::adapt(double corr) {
  if (!adaptive) return;
  if (std::isnan(corr)) corr = 1.0;
  corr *= headroom;
  if (lower_bound >= upper_bound) throw Error(MPA_ERR_RUNTIME, "The lower bound should be lower than the upper bound.");
  corr = corr > upper_bound ? upper_bound : corr;
  corr = corr < lower_bound ? lower_bound : corr;
  dt *= corr;
}

// RungeKuttaStepper::step (src/core/rungekutta.cpp:25-84).  Every stage first
// sets the stage state of all variables (x0 + dt * sum_j a_sj k_j,
// normalised for magnetisations), then evaluates all right-hand sides; the
// final state uses weights1; the error is the largest max-norm of
// dt * sum_j (w1_j - w2_j) k_j over the variables.  For tableaus whose last
// stage is the final state (DormandPrince, BogackiShampine) the last stage's
// k is the next step's first k (first-same-as-last), and the final
// combination is the last stage state itself.
void TimeSolver::step() {
  if (dt <= 0)
    throw Error(MPA_ERR_RUNTIME,
                "Timesolver can not make a step because the timestep is smaller than or equal to zero.");
  std::vector<StateVar> vars = world_->state_vars();
  if (vars.empty()) {
    time += dt;
    return;
  }
  hipStream_t s = world_->stream;
  const Tableau& tb = tableau_by_name(method);
  const int S = (int)tb.nodes.size();
  Profiler* prof = &world_->prof;
  const double t0 = time;
  // thermal noise: one draw per step for every magnetisation with a
  // temperature, kept through rejected attempts and all stages (the stage
  // executor draws it once, rungekutta.cpp:85-93); its stage derivatives
  // hold this step's noise, so the last one is not the next step's first
  bool noisy = false;
  for (auto& sv : vars)
    if (sv.kind == StateVar::MAG && (sv.fm->active_terms() & TERM_THERMAL)) {
      sv.fm->step_draw = sv.fm->noise_draws[0]++;
      noisy = true;
    }
  bool k0_valid = !noisy && fsal_valid_ && fsal_version_ == world_->version && fsal_time_ == t0;
  // One independent magnet on the uniform-parameter LLG path with a
  // first-same-as-last tableau: each evaluation also writes the next stage
  // state (the combination's last derivative is the one just computed), so
  // no separate combination launch runs.  With a fixed step the step's last
  // evaluation writes the next step's first stage state as well.
  Ferromagnet* fused = nullptr;
  if (fuse_stages && !noisy && tb.fsal && S > 1 && vars.size() == 1 && vars[0].kind == StateVar::MAG && !world_->relaxing() &&
      vars[0].fm->next_fusable())
    fused = vars[0].fm;
  // One independent magnetisation with demag: each stage state is formed in
  // the demag's first FFT pass, which reads x0 and the derivatives instead of
  // the state (no combination launch).
  const bool xfuse = !fused && fuse_combine && vars.size() == 1 && vars[0].kind == StateVar::MAG &&
                     !vars[0].fm->host && vars[0].normalize && (vars[0].fm->active_terms() & TERM_DEMAG);
  const bool next_kept = k0_valid && next_valid_ && next_h_ == dt;
  bool next_ready = fused && next_kept;
  next_valid_ = false;
  if (fused && fused->xn.n == 0) fused->xn.allocate((size_t)3 * fused->N);
  // One antiferromagnet on the sublattice-pair kernel (with its displacement
  // and velocity when it has elastodynamics) and nothing else: each
  // evaluation but the last also writes the next stage state of every
  // variable, the sublattices' in the field launch and (u, v) in the
  // acceleration launch, with the combination's arithmetic.
  Antiferromagnet* afused = nullptr;
  if (fuse_stages && !fused && !xfuse && !noisy && S > 1 && !world_->relaxing() && world_->magnets.empty() &&
      world_->afms.size() == 1 && !world_->split_elastic && !world_->split_demag) {
    Antiferromagnet* a = world_->afms[0].get();
    const bool el = a->el.enabled;
    const bool layout = vars.size() == (el ? 4u : 2u) && vars[0].fm == a->sub[0].get() &&
                        vars[1].fm == a->sub[1].get() &&
                        (!el || (vars[2].kind == StateVar::DISP && vars[3].kind == StateVar::VEL &&
                                 vars[2].owner == a && vars[3].owner == a));
    if (layout && a->next_fusable()) afused = a;
  }
  // with a fixed step the previous step's last evaluation formed this step's
  // first stage state (xn, un, vn and the displacement's k[1])
  bool pair_ready = afused && next_kept;
  if (afused) {
    for (auto& f : afused->sub)
      if (f->xn.n == 0) f->xn.allocate((size_t)3 * f->N);
    if (afused->el.enabled && afused->el.un.n == 0) {
      afused->el.un.allocate((size_t)3 * afused->N);
      afused->el.vn.allocate((size_t)3 * afused->N);
    }
  }
  // x0 = the state at the step start.  When the first stage's derivative is
  // reused, nothing reads the state before the second stage's combination
  // overwrites it, so the state array and x0 trade buffers instead of copying.
  for (auto& sv : vars) {
    if (sv.x0->n == 0) sv.x0->allocate((size_t)3 * sv.N);
    while ((int)sv.k->size() < S) sv.k->emplace_back((size_t)3 * sv.N);
    if (k0_valid && S > 1) {
      std::swap(*sv.xa, *sv.x0);
      sv.x = sv.xa->get();
    } else {
      Timed t(prof, "copy_x0", 24.0 * sv.N, s);
      HIP_CHECK(hipMemcpyAsync(sv.x0->get(), sv.x, (size_t)3 * sv.N * sizeof(float), hipMemcpyDeviceToDevice, s));
    }
  }
  // The displacement's derivative at a stage is that stage's velocity: the
  // velocity's combination writes it into the displacement's k as well
  // (du/dt = v, elastodynamics.cu:93-96), so no copy is needed.
  std::vector<int> disp_of(vars.size(), -1);
  for (size_t i = 0; i < vars.size(); ++i)
    if (vars[i].kind == StateVar::VEL)
      for (size_t j = 0; j < vars.size(); ++j)
        if (vars[j].kind == StateVar::DISP && vars[j].owner == vars[i].owner) disp_of[i] = (int)j;
  // All variables' combinations in one launch per batch of MAX_JOBS;
  // `stage` >= 0 names the k slot the velocity also fills.
  auto combine = [&](const std::vector<double>& w, int nk, double h, int stage) {
    double bytes = 0.0;
    std::vector<CombineJob> jobs;
    for (size_t i = 0; i < vars.size(); ++i) {
      const StateVar& sv = vars[i];
      CombineJob jb{};
      jb.out = sv.x;
      jb.x0 = sv.x0->get();
      jb.geom = sv.geom;
      jb.N = sv.N;
      jb.normalize = sv.normalize ? 1 : 0;
      jb.co.nk = nk;
      for (int j = 0; j < nk; ++j) {
        jb.co.k[j] = (*sv.k)[j].get();
        jb.co.c[j] = (float)(h * w[j]);
      }
      bytes += 12.0 * sv.N * (nk + 2);
      if (stage >= 0 && disp_of[i] >= 0) {
        jb.out2 = (*vars[disp_of[i]].k)[stage].get();
        bytes += 12.0 * sv.N;
      }
      jobs.push_back(jb);
    }
    for (size_t j0 = 0; j0 < jobs.size(); j0 += MAX_JOBS) {
      const int n = (int)std::min<size_t>(MAX_JOBS, jobs.size() - j0);
      Timed t(prof, "k_combine", bytes * n / jobs.size(), s);
      launch_combine_batch(jobs.data() + j0, n, s);
    }
  };
  while (true) {
    const double h = dt;
    bool have_next = next_ready;  // the magnet's xn holds the coming stage's state
    next_ready = false;
    bool have_pair_next = pair_ready;  // the antiferromagnet's xn, un, vn hold it
    pair_ready = false;
    for (int st = 0; st < S; ++st) {
      time = t0 + h * tb.nodes[st];
      World::StageIn sin{};
      const World::StageIn* stage_in = nullptr;
      if (st > 0) {
        if (have_next) {
          std::swap(fused->m, fused->xn);
          vars[0].x = fused->m.get();
          have_next = false;
        } else if (have_pair_next) {
          for (auto& f : afused->sub) std::swap(f->m, f->xn);
          if (afused->el.enabled) {
            std::swap(afused->el.u, afused->el.un);
            std::swap(afused->el.v, afused->el.vn);
          }
          for (auto& sv : vars) sv.x = sv.xa->get();
          have_pair_next = false;
        } else if (xfuse) {
          sin.x0 = vars[0].x0->get();
          sin.co.nk = st;
          for (int j = 0; j < st; ++j) {
            sin.co.k[j] = (*vars[0].k)[j].get();
            sin.co.c[j] = (float)(h * tb.A[st][j]);
          }
          stage_in = &sin;
        } else {
          combine(tb.A[st], st, h, st);
        }
      }
      if (st == 0 && k0_valid) continue;
      const bool cross = st + 1 == S;  // the next step's first stage
      if (fused && !(cross && adaptive)) {
        NextStage ns{};
        const std::vector<double>& a = tb.A[cross ? 1 : st + 1];
        ns.x0 = cross ? nullptr : vars[0].x0->get();
        ns.nk = cross ? 0 : st;
        for (int j = 0; j < ns.nk; ++j) {
          ns.k[j] = (*vars[0].k)[j].get();
          ns.c[j] = (float)(h * a[j]);
        }
        ns.cself = (float)(h * a[cross ? 0 : st]);
        ns.out = fused->xn.get();
        ns.N = fused->N;
        world_->eval_rhs_all(vars, st, s, st > 0, 0.0, &ns);
        if (cross) {
          next_valid_ = true;
          next_h_ = h;
        } else {
          have_next = true;
        }
      } else if (afused && (st + 1 < S || (st + 1 == S && !adaptive && tb.fsal))) {
        // st + 1 == S: the next step's first stage state, from this step's
        // final state (the stage state itself) and its derivative, for the
        // same fixed step
        const bool cross = st + 1 == S;
        World::PairNext pn{};
        pn.afm = afused;
        const std::vector<double>& a = tb.A[cross ? 1 : st + 1];
        const int nk = cross ? 0 : st;
        const float cself = (float)(h * a[cross ? 0 : st]);
        for (int i = 0; i < 2; ++i) {
          NextStage& ns = pn.m[i];
          ns.x0 = cross ? nullptr : vars[i].x0->get();
          ns.nk = nk;
          for (int j = 0; j < nk; ++j) {
            ns.k[j] = (*vars[i].k)[j].get();
            ns.c[j] = (float)(h * a[j]);
          }
          ns.cself = cself;
          ns.out = afused->sub[i]->xn.get();
          ns.N = vars[i].N;
        }
        if (afused->el.enabled) {
          pn.has_el = true;
          ElNext& e = pn.el;
          const StateVar &du = vars[2], &dv = vars[3];
          // across steps: x0 = the current state; the displacement's first
          // derivative is the current velocity
          e.u0 = cross ? du.x : du.x0->get();
          e.v0 = cross ? dv.x : dv.x0->get();
          e.nk = nk;
          for (int j = 0; j < nk; ++j) {
            e.ka[j] = (*dv.k)[j].get();
            e.c[j] = (float)(h * a[j]);
          }
          for (int j = 0; j <= nk; ++j) e.ku[j] = (*du.k)[j].get();
          if (cross) e.ku[0] = dv.x;
          e.cself = cself;
          e.uout = afused->el.un.get();
          e.vout = afused->el.vn.get();
          e.kdisp = (*du.k)[cross ? 1 : st + 1].get();
        }
        world_->eval_rhs_all(vars, st, s, st > 0, 0.0, nullptr, nullptr, &pn);
        if (cross) {
          next_valid_ = true;
          next_h_ = h;
        } else {
          have_pair_next = true;
        }
      } else {
        world_->eval_rhs_all(vars, st, s, st > 0, noisy ? h : 0.0, nullptr, stage_in);
      }
      nrhs += (long long)vars.size();
    }
    time = t0 + h;
    if (!tb.fsal) combine(tb.w1, S, h, -1);
    ++nsteps;
    if (!adaptive) break;
    double err = 0.0;
    for (auto& sv : vars) {
      Combo co{};
      co.nk = S;
      for (int j = 0; j < S; ++j) {
        co.k[j] = (*sv.k)[j].get();
        co.c[j] = (float)(h * (tb.w1[j] - tb.w2[j]));
      }
      Timed t(prof, "k_combo_maxnorm", 12.0 * sv.N * S, s);
      err = std::max(err, (double)world_->reducer.max_combo_norm(co, sv.N, sv.geom, s));
    }
    const bool success = err < max_error;
    const double ratio = max_error / err;  // err == 0 gives +inf, clamped by adapt()
    const double corr = success ? std::pow(ratio, 1.0 / (tb.order2 + 1)) : std::pow(ratio, 1.0 / (tb.order1 + 1));
    adapt(corr);
    if (success) break;
    ++nrejected;
    for (auto& sv : vars)
      HIP_CHECK(hipMemcpyAsync(sv.x, sv.x0->get(), (size_t)3 * sv.N * sizeof(float), hipMemcpyDeviceToDevice, s));
    time = t0;
    // k[0] = rhs(x0, t0) was computed in this attempt; with noise it holds
    // noise / sqrt(dt) of the rejected dt and is computed again
    k0_valid = !noisy;
  }
  if (tb.fsal && !noisy) {
    for (auto& sv : vars) std::swap((*sv.k)[0], (*sv.k)[S - 1]);
    fsal_valid_ = true;
    fsal_time_ = time;
    fsal_version_ = world_->version;
  } else {
    fsal_valid_ = false;
  }
}

void TimeSolver::steps(long long nstep) {
  for (long long i = 0; i < nstep; ++i) step();
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
  HIP_CHECK(hipStreamCreateWithFlags(&stream2, hipStreamNonBlocking));
  HIP_CHECK(hipStreamCreateWithFlags(&stream3, hipStreamNonBlocking));
  for (hipEvent_t* e : {&ev_fork, &ev_strain, &ev_local, &ev_join, &ev_join3})
    HIP_CHECK(hipEventCreateWithFlags(e, hipEventDisableTiming));
}

World::~World() {
  magnets.clear();
  afms.clear();
  for (hipEvent_t e : {ev_fork, ev_strain, ev_local, ev_join, ev_join3})
    if (e) (void)hipEventDestroy(e);
  if (stream3) (void)hipStreamDestroy(stream3);
  if (stream2) (void)hipStreamDestroy(stream2);
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

// MumaxWorld::checkAddibility / checkAllMagnetsInMastergrid (mumaxworld.cpp:35-46, :228-236):
// std::out_of_range in the reference, an index error at this boundary.
void World::check_in_mastergrid(const int size[3], const int org[3], const int mg_size[3], const int mg_origin[3],
                                bool adding) const {
  for (int a = 0; a < 3; ++a) {
    if (mg_size[a] == 0) continue;
    if (org[a] < mg_origin[a] || org[a] + size[a] > mg_origin[a] + mg_size[a])
      throw Error(MPA_ERR_INDEX, adding ? "Can not add magnet because the grid does not fit in the mastergrid"
                                        : "Not all magnets of the world fit inside the mastergrid.");
  }
}

void World::check_no_overlap(const int size[3], const int org[3]) const {
  auto overlaps = [&](const MagnetCore& m) {
    for (int a = 0; a < 3; ++a)
      if (org[a] >= m.origin[a] + m.n[a] || m.origin[a] >= org[a] + size[a]) return false;
    return true;
  };
  for (auto& m : magnets)
    if (overlaps(*m)) throw Error(MPA_ERR_INDEX, "Can not add magnet because it overlaps with another magnet.");
  for (auto& a : afms)
    if (overlaps(*a)) throw Error(MPA_ERR_INDEX, "Can not add magnet because it overlaps with another magnet.");
}

void World::check_name(const std::string& name) {
  if (find(name) || find_afm(name)) throw Error(MPA_ERR_RUNTIME, "A magnet with the name '" + name + "' already exists");
}

Ferromagnet* World::add_ferromagnet(const int size[3], const int org[3], const uint8_t* geometry,
                                    const uint32_t* regions, const char* nm) {
  static int unnamed = 1;  // ferromagnet_<n> naming of mumaxworld.cpp:61-64
  std::string name = nm ? nm : "";
  if (name.empty()) name = "ferromagnet_" + std::to_string(unnamed++);
  check_in_mastergrid(size, org, mastergrid_size, mastergrid_origin, true);
  check_no_overlap(size, org);
  check_name(name);
  magnets.push_back(std::make_unique<Ferromagnet>(this, size, org, geometry, name));
  magnets.back()->set_regions(regions);
  touch();
  solver.equations_changed();
  return magnets.back().get();
}

Antiferromagnet* World::add_antiferromagnet(const int size[3], const int org[3], const uint8_t* geometry,
                                            const uint32_t* regions, const char* nm) {
  static int unnamed = 1;  // antiferromagnet_<n> naming of mumaxworld.cpp:86-89
  std::string name = nm ? nm : "";
  if (name.empty()) name = "antiferromagnet_" + std::to_string(unnamed++);
  check_in_mastergrid(size, org, mastergrid_size, mastergrid_origin, true);
  check_no_overlap(size, org);
  check_name(name);
  afms.push_back(std::make_unique<Antiferromagnet>(this, size, org, geometry, name));
  afms.back()->set_regions(regions);
  touch();
  solver.equations_changed();
  return afms.back().get();
}

Ferromagnet* World::find(const std::string& name) {
  for (auto& m : magnets)
    if (m->name == name) return m.get();
  return nullptr;
}

Antiferromagnet* World::find_afm(const std::string& name) {
  for (auto& a : afms)
    if (a->name == name) return a.get();
  return nullptr;
}

std::vector<Ferromagnet*> World::all_ferromagnets() {
  std::vector<Ferromagnet*> v;
  for (auto& m : magnets) v.push_back(m.get());
  for (auto& a : afms) {
    v.push_back(a->sub[0].get());
    v.push_back(a->sub[1].get());
  }
  return v;
}

void World::bounding_grid(int size[3], int org[3], bool require_magnets) const {
  std::vector<const MagnetCore*> all;
  for (auto& m : magnets) all.push_back(m.get());
  for (auto& a : afms) all.push_back(a.get());
  if (all.empty() && require_magnets)  // MumaxWorld::boundingGrid (mumaxworld.cpp:264-267)
    throw Error(MPA_ERR_INDEX, "Cannot find minimum bounding box if there are no magnets in the world.");
  if (all.empty()) {
    for (int a = 0; a < 3; ++a) size[a] = org[a] = 0;
    return;
  }
  int lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = all[0]->origin[a];
    hi[a] = lo[a] + all[0]->n[a];
  }
  for (auto* m : all)
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
  for (auto& m : magnets) {
    m->invalidate_demag();
    m->stray_plans.clear();
  }
  for (auto& a : afms) {
    a->invalidate_demag();
    a->stray_plans.clear();
  }
  touch();
}

// MumaxWorld::setPBC(mastergrid, reps) (mumaxworld.cpp:311-320)
void World::set_pbc(const int mg_size[3], const int mg_origin[3], const int reps[3]) {
  check_pbc(mg_size, reps);
  for (auto& m : magnets) check_in_mastergrid(m->n, m->origin, mg_size, mg_origin, false);
  for (auto& a : afms) check_in_mastergrid(a->n, a->origin, mg_size, mg_origin, false);
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
  bounding_grid(size, org, true);
  for (int a = 0; a < 3; ++a)
    if (reps[a] == 0) size[a] = 0;
  set_pbc(size, org, reps);
}

void World::unset_pbc() {
  for (int a = 0; a < 3; ++a) mastergrid_size[a] = mastergrid_origin[a] = pbc[a] = 0;
  pbc_changed();
}

}  // namespace mpa
