// extern "C" boundary of the library; declarations in include/mumaxplus_amd.h.
#include "../../include/mumaxplus_amd.h"

#include <algorithm>
#include <climits>
#include <cstring>
#include <string>
#include <vector>

#include "batch.hpp"
#include "world.hpp"

using namespace mpa;

struct mpa_world {
  World w;
  mpa_world(const double cs[3], const int mgs[3], const int mgo[3], const int reps[3]) : w(cs, mgs, mgo, reps) {}
};
struct mpa_ferromagnet {};      // opaque: the pointer is the Ferromagnet itself
struct mpa_antiferromagnet {};  // opaque: the pointer is the Antiferromagnet itself
struct mpa_batch {
  Batch b;
  mpa_batch(long long cap, const std::string& method) : b(cap, method) {}
};
struct mpa_demag_plan {
  DemagPlan plan;
  explicit mpa_demag_plan(const DemagGeometry& g) : plan(g) {}
};

namespace {
thread_local std::string g_last_error;

template <typename F>
int guard(F&& f) {
  try {
    f();
    g_last_error.clear();
    return MPA_OK;
  } catch (const Error& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "out of host memory";
    return MPA_ERR_RUNTIME;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return MPA_ERR_RUNTIME;
  } catch (...) {
    g_last_error = "unknown error";
    return MPA_ERR_RUNTIME;
  }
}

void need(const void* p, const char* what) {
  if (!p) throw Error(MPA_ERR_INVALID, std::string(what) + " must not be NULL");
}

Ferromagnet* FM(mpa_ferromagnet* m) {
  need(m, "ferromagnet");
  return reinterpret_cast<Ferromagnet*>(m);
}
const Ferromagnet* FM(const mpa_ferromagnet* m) {
  need(m, "ferromagnet");
  return reinterpret_cast<const Ferromagnet*>(m);
}
mpa_ferromagnet* H(Ferromagnet* f) { return reinterpret_cast<mpa_ferromagnet*>(f); }

Antiferromagnet* AFM(mpa_antiferromagnet* a) {
  need(a, "antiferromagnet");
  return reinterpret_cast<Antiferromagnet*>(a);
}
const Antiferromagnet* AFM(const mpa_antiferromagnet* a) {
  need(a, "antiferromagnet");
  return reinterpret_cast<const Antiferromagnet*>(a);
}
mpa_antiferromagnet* HA(Antiferromagnet* a) { return reinterpret_cast<mpa_antiferromagnet*>(a); }

// Parameter access shared by ferromagnets and antiferromagnets.
void param_info(const MagnetCore* m, const char* key, int* ncomp, int* is_uniform, char* name, int namelen,
                char* unit, int unitlen);
void param_set_uniform(MagnetCore* m, const char* key, const double* value);
void param_set_cells(MagnetCore* m, const char* key, const float* values);
void param_get(const MagnetCore* m, const char* key, float* out);
void param_get_uniform(const MagnetCore* m, const char* key, double* out);
void param_add_time_term(MagnetCore* m, const char* key, TimeFn fn, void* ctx, const float* mask);
void param_remove_time_terms(MagnetCore* m, const char* key);
void param_is_dynamic(const MagnetCore* m, const char* key, int* out);
void magnet_regions(const MagnetCore* m, unsigned int* out);
void magnet_unique_regions(const MagnetCore* m, unsigned int* out, int cap, int* count);
void inter_info(const MagnetCore* m, const char* key, int* is_uniform, double* uniform_value, long long* size,
                char* name, int namelen, char* unit, int unitlen);
void inter_eval(const MagnetCore* m, const char* key, float* out);
void magnet_grid(const MagnetCore* m, int size[3], int origin[3]) {
  for (int a = 0; a < 3; ++a) {
    if (size) size[a] = m->n[a];
    if (origin) origin[a] = m->origin[a];
  }
}
void magnet_geometry(const MagnetCore* m, uint8_t* out) {
  need(out, "out");
  if (m->geom_host.empty()) std::memset(out, 1, m->N);
  else std::memcpy(out, m->geom_host.data(), m->N);
}

void copy_str(const std::string& s, char* buf, int buflen) {
  if (!buf || buflen <= 0) return;
  const size_t n = std::min<size_t>(s.size(), (size_t)buflen - 1);
  std::memcpy(buf, s.data(), n);
  buf[n] = 0;
}
void param_info(const MagnetCore* m, const char* key, int* ncomp, int* is_uniform, char* name, int namelen,
                char* unit, int unitlen) {
  need(key, "key");
  const Param& p = m->param(key);
  if (ncomp) *ncomp = p.ncomp;
  if (is_uniform) *is_uniform = p.is_uniform();
  copy_str(p.name, name, namelen);
  copy_str(p.unit, unit, unitlen);
}

void param_set_uniform(MagnetCore* m, const char* key, const double* value) {
  need(key, "key");
  need(value, "value");
  Param& p = m->param(key);
  p.cells.release();
  for (int c = 0; c < 3; ++c) p.uniform[c] = c < p.ncomp ? (float)value[c] : 0.f;
  p.changed();
  m->world->touch();
}

void param_set_cells(MagnetCore* m, const char* key, const float* values) {
  need(key, "key");
  need(values, "values");
  Param& p = m->param(key);
  const size_t n = (size_t)p.ncomp * m->N;
  if (p.cells.n != n) p.cells.allocate(n);
  HIP_CHECK(hipMemcpy(p.cells.get(), values, n * sizeof(float), hipMemcpyHostToDevice));
  for (int c = 0; c < 3; ++c) p.uniform[c] = 0.f;
  p.changed();
  m->world->touch();
}

void param_get(const MagnetCore* m, const char* key, float* out) {
  need(key, "key");
  need(out, "out");
  const Param& p = m->param(key);
  const size_t n = (size_t)p.ncomp * m->N;
  // the value at the solver's time, time-dependent terms included (parameter.cu:65-91)
  const PV v = p.pv();
  if (!v.p) {
    const float u[3] = {v.x, v.y, v.z};
    for (int c = 0; c < p.ncomp; ++c)
      for (int i = 0; i < m->N; ++i) out[(size_t)c * m->N + i] = u[c];
  } else {
    hipStream_t s = m->world->stream;
    HIP_CHECK(hipMemcpyAsync(out, v.p, n * sizeof(float), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
  }
}

// DynamicParameter::addTimeDependentTerm (dynamic_parameter.hpp:20-48)
void param_add_time_term(MagnetCore* m, const char* key, TimeFn fn, void* ctx, const float* mask) {
  need(key, "key");
  if (!fn) throw Error(MPA_ERR_INVALID, "the time-dependent term is a null function");
  Param& p = m->param(key);
  TimeTerm t{fn, ctx, DeviceArray<float>()};
  if (mask) {
    const size_t n = (size_t)p.ncomp * m->N;
    t.mask.allocate(n);
    HIP_CHECK(hipMemcpy(t.mask.get(), mask, n * sizeof(float), hipMemcpyHostToDevice));
  }
  p.terms.push_back(std::move(t));
  p.changed();
  m->world->touch();
}

void param_remove_time_terms(MagnetCore* m, const char* key) {
  need(key, "key");
  Param& p = m->param(key);
  HIP_CHECK(hipStreamSynchronize(m->world->stream));  // masks may still be read
  p.terms.clear();
  p.changed();
  m->world->touch();
}

void magnet_regions(const MagnetCore* m, unsigned int* out) {
  need(out, "out");
  if (m->regions_host.empty()) std::memset(out, 0, (size_t)m->N * sizeof(unsigned int));
  else std::memcpy(out, m->regions_host.data(), (size_t)m->N * sizeof(unsigned int));
}

void magnet_unique_regions(const MagnetCore* m, unsigned int* out, int cap, int* count) {
  need(count, "count");
  *count = (int)m->unique_regions.size();
  if (out)
    for (int i = 0; i < std::min(cap, *count); ++i) out[i] = m->unique_regions[i];
}

void inter_info(const MagnetCore* m, const char* key, int* is_uniform, double* uniform_value, long long* size,
                char* name, int namelen, char* unit, int unitlen) {
  need(key, "key");
  const InterParam& p = m->inter(key);
  if (is_uniform) *is_uniform = p.is_uniform() ? 1 : 0;
  if (uniform_value) *uniform_value = p.uniform;
  if (size) *size = (long long)p.limit;
  copy_str(p.name, name, namelen);
  copy_str(p.unit, unit, unitlen);
}

void inter_eval(const MagnetCore* m, const char* key, float* out) {
  need(key, "key");
  need(out, "out");
  const InterParam& p = m->inter(key);
  for (size_t k = 0; k < p.limit; ++k) out[k] = p.is_uniform() ? p.uniform : p.host[k];
}

void param_is_dynamic(const MagnetCore* m, const char* key, int* out) {
  need(key, "key");
  need(out, "out");
  *out = m->param(key).is_dynamic() ? 1 : 0;
}

void param_get_uniform(const MagnetCore* m, const char* key, double* out) {
  need(key, "key");
  need(out, "out");
  const Param& p = m->param(key);
  if (!p.is_uniform()) throw Error(MPA_ERR_RUNTIME, "parameter '" + std::string(key) + "' is not uniform");
  for (int c = 0; c < p.ncomp; ++c) out[c] = p.uniform[c];
}

// Flags every magnet has: stray-field coupling and elastodynamics.
bool magnet_set_flag(MagnetCore* m, const std::string& k, int value) {
  if (k == "enable_as_stray_field_source") m->stray_source = value != 0;
  else if (k == "enable_as_stray_field_destination") m->stray_destination = value != 0;
  else if (k == "enable_elastodynamics") {
    // Magnet::setEnableElastodynamics (magnet.cpp:177-213)
    if (value && m->rigid_strain_applied())
      throw Error(MPA_ERR_INVALID, "Cannot enable elastodynamics when rigid strain is set.");
    m->set_elastodynamics(value != 0);
  } else
    return false;
  m->world->touch();
  return true;
}

bool magnet_get_flag(const MagnetCore* m, const std::string& k, int* out) {
  if (k == "enable_as_stray_field_source") *out = m->stray_source;
  else if (k == "enable_as_stray_field_destination") *out = m->stray_destination;
  else if (k == "enable_elastodynamics") *out = m->el.enabled;
  else return false;
  return true;
}
}  // namespace

extern "C" {

const char* mpa_last_error(void) { return g_last_error.c_str(); }
int mpa_abi_version(void) { return 1; }

int mpa_device_available(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n > 0 ? 1 : 0;
}

// ---------------------------------------------------------------- World
int mpa_world_create(const double cellsize[3], const int mastergrid_size[3], const int mastergrid_origin[3],
                     const int pbc_repetitions[3], mpa_world** out) {
  return guard([&] {
    need(cellsize, "cellsize");
    need(out, "out");
    const int zero[3] = {0, 0, 0};
    *out = new mpa_world(cellsize, mastergrid_size ? mastergrid_size : zero,
                         mastergrid_origin ? mastergrid_origin : zero, pbc_repetitions ? pbc_repetitions : zero);
  });
}

int mpa_world_destroy(mpa_world* w) {
  return guard([&] { delete w; });
}

int mpa_world_get_cellsize(const mpa_world* w, double out[3]) {
  return guard([&] {
    need(w, "world");
    for (int a = 0; a < 3; ++a) out[a] = w->w.cellsize[a];
  });
}

int mpa_world_set_bias_magnetic_field(mpa_world* w, const double b[3]) {
  return guard([&] {
    need(w, "world");
    for (int a = 0; a < 3; ++a) w->w.bias[a] = (double)(float)b[a];  // FP32 real3 in the reference
    w->w.touch();
  });
}

int mpa_world_get_bias_magnetic_field(const mpa_world* w, double out[3]) {
  return guard([&] {
    need(w, "world");
    for (int a = 0; a < 3; ++a) out[a] = w->w.bias[a];
  });
}

int mpa_world_add_ferromagnet(mpa_world* w, const int size[3], const int origin[3], const uint8_t* geometry,
                              const unsigned int* regions, const char* name, mpa_ferromagnet** out) {
  return guard([&] {
    need(w, "world");
    need(size, "size");
    const int zero[3] = {0, 0, 0};
    Ferromagnet* f = w->w.add_ferromagnet(size, origin ? origin : zero, geometry, regions, name);
    if (out) *out = H(f);
  });
}

int mpa_world_num_ferromagnets(const mpa_world* w, int* out) {
  return guard([&] {
    need(w, "world");
    *out = (int)w->w.magnets.size();
  });
}

int mpa_world_ferromagnet_at(mpa_world* w, int index, mpa_ferromagnet** out) {
  return guard([&] {
    need(w, "world");
    if (index < 0 || index >= (int)w->w.magnets.size()) throw Error(MPA_ERR_INDEX, "ferromagnet index out of range");
    *out = H(w->w.magnets[index].get());
  });
}

int mpa_world_find_ferromagnet(mpa_world* w, const char* name, mpa_ferromagnet** out) {
  return guard([&] {
    need(w, "world");
    need(name, "name");
    *out = H(w->w.find(name));
  });
}

int mpa_world_add_antiferromagnet(mpa_world* w, const int size[3], const int origin[3], const uint8_t* geometry,
                                  const unsigned int* regions, const char* name, mpa_antiferromagnet** out) {
  return guard([&] {
    need(w, "world");
    need(size, "size");
    const int zero[3] = {0, 0, 0};
    Antiferromagnet* a = w->w.add_antiferromagnet(size, origin ? origin : zero, geometry, regions, name);
    if (out) *out = HA(a);
  });
}

int mpa_world_num_antiferromagnets(const mpa_world* w, int* out) {
  return guard([&] {
    need(w, "world");
    need(out, "out");
    *out = (int)w->w.afms.size();
  });
}

int mpa_world_antiferromagnet_at(mpa_world* w, int index, mpa_antiferromagnet** out) {
  return guard([&] {
    need(w, "world");
    need(out, "out");
    if (index < 0 || index >= (int)w->w.afms.size())
      throw Error(MPA_ERR_INDEX, "antiferromagnet index out of range");
    *out = HA(w->w.afms[index].get());
  });
}

int mpa_world_find_antiferromagnet(mpa_world* w, const char* name, mpa_antiferromagnet** out) {
  return guard([&] {
    need(w, "world");
    need(name, "name");
    need(out, "out");
    *out = HA(w->w.find_afm(name));
  });
}

int mpa_world_set_pbc(mpa_world* w, const int mastergrid_size[3], const int mastergrid_origin[3], const int reps[3]) {
  return guard([&] {
    need(w, "world");
    const int zero[3] = {0, 0, 0};
    w->w.set_pbc(mastergrid_size, mastergrid_origin ? mastergrid_origin : zero, reps);
  });
}

int mpa_world_set_pbc_auto(mpa_world* w, const int reps[3]) {
  return guard([&] {
    need(w, "world");
    w->w.set_pbc_auto(reps);
  });
}

int mpa_world_unset_pbc(mpa_world* w) {
  return guard([&] {
    need(w, "world");
    w->w.unset_pbc();
  });
}

int mpa_world_get_pbc(const mpa_world* w, int mastergrid_size[3], int mastergrid_origin[3], int reps[3]) {
  return guard([&] {
    need(w, "world");
    for (int a = 0; a < 3; ++a) {
      if (mastergrid_size) mastergrid_size[a] = w->w.mastergrid_size[a];
      if (mastergrid_origin) mastergrid_origin[a] = w->w.mastergrid_origin[a];
      if (reps) reps[a] = w->w.pbc[a];
    }
  });
}

int mpa_world_bounding_grid(const mpa_world* w, int size[3], int origin[3]) {
  return guard([&] {
    need(w, "world");
    w->w.bounding_grid(size, origin, true);
  });
}

int mpa_world_stream(mpa_world* w, void** stream) {
  return guard([&] {
    need(w, "world");
    *stream = (void*)w->w.stream;
  });
}

// ----------------------------------------------------------- Ferromagnet
int mpa_fm_name(const mpa_ferromagnet* m, char* buf, int buflen) {
  return guard([&] { copy_str(FM(m)->name, buf, buflen); });
}

int mpa_fm_grid(const mpa_ferromagnet* m, int size[3], int origin[3]) {
  return guard([&] { magnet_grid(FM(m), size, origin); });
}

int mpa_fm_geometry(const mpa_ferromagnet* m, uint8_t* out) {
  return guard([&] { magnet_geometry(FM(m), out); });
}

int mpa_fm_cellsingeo(const mpa_ferromagnet* m, int* out) {
  return guard([&] { *out = FM(m)->cellsingeo; });
}

int mpa_fm_set_flag(mpa_ferromagnet* m, const char* flag, int value) {
  return guard([&] {
    Ferromagnet* f = FM(m);
    need(flag, "flag");
    const std::string k = flag;
    if (k == "enable_demag") f->enable_demag = value != 0;
    else if (k == "enable_openbc") f->openbc = value != 0;
    else if (k == "enable_zhang_li_torque") f->enable_zhangli = value != 0;
    else if (k == "enable_slonczewski_torque") f->enable_slonc = value != 0;
    else if (k == "fixed_layer_on_top") f->fixed_on_top = value != 0;
    else if (k == "noise_seed") f->noise_seed = (unsigned)value;
    else if (k == "enable_elastodynamics" && f->host)
      throw Error(MPA_ERR_INVALID, "Cannot enable/disable elastodynamics for a sublattice.");
    else if (!magnet_set_flag(f, k, value)) throw Error(MPA_ERR_INVALID, "unknown flag '" + k + "'");
    f->world->touch();
  });
}

int mpa_fm_get_flag(const mpa_ferromagnet* m, const char* flag, int* out) {
  return guard([&] {
    const Ferromagnet* f = FM(m);
    need(flag, "flag");
    const std::string k = flag;
    if (k == "enable_demag") *out = f->enable_demag;
    else if (k == "enable_openbc") *out = f->openbc;
    else if (k == "enable_zhang_li_torque") *out = f->enable_zhangli;
    else if (k == "enable_slonczewski_torque") *out = f->enable_slonc;
    else if (k == "fixed_layer_on_top") *out = f->fixed_on_top;
    else if (k == "noise_seed") *out = (int)f->noise_seed;
    else if (k == "noise_draws") {
      // the 64-bit counter as an int flag; mpa_fm_get_noise_state gives it whole
      if (f->noise_draws[0] > (unsigned long long)INT_MAX)
        throw Error(MPA_ERR_INDEX, "noise_draws exceeds an int: read it with mpa_fm_get_noise_state");
      *out = (int)f->noise_draws[0];
    }
    else if (!magnet_get_flag(f, k, out)) throw Error(MPA_ERR_INVALID, "unknown flag '" + k + "'");
  });
}

int mpa_fm_get_noise_state(const mpa_ferromagnet* m, uint32_t* seed, uint64_t* step_draws,
                           uint64_t* quantity_draws) {
  return guard([&] {
    const Ferromagnet* f = FM(m);
    need(seed, "seed");
    need(step_draws, "step_draws");
    need(quantity_draws, "quantity_draws");
    *seed = f->noise_seed;
    *step_draws = f->noise_draws[0];
    *quantity_draws = f->noise_draws[1];
  });
}

int mpa_fm_set_noise_state(mpa_ferromagnet* m, uint32_t seed, uint64_t step_draws, uint64_t quantity_draws) {
  return guard([&] {
    Ferromagnet* f = FM(m);
    f->noise_seed = seed;
    f->noise_draws[0] = step_draws;
    f->noise_draws[1] = quantity_draws;
  });
}

int mpa_fm_param_info(const mpa_ferromagnet* m, const char* key, int* ncomp, int* is_uniform, char* name,
                      int namelen, char* unit, int unitlen) {
  return guard([&] { param_info(FM(m), key, ncomp, is_uniform, name, namelen, unit, unitlen); });
}

int mpa_fm_param_set_uniform(mpa_ferromagnet* m, const char* key, const double* value) {
  return guard([&] { param_set_uniform(FM(m), key, value); });
}

int mpa_fm_param_set_cells(mpa_ferromagnet* m, const char* key, const float* values) {
  return guard([&] { param_set_cells(FM(m), key, values); });
}

int mpa_fm_param_get(const mpa_ferromagnet* m, const char* key, float* out) {
  return guard([&] { param_get(FM(m), key, out); });
}

int mpa_fm_param_get_uniform(const mpa_ferromagnet* m, const char* key, double* out) {
  return guard([&] { param_get_uniform(FM(m), key, out); });
}

int mpa_fm_param_add_time_term(mpa_ferromagnet* m, const char* key, mpa_time_fn fn, void* ctx, const float* mask) {
  return guard([&] { param_add_time_term(FM(m), key, fn, ctx, mask); });
}

int mpa_fm_param_remove_time_terms(mpa_ferromagnet* m, const char* key) {
  return guard([&] { param_remove_time_terms(FM(m), key); });
}

int mpa_fm_param_is_dynamic(const mpa_ferromagnet* m, const char* key, int* out) {
  return guard([&] { param_is_dynamic(FM(m), key, out); });
}

int mpa_fm_poisson_init(mpa_ferromagnet* m) {
  return guard([&] {
    Ferromagnet* f = FM(m);
    f->poisson.init(f->poisson_desc(), f->world->stream);
  });
}

int mpa_fm_poisson_solve(mpa_ferromagnet* m, float* out) {
  return guard([&] {
    need(out, "out");
    FM(m)->eval_quantity("electrical_potential", out, FM(m)->world->stream);
  });
}

int mpa_fm_poisson_step(mpa_ferromagnet* m) {
  return guard([&] {
    Ferromagnet* f = FM(m);
    f->poisson.step(f->world->stream);
    HIP_CHECK(hipStreamSynchronize(f->world->stream));
  });
}

int mpa_fm_poisson_restart(mpa_ferromagnet* m) {
  return guard([&] {
    Ferromagnet* f = FM(m);
    f->poisson.restart(f->world->stream);
    HIP_CHECK(hipStreamSynchronize(f->world->stream));
  });
}

int mpa_fm_poisson_solver_solve(mpa_ferromagnet* m, double* out) {
  return guard([&] {
    Ferromagnet* f = FM(m);
    hipStream_t s = f->world->stream;
    f->poisson.solve(s);
    if (out && f->poisson.ready())
      HIP_CHECK(hipMemcpyAsync(out, f->poisson.x(), (size_t)f->poisson.rows() * sizeof(double),
                               hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
  });
}

int mpa_fm_poisson_residual(mpa_ferromagnet* m, double* out) {
  return guard([&] {
    need(out, "out");
    Ferromagnet* f = FM(m);
    *out = f->poisson.residual_max_norm(f->world->stream);
  });
}

int mpa_fm_poisson_set_method(mpa_ferromagnet* m, const char* name) {
  return guard([&] {
    need(name, "name");
    Ferromagnet* f = FM(m);
    f->poisson.set_method(name);
    // LinSolver::setMethod: a new stepper, restarted (linsolver.cpp)
    f->poisson.restart(f->world->stream);
    HIP_CHECK(hipStreamSynchronize(f->world->stream));
  });
}

int mpa_fm_poisson_set_option(mpa_ferromagnet* m, const char* key, double value) {
  return guard([&] {
    need(key, "key");
    const std::string k = key;
    if (k == "max_iter") FM(m)->poisson.max_iter = (int)value;
    else if (k == "tol") FM(m)->poisson.tol = value;
    else throw Error(MPA_ERR_INVALID, "unknown poisson solver option '" + k + "'");
  });
}

int mpa_fm_poisson_get_option(const mpa_ferromagnet* m, const char* key, double* out) {
  return guard([&] {
    need(key, "key");
    need(out, "out");
    const std::string k = key;
    if (k == "max_iter") *out = FM(m)->poisson.max_iter;
    else if (k == "tol") *out = FM(m)->poisson.tol;
    else throw Error(MPA_ERR_INVALID, "unknown poisson solver option '" + k + "'");
  });
}

int mpa_fm_param_set_in_region(mpa_ferromagnet* m, const char* key, unsigned int region, const double* value) {
  return guard([&] {
    need(key, "key");
    need(value, "value");
    FM(m)->param_set_in_region(key, region, value);
  });
}

int mpa_fm_regions(const mpa_ferromagnet* m, unsigned int* out) {
  return guard([&] { magnet_regions(FM(m), out); });
}

int mpa_fm_unique_regions(const mpa_ferromagnet* m, unsigned int* out, int cap, int* count) {
  return guard([&] { magnet_unique_regions(FM(m), out, cap, count); });
}

int mpa_fm_inter_info(const mpa_ferromagnet* m, const char* key, int* is_uniform, double* uniform_value,
                       long long* size, char* name, int namelen, char* unit, int unitlen) {
  return guard([&] { inter_info(FM(m), key, is_uniform, uniform_value, size, name, namelen, unit, unitlen); });
}

int mpa_fm_inter_set(mpa_ferromagnet* m, const char* key, double value) {
  return guard([&] {
    need(key, "key");
    FM(m)->inter_set(key, value);
  });
}

int mpa_fm_inter_set_between(mpa_ferromagnet* m, const char* key, unsigned int i, unsigned int j, double value) {
  return guard([&] {
    need(key, "key");
    FM(m)->inter_set_between(key, i, j, value);
  });
}

int mpa_fm_inter_get_between(const mpa_ferromagnet* m, const char* key, unsigned int i, unsigned int j, double* out) {
  return guard([&] {
    need(key, "key");
    need(out, "out");
    *out = FM(m)->inter_get_between(key, i, j);
  });
}

int mpa_fm_inter_eval(const mpa_ferromagnet* m, const char* key, float* out) {
  return guard([&] { inter_eval(FM(m), key, out); });
}

int mpa_fm_set_magnetization(mpa_ferromagnet* m, const float* values, long long count) {
  return guard([&] {
    Ferromagnet* f = FM(m);
    need(values, "values");
    f->set_magnetization(values, count, f->world->stream);
  });
}

int mpa_fm_get_magnetization(mpa_ferromagnet* m, float* out) {
  return guard([&] {
    Ferromagnet* f = FM(m);
    need(out, "out");
    f->get_magnetization(out, f->world->stream);
  });
}

int mpa_fm_quantity_ncomp(const mpa_ferromagnet* m, const char* q, int* ncomp) {
  return guard([&] {
    need(q, "quantity");
    *ncomp = FM(m)->quantity_ncomp(q);
  });
}

int mpa_fm_eval(mpa_ferromagnet* m, const char* q, float* out) {
  return guard([&] {
    Ferromagnet* f = FM(m);
    need(q, "quantity");
    need(out, "out");
    f->eval_quantity(q, out, f->world->stream);
  });
}

int mpa_fm_average(mpa_ferromagnet* m, const char* q, double* out) {
  return guard([&] {
    Ferromagnet* f = FM(m);
    need(q, "quantity");
    need(out, "out");
    f->average_quantity(q, out, f->world->stream);
  });
}

int mpa_fm_scalar(mpa_ferromagnet* m, const char* q, double* out) {
  return guard([&] {
    Ferromagnet* f = FM(m);
    need(q, "quantity");
    need(out, "out");
    *out = f->scalar_quantity(q, f->world->stream);
  });
}

int mpa_fm_demag_kernel(mpa_ferromagnet* m, int ksize[3], float* out) {
  return guard([&] {
    Ferromagnet* f = FM(m);
    DemagGeometry g;
    for (int a = 0; a < 3; ++a) {
      g.n[a] = f->n[a];
      g.cs[a] = f->world->cellsize[a];
      g.mastergrid[a] = f->world->mastergrid_size[a];
      g.pbc[a] = f->world->pbc[a];
    }
    if (!out) reference_kernel_size(g, ksize);
    else reference_kernel(g, ksize, out);
  });
}

int mpa_fm_host(mpa_ferromagnet* m, mpa_antiferromagnet** out) {
  return guard([&] {
    need(out, "out");
    *out = HA(FM(m)->host);
  });
}

int mpa_fm_set_variable(mpa_ferromagnet* m, const char* which, const float* values, long long count) {
  return guard([&] {
    Ferromagnet* f = FM(m);
    need(which, "variable");
    need(values, "values");
    if (f->host) throw Error(MPA_ERR_INVALID, "a sublattice has no elastic variables; use its antiferromagnet");
    f->set_variable(which, values, count, f->world->stream);
  });
}

int mpa_fm_get_variable(mpa_ferromagnet* m, const char* which, float* out) {
  return guard([&] {
    Ferromagnet* f = FM(m);
    need(which, "variable");
    need(out, "out");
    if (f->host) throw Error(MPA_ERR_INVALID, "a sublattice has no elastic variables; use its antiferromagnet");
    f->get_variable(which, out, f->world->stream);
  });
}

int mpa_fm_stray_field(mpa_ferromagnet* m, const int size[3], const int origin[3], int method, float* out) {
  return guard([&] {
    Ferromagnet* f = FM(m);
    need(size, "size");
    need(out, "out");
    const int zero[3] = {0, 0, 0};
    f->world->stray_field_on_grid(f, size, origin ? origin : zero, method, out);
  });
}

// ------------------------------------------------------- Antiferromagnet
int mpa_afm_name(const mpa_antiferromagnet* a, char* buf, int buflen) {
  return guard([&] { copy_str(AFM(a)->name, buf, buflen); });
}

int mpa_afm_grid(const mpa_antiferromagnet* a, int size[3], int origin[3]) {
  return guard([&] { magnet_grid(AFM(a), size, origin); });
}

int mpa_afm_geometry(const mpa_antiferromagnet* a, uint8_t* out) {
  return guard([&] { magnet_geometry(AFM(a), out); });
}

int mpa_afm_cellsingeo(const mpa_antiferromagnet* a, int* out) {
  return guard([&] {
    need(out, "out");
    *out = AFM(a)->cellsingeo;
  });
}

int mpa_afm_sublattice(mpa_antiferromagnet* a, int index, mpa_ferromagnet** out) {
  return guard([&] {
    Antiferromagnet* af = AFM(a);
    need(out, "out");
    if (index < 0 || index > 1) throw Error(MPA_ERR_INDEX, "sublattice index should be 0 or 1");
    *out = H(af->sub[index].get());
  });
}

int mpa_afm_set_flag(mpa_antiferromagnet* a, const char* flag, int value) {
  return guard([&] {
    need(flag, "flag");
    if (!magnet_set_flag(AFM(a), flag, value)) throw Error(MPA_ERR_INVALID, "unknown flag '" + std::string(flag) + "'");
  });
}

int mpa_afm_get_flag(const mpa_antiferromagnet* a, const char* flag, int* out) {
  return guard([&] {
    need(flag, "flag");
    need(out, "out");
    if (!magnet_get_flag(AFM(a), flag, out)) throw Error(MPA_ERR_INVALID, "unknown flag '" + std::string(flag) + "'");
  });
}

int mpa_afm_param_info(const mpa_antiferromagnet* a, const char* key, int* ncomp, int* is_uniform, char* name,
                       int namelen, char* unit, int unitlen) {
  return guard([&] { param_info(AFM(a), key, ncomp, is_uniform, name, namelen, unit, unitlen); });
}

int mpa_afm_param_set_uniform(mpa_antiferromagnet* a, const char* key, const double* value) {
  return guard([&] { param_set_uniform(AFM(a), key, value); });
}

int mpa_afm_param_set_cells(mpa_antiferromagnet* a, const char* key, const float* values) {
  return guard([&] { param_set_cells(AFM(a), key, values); });
}

int mpa_afm_param_get(const mpa_antiferromagnet* a, const char* key, float* out) {
  return guard([&] { param_get(AFM(a), key, out); });
}

int mpa_afm_param_get_uniform(const mpa_antiferromagnet* a, const char* key, double* out) {
  return guard([&] { param_get_uniform(AFM(a), key, out); });
}

int mpa_afm_param_add_time_term(mpa_antiferromagnet* a, const char* key, mpa_time_fn fn, void* ctx,
                                const float* mask) {
  return guard([&] { param_add_time_term(AFM(a), key, fn, ctx, mask); });
}

int mpa_afm_param_remove_time_terms(mpa_antiferromagnet* a, const char* key) {
  return guard([&] { param_remove_time_terms(AFM(a), key); });
}

int mpa_afm_param_is_dynamic(const mpa_antiferromagnet* a, const char* key, int* out) {
  return guard([&] { param_is_dynamic(AFM(a), key, out); });
}

int mpa_afm_param_set_in_region(mpa_antiferromagnet* m, const char* key, unsigned int region, const double* value) {
  return guard([&] {
    need(key, "key");
    need(value, "value");
    AFM(m)->param_set_in_region(key, region, value);
  });
}

int mpa_afm_regions(const mpa_antiferromagnet* m, unsigned int* out) {
  return guard([&] { magnet_regions(AFM(m), out); });
}

int mpa_afm_unique_regions(const mpa_antiferromagnet* m, unsigned int* out, int cap, int* count) {
  return guard([&] { magnet_unique_regions(AFM(m), out, cap, count); });
}

int mpa_afm_inter_info(const mpa_antiferromagnet* m, const char* key, int* is_uniform, double* uniform_value,
                       long long* size, char* name, int namelen, char* unit, int unitlen) {
  return guard([&] { inter_info(AFM(m), key, is_uniform, uniform_value, size, name, namelen, unit, unitlen); });
}

int mpa_afm_inter_set(mpa_antiferromagnet* m, const char* key, double value) {
  return guard([&] {
    need(key, "key");
    AFM(m)->inter_set(key, value);
  });
}

int mpa_afm_inter_set_between(mpa_antiferromagnet* m, const char* key, unsigned int i, unsigned int j, double value) {
  return guard([&] {
    need(key, "key");
    AFM(m)->inter_set_between(key, i, j, value);
  });
}

int mpa_afm_inter_get_between(const mpa_antiferromagnet* m, const char* key, unsigned int i, unsigned int j, double* out) {
  return guard([&] {
    need(key, "key");
    need(out, "out");
    *out = AFM(m)->inter_get_between(key, i, j);
  });
}

int mpa_afm_inter_eval(const mpa_antiferromagnet* m, const char* key, float* out) {
  return guard([&] { inter_eval(AFM(m), key, out); });
}

int mpa_afm_set_variable(mpa_antiferromagnet* a, const char* which, const float* values, long long count) {
  return guard([&] {
    Antiferromagnet* af = AFM(a);
    need(which, "variable");
    need(values, "values");
    af->set_variable(which, values, count, af->world->stream);
  });
}

int mpa_afm_get_variable(mpa_antiferromagnet* a, const char* which, float* out) {
  return guard([&] {
    Antiferromagnet* af = AFM(a);
    need(which, "variable");
    need(out, "out");
    af->get_variable(which, out, af->world->stream);
  });
}

int mpa_afm_stray_field(mpa_antiferromagnet* a, const int size[3], const int origin[3], int method, float* out) {
  return guard([&] {
    Antiferromagnet* af = AFM(a);
    need(size, "size");
    need(out, "out");
    const int zero[3] = {0, 0, 0};
    af->world->stray_field_on_grid(af, size, origin ? origin : zero, method, out);
  });
}

int mpa_afm_quantity_ncomp(const mpa_antiferromagnet* a, const char* q, int* ncomp) {
  return guard([&] {
    need(q, "quantity");
    need(ncomp, "ncomp");
    *ncomp = AFM(a)->quantity_ncomp(q);
  });
}

int mpa_afm_eval(mpa_antiferromagnet* a, const char* q, float* out) {
  return guard([&] {
    Antiferromagnet* af = AFM(a);
    need(q, "quantity");
    need(out, "out");
    af->eval_quantity(q, out, af->world->stream);
  });
}

int mpa_afm_average(mpa_antiferromagnet* a, const char* q, double* out) {
  return guard([&] {
    Antiferromagnet* af = AFM(a);
    need(q, "quantity");
    need(out, "out");
    af->average_quantity(q, out, af->world->stream);
  });
}

int mpa_afm_scalar(mpa_antiferromagnet* a, const char* q, double* out) {
  return guard([&] {
    Antiferromagnet* af = AFM(a);
    need(q, "quantity");
    need(out, "out");
    *out = af->scalar_quantity(q, af->world->stream);
  });
}

// ------------------------------------------------------------ TimeSolver
int mpa_ts_get(const mpa_world* w, const char* key, double* out) {
  return guard([&] {
    need(w, "world");
    need(key, "key");
    const TimeSolver& t = w->w.solver;
    const std::string k = key;
    if (k == "time") *out = t.time;
    else if (k == "timestep") *out = t.dt;
    else if (k == "adaptive_timestep") *out = t.adaptive ? 1.0 : 0.0;
    else if (k == "max_error") *out = t.max_error;
    else if (k == "headroom") *out = t.headroom;
    else if (k == "lower_bound") *out = t.lower_bound;
    else if (k == "upper_bound") *out = t.upper_bound;
    else if (k == "sensible_factor") *out = t.sensible_factor;
    else if (k == "sensible_timestep_default") *out = t.sensible_default;
    else if (k == "nsteps") *out = (double)t.nsteps;
    else if (k == "nrejected") *out = (double)t.nrejected;
    else if (k == "nrhs") *out = (double)t.nrhs;
    else throw Error(MPA_ERR_INVALID, "unknown time solver property '" + k + "'");
  });
}

int mpa_ts_set(mpa_world* w, const char* key, double value) {
  return guard([&] {
    need(w, "world");
    need(key, "key");
    TimeSolver& t = w->w.solver;
    const std::string k = key;
    if (k == "time") { t.time = value; w->w.touch(); }
    else if (k == "timestep") t.dt = value;
    else if (k == "adaptive_timestep") t.adaptive = value != 0.0;
    else if (k == "max_error") t.max_error = value;
    else if (k == "headroom") t.headroom = value;
    else if (k == "lower_bound") t.lower_bound = value;
    else if (k == "upper_bound") t.upper_bound = value;
    else if (k == "sensible_factor") t.sensible_factor = value;
    else if (k == "sensible_timestep_default") {
      if (value < 0) throw Error(MPA_ERR_RUNTIME, "The sensible timestep should be larger than zero.");
      t.sensible_default = value;
    } else throw Error(MPA_ERR_INVALID, "unknown or read-only time solver property '" + k + "'");
  });
}

int mpa_ts_sensible_timestep(mpa_world* w, double* out) {
  return guard([&] {
    need(w, "world");
    *out = w->w.solver.sensible_timestep();
  });
}

int mpa_ts_set_method(mpa_world* w, const char* name) {
  return guard([&] {
    need(w, "world");
    need(name, "method name");
    w->w.solver.set_method(name);
  });
}

int mpa_ts_get_method(const mpa_world* w, char* buf, int buflen) {
  return guard([&] {
    need(w, "world");
    copy_str(w->w.solver.method, buf, buflen);
  });
}

int mpa_ts_step(mpa_world* w) {
  return guard([&] {
    need(w, "world");
    w->w.solver.step();
    HIP_CHECK(hipStreamSynchronize(w->w.stream));
  });
}

int mpa_ts_steps(mpa_world* w, long long n) {
  return guard([&] {
    need(w, "world");
    if (n < 0) throw Error(MPA_ERR_INVALID, "number of steps must be non-negative");
    w->w.solver.steps(n);
  });
}

int mpa_ts_run(mpa_world* w, double duration) {
  return guard([&] {
    need(w, "world");
    w->w.solver.run(duration);
  });
}

// --------------------------------------------------------- relax / minimize
int mpa_world_relax(mpa_world* w, double tol) {
  return guard([&] {
    need(w, "world");
    w->w.relax(nullptr, tol);
  });
}

int mpa_world_minimize(mpa_world* w, double tol, int nsamples) {
  return guard([&] {
    need(w, "world");
    w->w.minimize(nullptr, tol, nsamples);
  });
}

int mpa_world_set_relax_torque_threshold(mpa_world* w, double value) {
  return guard([&] {
    need(w, "world");
    w->w.relax_threshold = value;
  });
}

int mpa_world_get_relax_torque_threshold(const mpa_world* w, double* out) {
  return guard([&] {
    need(w, "world");
    need(out, "out");
    *out = w->w.relax_threshold;
  });
}

int mpa_fm_relax(mpa_ferromagnet* m, double tol) {
  return guard([&] {
    Ferromagnet* f = FM(m);
    f->world->relax(f, tol);
  });
}

int mpa_fm_minimize(mpa_ferromagnet* m, double tol, int nsamples) {
  return guard([&] {
    Ferromagnet* f = FM(m);
    f->world->minimize(f, tol, nsamples);
  });
}

int mpa_fm_set_relax_torque_threshold(mpa_ferromagnet* m, double value) {
  return guard([&] { FM(m)->relax_threshold = value; });
}

int mpa_fm_get_relax_torque_threshold(const mpa_ferromagnet* m, double* out) {
  return guard([&] {
    need(out, "out");
    *out = FM(m)->relax_threshold;
  });
}

int mpa_afm_relax(mpa_antiferromagnet* a, double tol) {
  return guard([&] {
    Antiferromagnet* af = AFM(a);
    af->world->relax(af, tol);
  });
}

int mpa_afm_minimize(mpa_antiferromagnet* a, double tol, int nsamples) {
  return guard([&] {
    Antiferromagnet* af = AFM(a);
    af->world->minimize(af, tol, nsamples);
  });
}

// -------------------------------------------------------- Demag operator
int mpa_demag_plan_create(const int size[3], const double cellsize[3], const int mastergrid_size[3],
                          const int pbc_repetitions[3], mpa_demag_plan** out) {
  return guard([&] {
    need(size, "size");
    need(cellsize, "cellsize");
    need(out, "out");
    DemagGeometry g;
    for (int a = 0; a < 3; ++a) {
      g.n[a] = size[a];
      g.cs[a] = (double)(float)cellsize[a];  // the reference's FP32 cell size
      g.mastergrid[a] = mastergrid_size ? mastergrid_size[a] : 0;
      g.pbc[a] = pbc_repetitions ? pbc_repetitions[a] : 0;
      if ((g.mastergrid[a] == 0) != (g.pbc[a] == 0))
        throw Error(MPA_ERR_INVALID, "0 in size of mastergrid should match 0 in pbcRepetitions.");
    }
    *out = new mpa_demag_plan(g);
  });
}

int mpa_demag_plan_destroy(mpa_demag_plan* p) {
  return guard([&] { delete p; });
}

int mpa_demag_plan_exec(mpa_demag_plan* p, const float* m, const float* msat_cells, float msat_uniform,
                        const uint8_t* geometry, float* h, void* stream) {
  return guard([&] {
    need(p, "plan");
    need(m, "m");
    need(h, "h");
    p->plan.exec(m, PS{msat_cells, msat_uniform}, geometry, h, (hipStream_t)stream);
  });
}

int mpa_demag_plan_exec_noalloc(mpa_demag_plan* p, const float* m, const float* msat_cells, float msat_uniform,
                                const uint8_t* geometry, float* h, void* stream) {
  return guard([&] {
    need(p, "plan");
    need(m, "m");
    need(h, "h");
    p->plan.exec_noalloc(m, PS{msat_cells, msat_uniform}, geometry, h, (hipStream_t)stream);
  });
}

int mpa_demag_plan_info(const mpa_demag_plan* p, int fft_size[3], double* algorithmic_bytes, double* device_bytes) {
  return guard([&] {
    need(p, "plan");
    for (int a = 0; a < 3; ++a)
      if (fft_size) fft_size[a] = p->plan.K(a);
    if (algorithmic_bytes) *algorithmic_bytes = p->plan.algorithmic_bytes();
    if (device_bytes) *device_bytes = (double)p->plan.device_bytes();
  });
}

int mpa_world_set_option(mpa_world* w, const char* key, int value) {
  return guard([&] {
    need(w, "world");
    need(key, "key");
    const std::string k = key;
    if (k == "fuse_stages") w->w.solver.fuse_stages = value != 0;
    else if (k == "fuse_combine") w->w.solver.fuse_combine = value != 0;
    else if (k == "split_elastic") w->w.split_elastic = value != 0;
    else if (k == "split_demag") w->w.split_demag = value != 0;
    else if (k == "field_zchunk") {
      if (value < -1) throw Error(MPA_ERR_INVALID, "field_zchunk must be >= -1");
      w->w.field_zchunk = value;
    } else throw Error(MPA_ERR_INVALID, "unknown world option '" + k + "'");
  });
}

int mpa_world_get_option(const mpa_world* w, const char* key, int* value) {
  return guard([&] {
    need(w, "world");
    need(key, "key");
    need(value, "value");
    const std::string k = key;
    if (k == "fuse_stages") *value = w->w.solver.fuse_stages ? 1 : 0;
    else if (k == "fuse_combine") *value = w->w.solver.fuse_combine ? 1 : 0;
    else if (k == "split_elastic") *value = w->w.split_elastic ? 1 : 0;
    else if (k == "split_demag") *value = w->w.split_demag ? 1 : 0;
    else if (k == "field_zchunk") *value = w->w.field_zchunk;
    else throw Error(MPA_ERR_INVALID, "unknown world option '" + k + "'");
  });
}

// ---------------------------------------------------------------- Timing
int mpa_world_set_profiling(mpa_world* w, int enable) {
  return guard([&] {
    need(w, "world");
    w->w.prof.flush_sync(w->w.stream);
    w->w.prof.enabled = enable != 0;
  });
}

int mpa_world_profile_reset(mpa_world* w) {
  return guard([&] {
    need(w, "world");
    w->w.prof.flush_sync(w->w.stream);
    w->w.prof.reset();
  });
}

int mpa_world_profile_report(mpa_world* w, int cap, char* names, double* total_ms, long long* launches,
                             double* bytes, int* count) {
  return guard([&] {
    need(w, "world");
    need(count, "count");
    w->w.prof.flush_sync(w->w.stream);
    int i = 0;
    for (const auto& kv : w->w.prof.stats) {
      if (i < cap) {
        if (names) copy_str(kv.first, names + 32 * i, 32);
        if (total_ms) total_ms[i] = kv.second.total_ms;
        if (launches) launches[i] = kv.second.count;
        if (bytes) bytes[i] = kv.second.bytes;
      }
      ++i;
    }
    *count = i;
  });
}

int mpa_batch_create(long long capacity, const char* method, mpa_batch** out) {
  return guard([&] {
    need(method, "method");
    need(out, "out");
    *out = nullptr;
    *out = new mpa_batch(capacity, method);
  });
}

int mpa_batch_destroy(mpa_batch* b) {
  return guard([&] { delete b; });
}

int mpa_batch_add(mpa_batch* b, mpa_ferromagnet* m, int mode, double arg, int* id) {
  return guard([&] {
    need(b, "batch");
    need(id, "id");
    *id = b->b.add(FM(m), mode, arg);
  });
}

int mpa_batch_remove(mpa_batch* b, int id) {
  return guard([&] {
    need(b, "batch");
    b->b.remove(id);
  });
}

int mpa_batch_advance(mpa_batch* b, long long n) {
  return guard([&] {
    need(b, "batch");
    b->b.advance(n);
  });
}

int mpa_batch_finished(mpa_batch* b, int cap, int* ids, int* count) {
  return guard([&] {
    need(b, "batch");
    need(count, "count");
    const std::vector<int> f = b->b.finished();
    for (int i = 0; i < (int)f.size() && i < cap; ++i)
      if (ids) ids[i] = f[(size_t)i];
    *count = (int)f.size();
  });
}

int mpa_batch_member(mpa_batch* b, int id, double* time, double* dt, long long* steps, long long* rejected,
                     int* done, float* last_error, float* last_torque) {
  return guard([&] {
    need(b, "batch");
    const BatchMember& S = b->b.member_state(id);
    if (time) *time = S.time;
    if (dt) *dt = S.dt;
    if (steps) *steps = S.steps;
    if (rejected) *rejected = S.rejected;
    if (done) *done = S.done;
    if (last_error) *last_error = S.last_error;
    if (last_torque) *last_torque = S.last_torque;
  });
}

int mpa_batch_info(mpa_batch* b, long long* capacity, long long* free_cells, int* members, double* step_bytes,
                   int* launches_per_step) {
  return guard([&] {
    need(b, "batch");
    if (capacity) *capacity = b->b.capacity();
    if (free_cells) *free_cells = b->b.free_cells();
    if (members) *members = b->b.nmembers();
    if (step_bytes) *step_bytes = b->b.nmembers() ? b->b.step_bytes() : 0.0;
    if (launches_per_step) *launches_per_step = b->b.nmembers() ? b->b.launches_per_step() : 0;
  });
}

int mpa_batch_set_option(mpa_batch* b, const char* key, int value) {
  return guard([&] {
    need(b, "batch");
    need(key, "key");
    const std::string k = key;
    if (k == "graph") {
      b->b.use_graph = value != 0;
    } else if (k == "profiling") {
      b->b.prof.flush_sync(b->b.stream);
      b->b.prof.enabled = value != 0;
      if (value) b->b.prof.reset();
    } else if (k == "step_mode") {
      if (value < Batch::STEP_PASSES || value > Batch::STEP_AUTO)
        throw Error(MPA_ERR_INVALID, "step_mode must be 0 (passes), 1 (persistent) or 2 (auto)");
      b->b.step_mode = value;
    } else if (k == "persist_blocks_per_cu") {
      if (value < 0) throw Error(MPA_ERR_INVALID, "persist_blocks_per_cu must be >= 0");
      b->b.persist_blocks_per_cu = value;
    } else if (k == "persist_max_grid") {
      if (value < 0) throw Error(MPA_ERR_INVALID, "persist_max_grid must be >= 0");
      b->b.persist_max_grid = value;
    } else if (k == "persist_spin") {
      if (value < 0) throw Error(MPA_ERR_INVALID, "persist_spin must be >= 0");
      b->b.persist_spin = (unsigned)value;
      b->b.invalidate_persist();
    } else {
      throw Error(MPA_ERR_INVALID, "unknown batch option '" + k + "'");
    }
  });
}

int mpa_batch_tune_save(const char* path) {
  return guard([&] {
    need(path, "path");
    batch_tune_save(path);
  });
}

int mpa_batch_tune_load(const char* path) {
  return guard([&] {
    need(path, "path");
    batch_tune_load(path);
  });
}

int mpa_batch_tune_clear(void) {
  return guard([&] { batch_tune_clear(); });
}

int mpa_field_tile_report(int cap, char* text, int* len) {
  return guard([&] {
    need(len, "len");
    const std::string r = field_tile_report();
    *len = (int)r.size();
    if (text && cap > 0) copy_str(r, text, cap);
  });
}

int mpa_field_tile_clear(void) {
  return guard([&] { field_tile_clear(); });
}

int mpa_bench_dft8(int variant, int nlines, int rounds, const float* in, float* out, double* ms) {
  return guard([&] {
    need(in, "in");
    need(out, "out");
    const double t = bench_dft8(variant, nlines, rounds, in, out);
    if (ms) *ms = t;
  });
}

int mpa_batch_step_path(mpa_batch* b, int* last_path, int* persist_grid) {
  return guard([&] {
    need(b, "batch");
    if (last_path) *last_path = b->b.last_path();
    if (persist_grid) *persist_grid = b->b.nmembers() ? b->b.persist_grid() : 0;
  });
}

int mpa_batch_tune_report(mpa_batch* b, int cap, char* text, int* len) {
  return guard([&] {
    need(b, "batch");
    need(len, "len");
    const std::string r = b->b.tune_report();
    *len = (int)r.size();
    if (text && cap > 0) copy_str(r, text, cap);
  });
}

int mpa_batch_profile_report(mpa_batch* b, int cap, char* names, double* total_ms, long long* launches,
                             double* bytes, int* count) {
  return guard([&] {
    need(b, "batch");
    need(count, "count");
    b->b.prof.flush_sync(b->b.stream);
    int i = 0;
    for (const auto& kv : b->b.prof.stats) {
      if (i < cap) {
        if (names) copy_str(kv.first, names + 32 * i, 32);
        if (total_ms) total_ms[i] = kv.second.total_ms;
        if (launches) launches[i] = kv.second.count;
        if (bytes) bytes[i] = kv.second.bytes;
      }
      ++i;
    }
    *count = i;
  });
}

}  // extern "C"
