// Host object model: World, magnets (ferromagnets, antiferromagnets and their
// sublattices) and the Runge-Kutta TimeSolver.
//
// Mirrors the reference's object graph (MumaxWorld src/physics/mumaxworld.cpp,
// Magnet src/physics/magnet.cpp, Ferromagnet src/physics/ferromagnet.cpp:18-82,
// Antiferromagnet src/physics/antiferromagnet.cpp, TimeSolver
// src/core/timesolver.cpp, RungeKuttaStepper src/core/rungekutta.cpp) with
// device memory owned per magnet and all work on one HIP stream per world.
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "demag.hpp"
#include "elastic.hpp"
#include "fields.hpp"
#include "poisson.hpp"
#include "profiler.hpp"
#include "rk.hpp"
#include "stray.hpp"

namespace mpa {

class World;
class Ferromagnet;
class Antiferromagnet;

// A scalar (ncomp 1) or vector (ncomp 3) material parameter: a uniform value
// or one value per cell (Parameter / VectorParameter, src/core/parameter.hpp).
//
// Time-dependent terms (DynamicParameter, src/core/dynamic_parameter.hpp:10-68):
// the value at time t is static + sum_k f_k(t) (* mask_k, per cell and
// component, when the term has a mask), with t the time solver's (stage)
// time.  ps() / pv() evaluate the terms on the world's stream when the time
// or the terms changed since the last evaluation: a uniform sum when no term
// has a mask and the static value is uniform, otherwise one fill kernel into
// a per-cell array.
typedef int (*TimeFn)(double t, void* ctx, double* out /* ncomp */);
struct TimeTerm {
  TimeFn fn;
  void* ctx;
  DeviceArray<float> mask;  // ncomp * N, or empty
};

struct Param {
  std::string name, unit;
  int ncomp = 1;
  int N = 0;
  World* world = nullptr;
  float uniform[3] = {0.f, 0.f, 0.f};
  DeviceArray<float> cells;  // ncomp * N when non-uniform
  std::vector<TimeTerm> terms;
  bool is_dynamic() const { return !terms.empty(); }
  bool has_masks() const {
    for (const auto& t : terms)
      if (t.mask.n) return true;
    return false;
  }
  // Parameter::isUniform (parameter.cu:47-49): no per-cell values and no masked term
  bool is_uniform() const { return cells.n == 0 && !has_masks(); }
  // Parameter::assuredZero (parameter.cu:51-53)
  bool assured_zero() const {
    return !is_dynamic() && cells.n == 0 && uniform[0] == 0.f && uniform[1] == 0.f && uniform[2] == 0.f;
  }
  PS ps() const {
    if (is_dynamic()) {
      refresh();
      return PS{dyn_cells_ ? dyn_.get() : nullptr, eff_[0]};
    }
    return PS{cells.get(), uniform[0]};
  }
  PV pv() const {
    if (is_dynamic()) {
      refresh();
      return PV{dyn_cells_ ? dyn_.get() : nullptr, eff_[0], eff_[1], eff_[2]};
    }
    return PV{cells.get(), uniform[0], uniform[1], uniform[2]};
  }
  // the static value or the terms changed: the next ps() / pv() evaluates again
  void changed() { fresh_ = false; }

 private:
  void refresh() const;
  mutable float eff_[3] = {0.f, 0.f, 0.f};  // uniform value, or the per-cell fill's uniform addend
  mutable DeviceArray<float> dyn_;          // ncomp * N values at at_ when dyn_cells_
  mutable bool dyn_cells_ = false;
  mutable bool fresh_ = false;
  mutable double at_ = 0.0;
};

// A value between pairs of regions (InterParameter,
// src/core/inter_parameter.hpp:13-55): uniform, or one value per pair i != j
// of region indices below max + 1, in IP's lower-triangle order.
struct InterParam {
  std::string name, unit;
  float uniform = 0.f;
  size_t limit = 0;            // N (N - 1) / 2, N = largest region index + 1 (at least 1)
  std::vector<float> host;     // the table, when non-uniform
  DeviceArray<float> lut;
  bool is_uniform() const { return lut.n == 0; }
  bool assured_zero() const { return is_uniform() && uniform == 0.f; }
  IP ip() const { return IP{lut.get(), uniform}; }
};

struct Tableau {
  std::vector<double> nodes;
  std::vector<std::vector<double>> A;
  std::vector<double> w1, w2;
  int order1, order2;
  bool fsal;  // last row of A equals w1, last node is 1, last weight is 0
};
const Tableau& tableau_by_name(const std::string& name);  // throws MPA_ERR_INVALID

// Elastic displacement and velocity of a magnet (magnet.cpp:177-213), with
// their Runge-Kutta buffers and the strain / stress of the current u.
struct ElasticState {
  bool enabled = false;
  DeviceArray<float> u, v;    // 3N each
  DeviceArray<float> u0, v0;  // step start
  DeviceArray<float> un, vn;  // next stage state formed in the acceleration launch
  std::vector<DeviceArray<float>> ku, kv;
  DeviceArray<float> strain, stress;  // 6N each
};

// Grid, geometry, parameters and elastic state common to every magnet.
class MagnetCore {
 public:
  MagnetCore(World* w, const int size[3], const int origin[3], const uint8_t* geometry, std::string name);
  virtual ~MagnetCore() = default;

  World* world;
  std::string name;
  int n[3], origin[3], N;
  std::vector<uint8_t> geom_host;  // empty: every cell in the geometry
  DeviceArray<uint8_t> geom;
  int cellsingeo;
  std::map<std::string, Param> params;
  // regions (System::regions_, system.cpp:13-36): one index per cell, or
  // empty (every cell in region 0, and no region is defined)
  std::vector<uint32_t> regions_host;
  DeviceArray<uint32_t> regions;
  std::vector<uint32_t> unique_regions;  // sorted
  std::map<std::string, InterParam> inter_params;
  ElasticState el;
  // Magnet::enableAsStrayFieldSource / Destination (magnet.cpp:23-24); a
  // sublattice follows its host's flags (zeeman.cu:13-22).
  bool stray_source = true;
  bool stray_destination = true;
  DeviceArray<float> qbuf;  // 6N, magnet-level quantity results
  // Stray field of the other magnets on this one (3N), and one plan per
  // source magnet.
  DeviceArray<float> hstray;
  std::vector<std::pair<const MagnetCore*, std::unique_ptr<StrayPlan>>> stray_plans;
  // This magnet as a stray-field source (Ms m, or Ms1 m1 + Ms2 m2); false
  // when it is disabled as a source or its msat is assured zero
  // (strayfield.cpp:88-95).
  virtual bool stray_source_of(StraySource& out) = 0;
  // The other magnets' stray fields enter this magnet's external field
  // (zeeman.cu:13-34, 70-88).
  std::vector<std::pair<MagnetCore*, StraySource>> stray_sources();
  bool stray_active() { return !stray_sources().empty(); }
  void compute_stray(hipStream_t s, Profiler* prof);

  const uint8_t* geom_ptr() const { return geom.get(); }
  Param& param(const std::string& key);
  const Param& param(const std::string& key) const;
  bool has_param(const std::string& key) const { return params.count(key) != 0; }
  InterParam& inter(const std::string& key);
  const InterParam& inter(const std::string& key) const;
  // Set once, right after construction (and passed on to sublattices);
  // sizes the inter-region tables (inter_parameter.cu:7-20).
  virtual void set_regions(const uint32_t* regions);
  // System::checkIdxInRegions (system.cpp:73-77)
  void check_region(unsigned idx) const;
  // Parameter::setInRegion (parameter.cu:41-47, 144-150): the static value
  // in the cells of region idx
  void param_set_in_region(const std::string& key, unsigned idx, const double* value);
  // InterParameter::set / setBetween / getBetween (inter_parameter.cu:38-82)
  void inter_set(const std::string& key, double value);
  void inter_set_between(const std::string& key, unsigned i, unsigned j, double value);
  double inter_get_between(const std::string& key, unsigned i, unsigned j) const;

  // Elastodynamics of a magnet that is not a sublattice.
  void set_elastodynamics(bool on);
  bool elasticity_assured_zero() const;  // elastodynamics.cu:47-51
  bool rigid_strain_applied() const;     // magnetoelasticfield.cu:29-42
  ElDesc eldesc() const;
  // Ferromagnets whose magnetoelastic force acts on this magnet.
  virtual int num_fm() const = 0;
  virtual Ferromagnet* fm_at(int i) = 0;
  MEForce me_force_of(int i);
  void compute_strain(hipStream_t s);  // el.strain and el.stress from el.u
  void ensure_elastic_stage_buffers(int S);
  void set_variable(const std::string& which, const float* host, long long count, hipStream_t s);
  void get_variable(const std::string& which, float* host, hipStream_t s);
  // Magnet-level quantities (strain_tensor, stress_tensor, elastic forces,
  // acceleration, velocity); ncomp, or 0 if q is not one of them.
  int magnet_quantity_ncomp(const std::string& q) const;
  const float* magnet_eval_device(const std::string& q, hipStream_t s, int* ncomp);
  // kinetic_energy, elastic_energy
  double magnet_scalar(const std::string& q, hipStream_t s);
  void add_elastic_energies(float* out, hipStream_t s);

 protected:
  void add_params(const void* specs, int count);
  void add_inter(const std::string& key, float value, const char* unit);
};

class Ferromagnet : public MagnetCore {
 public:
  Ferromagnet(World* world, const int size[3], const int origin[3], const uint8_t* geometry, std::string name,
              Antiferromagnet* host = nullptr);

  Antiferromagnet* host;  // non-null for a sublattice
  bool enable_demag = true;
  bool openbc = false;
  // spin-transfer torque switches (ferromagnet.cpp:40-44, 51-52)
  bool enable_zhangli = true, enable_slonc = true, fixed_on_top = true;
  // thermal field random streams: seed, and the draws used by time steps
  // (purpose 0) and by evaluations of the thermal_noise quantity (purpose 1)
  unsigned noise_seed = 0;
  unsigned long long noise_draws[2] = {0, 0};
  unsigned long long step_draw = 0;  // the draw of the step in progress
  double relax_threshold = -1.0;  // Ferromagnet.RelaxTorqueThreshold (wrap_ferromagnet.cpp:67)

  DeviceArray<float> m;       // magnetisation, 3N
  DeviceArray<float> x0;      // RK start state
  DeviceArray<float> xn;      // next stage state written by a fused field evaluation
  DeviceArray<float> hdemag;  // demag field scratch (independent magnets)
  DeviceArray<float> scratch; // quantity evaluation
  // electrical potential (Ferromagnet::poissonSystem, ferromagnet.hpp:55)
  PoissonSolver poisson;
  PoissonDesc poisson_desc();
  std::vector<DeviceArray<float>> k;

  Ferromagnet* sister() const;
  MagnetCore* owner();  // the magnet holding the displacement: the host, or this
  unsigned active_terms();  // the reference's "assured zero" rules
  bool magnetoelastic_assured_zero();  // magnetoelasticfield.cu:10-42
  FMDesc desc(unsigned terms);
  DemagPlan& demag();
  void invalidate_demag() { demag_.reset(); }
  void ensure_stage_buffers(int S);
  // Fills the demag field and the strain this magnet's terms need.
  void prepare(unsigned terms, hipStream_t s, Profiler* prof);

  // out (device, 3N) = torque of the current magnetisation (the LLG torque,
  // or the damping-only torque of a relaxation).  With prepared, the demag
  // field and strain are taken as already computed.
  // noise_dt > 0: add the thermal field of the step in progress divided by
  // sqrt(noise_dt) (the stepper's noise term).
  void eval_rhs(float* out, hipStream_t s, bool prepared = false, FieldMode mode = MODE_LLG, double noise_dt = 0.0,
                const NextStage* next = nullptr);
  // The right-hand side can also write the next stage state in the same
  // launch (the uniform-parameter LLG path of an independent magnet).
  bool next_fusable();
  double rhs_bytes(unsigned terms);
  FMDesc rhs_desc(unsigned terms, double noise_dt);
  int eval_quantity(const std::string& q, float* host, hipStream_t s);
  int quantity_ncomp(const std::string& q) const;
  void average_quantity(const std::string& q, double* out, hipStream_t s);
  double scalar_quantity(const std::string& q, hipStream_t s);

  void set_magnetization(const float* host, long long count, hipStream_t s);
  void get_magnetization(float* host, hipStream_t s);

  int num_fm() const override { return 1; }
  Ferromagnet* fm_at(int) override { return this; }
  bool stray_source_of(StraySource& out) override;

  const float* eval_device(const std::string& q, hipStream_t s, int* ncomp);
  // the demag field this magnet reads (its own, or its antiferromagnet's)
  const float* demag_field() const;

 private:
  std::unique_ptr<DemagPlan> demag_;
};

// Two ferromagnetic sublattices on one grid (antiferromagnet.cpp:15-28).
class Antiferromagnet : public MagnetCore {
 public:
  Antiferromagnet(World* world, const int size[3], const int origin[3], const uint8_t* geometry, std::string name);

  std::unique_ptr<Ferromagnet> sub[2];
  DeviceArray<float> hdemag;  // stray field of both sublattices
  DeviceArray<float> scratch;

  DemagPlan& demag();
  void invalidate_demag() { demag_.reset(); }
  void compute_demag(hipStream_t s, Profiler* prof);

  int num_fm() const override { return 2; }
  Ferromagnet* fm_at(int i) override { return sub[i].get(); }
  void set_regions(const uint32_t* regions) override;  // the sublattices share the host's regions
  void eval_rhs_pair(float* out0, float* out1, hipStream_t s, FieldMode mode, double noise_dt);
  // eval_rhs_pair that also writes both sublattices' next stage states;
  // next_fusable tells whether the pair takes it (uniform-or-not alike,
  // MODE_LLG, no thermal noise).
  bool next_fusable();
  void eval_rhs_pair_next(float* out0, float* out1, hipStream_t s, const NextStage& n0, const NextStage& n1);
  // eval_rhs_pair in two parts, so that the first can run beside the demag
  // passes: the sublattices' fields without demag (into hloc), then their
  // torques from those and the demag field.  pair_split tells whether the
  // two parts give what eval_rhs_pair gives (torque_pair_splittable).
  bool pair_split(FieldMode mode, double noise_dt);
  void eval_local_pair(hipStream_t s, double noise_dt);
  void eval_torque_pair(float* out0, float* out1, hipStream_t s, FieldMode mode, double noise_dt);
  DeviceArray<float> hloc[2];
  bool stray_source_of(StraySource& out) override;

  // Quantities of the whole antiferromagnet (neel_vector, full_magnetization
  // and the magnet-level elastic ones).
  int quantity_ncomp(const std::string& q) const;
  const float* eval_device(const std::string& q, hipStream_t s, int* ncomp);
  int eval_quantity(const std::string& q, float* host, hipStream_t s);
  void average_quantity(const std::string& q, double* out, hipStream_t s);
  double scalar_quantity(const std::string& q, hipStream_t s);  // total_energy, kinetic_energy, elastic_energy

 private:
  std::unique_ptr<DemagPlan> demag_;
};

// One state variable of the time integration (DynamicEquation,
// src/core/dynamicequation.cpp:10-37).
struct StateVar {
  enum Kind { MAG, DISP, VEL } kind;
  Ferromagnet* fm;    // MAG
  MagnetCore* owner;  // DISP / VEL
  float* x;
  DeviceArray<float>* xa;  // the array holding x
  DeviceArray<float>* x0;
  std::vector<DeviceArray<float>>* k;
  int N;
  const uint8_t* geom;
  bool normalize;
};

class TimeSolver {
 public:
  explicit TimeSolver(World* w);
  double time = 0.0;
  double dt = 0.0;
  bool adaptive = true;
  double max_error = 1e-5, headroom = 0.8, lower_bound = 0.5, upper_bound = 2.0;
  double sensible_factor = 0.01, sensible_default = 1e-14;
  std::string method = "Fehlberg";
  long long nsteps = 0, nrejected = 0, nrhs = 0;
  // stage combinations inside the field launch where possible (one read of
  // the last derivative less per stage): 2.30 against 2.32 ms per step at
  // 256x256x64 on MI355X; for a lone antiferromagnet inside its
  // sublattice-field and acceleration launches: 0.431 against 0.4545 ms per
  // step at 512x512 with elastodynamics
  bool fuse_stages = true;
  // the stage combination of a lone magnetisation with demag inside the
  // demag's first FFT pass (no combination launch, one read of the state);
  // off by default: slower on MI355X at 256x256x64 (132 us for the fused pass
  // against 54 + 38 us for the combination and the x-forward pass)
  bool fuse_combine = false;

  void set_method(const std::string& name);
  double sensible_timestep();
  void adapt(double corr);
  void step();
  void steps(long long n);
  void run(double duration);
  void equations_changed();

 private:
  World* world_;
  bool fsal_valid_ = false;
  double fsal_time_ = 0.0;
  unsigned long long fsal_version_ = 0;
  // the last step's final evaluation also wrote the next step's first stage
  // state (fixed step h) into the magnet's xn
  bool next_valid_ = false;
  double next_h_ = 0.0;
};

class World {
 public:
  World(const double cellsize[3], const int mastergrid_size[3], const int mastergrid_origin[3], const int pbc[3]);
  ~World();

  double cellsize[3];
  int mastergrid_size[3], mastergrid_origin[3], pbc[3];
  double bias[3] = {0, 0, 0};
  std::vector<std::unique_ptr<Ferromagnet>> magnets;  // independent ferromagnets
  std::vector<std::unique_ptr<Antiferromagnet>> afms;
  TimeSolver solver;
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr;  // strains and fields without demag of a right-hand side
  hipStream_t stream3 = nullptr;  // accelerations of a right-hand side
  hipEvent_t ev_fork = nullptr, ev_strain = nullptr, ev_local = nullptr, ev_join = nullptr, ev_join3 = nullptr;
  Reducer reducer;
  Profiler prof;
  unsigned long long version = 1;  // bumped by every state change (invalidates FSAL reuse)
  // Uniform-parameter field kernel: z planes per block of the LDS halo-tile
  // variant, 0 for the one-cell-per-thread variant.
  int field_zchunk = 4;
  // Elastodynamics: the strain and the acceleration of a right-hand side run
  // on a second stream, concurrently with the demag passes (joined before
  // the magnetoelastic field and at the end of the evaluation).  Off by
  // default: on a 512 x 512 antiferromagnet the overlap varied from step
  // to step between a little faster and a good deal slower than one stream.
  bool split_elastic = false;
  // Antiferromagnets with demag: the sublattice fields without demag run on
  // a second stream beside the demag passes; a torque launch after the x
  // inverse adds the demag field (the same values bit for bit).  Off by
  // default: at 512 x 512 the field launch fills every compute unit and the
  // demag passes beside it slow down more than the overlap gains.
  bool split_demag = false;

  // regions: one index per cell, or nullptr (mumaxworld.cpp:57-103)
  Ferromagnet* add_ferromagnet(const int size[3], const int origin[3], const uint8_t* geometry,
                               const uint32_t* regions, const char* name);
  Antiferromagnet* add_antiferromagnet(const int size[3], const int origin[3], const uint8_t* geometry,
                                       const uint32_t* regions, const char* name);
  Ferromagnet* find(const std::string& name);
  Antiferromagnet* find_afm(const std::string& name);
  // every ferromagnet with a magnetisation: independent ones and sublattices
  std::vector<Ferromagnet*> all_ferromagnets();
  // require_magnets: an index error for a world without magnets (else all zeros)
  void bounding_grid(int size[3], int origin[3], bool require_magnets = false) const;
  void set_pbc(const int mg_size[3], const int mg_origin[3], const int reps[3]);
  void set_pbc_auto(const int reps[3]);
  void unset_pbc();
  void touch() { ++version; }

  // Equations in the order of MumaxWorld::resetTimeSolverEquations
  // (mumaxworld.cpp:159-205), or the relaxation's equations while it runs.
  std::vector<StateVar> state_vars();
  // k[stage] of every variable at the current state (all stage states set).
  // disp_ready: the displacements' k[stage] (the stage velocity) is written.
  // noise_dt > 0: magnetisation right-hand sides include the thermal noise
  // term of the step in progress (rungekutta.cpp:95-106).
  // next: the single magnetisation variable's evaluation also writes the
  // next stage state (Ferromagnet::next_fusable).
  // stage_in: the single magnetisation variable's stage state is formed inside
  // its demag's first FFT pass (x = normalise(x0 + sum_j c_j k_j)).
  struct StageIn {
    const float* x0;
    Combo co;
  };
  // pair_next: an antiferromagnet's evaluation also writes the next stage
  // state of both sublattices (in its field launch) and, with elastodynamics,
  // of its displacement and velocity (in its acceleration launch).
  struct PairNext {
    Antiferromagnet* afm;
    NextStage m[2];
    bool has_el;
    ElNext el;
  };
  void eval_rhs_all(const std::vector<StateVar>& vars, int stage, hipStream_t s, bool disp_ready = false,
                    double noise_dt = 0.0, const NextStage* next = nullptr, const StageIn* stage_in = nullptr,
                    const PairNext* pair_next = nullptr);
  bool relaxing() const { return relaxing_; }

  // Field of `src` on the grid (size, origin); method 0 = FFT, 1 = direct sum.
  void stray_field_on_grid(MagnetCore* src, const int size[3], const int origin[3], int method, float* host_out);

  double relax_threshold = -1.0;  // World.RelaxTorqueThreshold (mumaxworld.cpp:26)
  // Relaxer::exec (relaxer.cpp:87-174) of one magnet, or of the whole world
  // when target is nullptr.
  void relax(MagnetCore* target, double tol);
  // Minimizer::exec (minimizer.cu:68-160) of one magnet, or of the whole
  // world when target is nullptr.
  void minimize(MagnetCore* target, double tol, int nsamples);

 private:
  void check_pbc(const int mg_size[3], const int reps[3]) const;
  void check_in_mastergrid(const int size[3], const int origin[3], const int mg_size[3], const int mg_origin[3],
                           bool adding) const;
  void check_no_overlap(const int size[3], const int origin[3]) const;
  void check_name(const std::string& name);
  void pbc_changed();
  double total_energy(MagnetCore* m);
  // Active while a relaxation runs: damping-only torques of these
  // ferromagnets are the equations (plus the elastic ones for a world relax).
  bool relaxing_ = false;
  bool relax_elastic_ = false;
  std::vector<Ferromagnet*> relax_fms_;
};

}  // namespace mpa
