// Host object model: World, Ferromagnet, TimeSolver.
//
// Mirrors the reference's object graph (MumaxWorld src/physics/mumaxworld.cpp,
// Ferromagnet src/physics/ferromagnet.cpp:18-82, TimeSolver
// src/core/timesolver.cpp, RungeKuttaStepper src/core/rungekutta.cpp) with
// device memory owned per magnet and all work on one HIP stream per world.
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "demag.hpp"
#include "fields.hpp"
#include "profiler.hpp"
#include "rk.hpp"

namespace mpa {

class World;

// A scalar (ncomp 1) or vector (ncomp 3) material parameter: a uniform value
// or one value per cell (Parameter / VectorParameter, src/core/parameter.hpp).
struct Param {
  std::string name, unit;
  int ncomp = 1;
  float uniform[3] = {0.f, 0.f, 0.f};
  DeviceArray<float> cells;  // ncomp * N when non-uniform
  bool is_uniform() const { return cells.n == 0; }
  bool assured_zero() const {
    return is_uniform() && uniform[0] == 0.f && uniform[1] == 0.f && uniform[2] == 0.f;
  }
  PS ps() const { return PS{cells.get(), uniform[0]}; }
  PV pv() const { return PV{cells.get(), uniform[0], uniform[1], uniform[2]}; }
};

struct Tableau {
  std::vector<double> nodes;
  std::vector<std::vector<double>> A;
  std::vector<double> w1, w2;
  int order1, order2;
  bool fsal;  // last row of A equals w1, last node is 1, last weight is 0
};
const Tableau& tableau_by_name(const std::string& name);  // throws MPA_ERR_INVALID

class Ferromagnet {
 public:
  Ferromagnet(World* world, const int size[3], const int origin[3], const uint8_t* geometry, std::string name);

  World* world;
  std::string name;
  int n[3], origin[3], N;
  std::vector<uint8_t> geom_host;  // empty: every cell in the geometry
  DeviceArray<uint8_t> geom;
  int cellsingeo;
  std::map<std::string, Param> params;
  bool enable_demag = true;
  bool openbc = false;
  bool stray_source = true;       // Magnet::enableAsStrayFieldSource (magnet.cpp:26-34)
  bool stray_destination = true;  // Magnet::enableAsStrayFieldDestination

  DeviceArray<float> m;       // magnetisation, 3N
  DeviceArray<float> x0;      // RK start state
  DeviceArray<float> hdemag;  // demag field scratch
  DeviceArray<float> scratch; // quantity evaluation
  std::vector<DeviceArray<float>> k;

  const uint8_t* geom_ptr() const { return geom.get(); }
  Param& param(const std::string& key);
  unsigned active_terms() const;  // the reference's "assured zero" rules
  // Throws when another magnet's stray field would enter this magnet's
  // external field (not built); called by every evaluation that includes it.
  void check_stray_coupling() const;
  FMDesc desc(unsigned terms) const;
  DemagPlan& demag();
  void invalidate_demag() { demag_.reset(); }
  void ensure_stage_buffers(int S);

  // out (device, 3N) = torque of the current magnetisation
  void eval_rhs(float* out, hipStream_t s, bool check_stray = true);
  // Evaluate a named quantity into host memory (ncomp*N floats); returns ncomp.
  int eval_quantity(const std::string& q, float* host, hipStream_t s);
  int quantity_ncomp(const std::string& q) const;
  // Average of each component over the geometry cells (FP64).
  void average_quantity(const std::string& q, double* out, hipStream_t s);
  double scalar_quantity(const std::string& q, hipStream_t s);

  void set_magnetization(const float* host, long long count, hipStream_t s);
  void get_magnetization(float* host, hipStream_t s);

 private:
  // field / density of the given term set into scratch (device)
  const float* eval_device(const std::string& q, hipStream_t s, int* ncomp);
  std::unique_ptr<DemagPlan> demag_;
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
};

class World {
 public:
  World(const double cellsize[3], const int mastergrid_size[3], const int mastergrid_origin[3], const int pbc[3]);
  ~World();

  double cellsize[3];
  int mastergrid_size[3], mastergrid_origin[3], pbc[3];
  double bias[3] = {0, 0, 0};
  std::vector<std::unique_ptr<Ferromagnet>> magnets;
  TimeSolver solver;
  hipStream_t stream = nullptr;
  Reducer reducer;
  Profiler prof;
  unsigned long long version = 1;  // bumped by every state change (invalidates FSAL reuse)

  Ferromagnet* add_ferromagnet(const int size[3], const int origin[3], const uint8_t* geometry, const char* name);
  Ferromagnet* find(const std::string& name);
  void bounding_grid(int size[3], int origin[3]) const;
  void set_pbc(const int mg_size[3], const int mg_origin[3], const int reps[3]);
  void set_pbc_auto(const int reps[3]);
  void unset_pbc();
  void touch() { ++version; }

 private:
  void check_pbc(const int mg_size[3], const int reps[3]) const;
  void check_in_mastergrid(const int size[3], const int origin[3], const int mg_size[3], const int mg_origin[3]) const;
  void pbc_changed();
};

}  // namespace mpa
