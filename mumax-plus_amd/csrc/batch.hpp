// Many small independent ferromagnets stepped together.
//
// The reference steps one World at a time (TimeSolver::steps / run,
// src/core/timesolver.cpp:37-111, one kernel launch per term per magnet).
// Here the members' magnetisations are packed into one SoA state of
// capacity `cs` cells (component c of cell i of a member at c * cs + off + i,
// every member starting on a multiple of 256 cells), so a Runge-Kutta stage
// of every member is a handful of launches: one stage combination, the
// batched demag passes (one launch per pass and FFT line length) and one
// batched field kernel.  Each member keeps its own step size, time,
// controller and stopping rule in device memory; the step controller, the
// accept / reject decision and the stopping rules run on the device, so the
// host only launches steps (a captured HIP graph once the membership is
// fixed) and polls for finished members now and then.
//
// A member's arithmetic does not depend on the other members: its field
// comes from its own cells and demag plan, its error norm and controller are
// its own.  Its trajectory is therefore the same alone and in any mix.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "demag.hpp"
#include "fields.hpp"
#include "profiler.hpp"

namespace mpa {

class Ferromagnet;

enum BatchMode {
  BATCH_STEPS = 0,  // a fixed number of accepted steps
  BATCH_RUN = 1,    // TimeSolver::run(duration): land on time + duration
  BATCH_RELAX = 2,  // damping-only dynamics until the max torque <= threshold
};
// A relaxation whose torque stalls above its threshold (the adaptive step's
// error floor) stops after this many accepted steps, not converged.
constexpr long long BATCH_RELAX_MAX_STEPS = 200000;
// Mode flag: start with the world's step as it is (a resumed member) instead
// of the sensible-step rule of TimeSolver.run / steps.
constexpr int BATCH_KEEP_STEP = 0x100;

// Per-member state in device memory.
struct BatchMember {
  double time, dt, stop, dt_saved;
  double max_error, headroom, lower, upper, sfactor, sdefault;
  long long steps, target, rejected;
  float threshold;
  float last_error, last_torque;
  int mode, adaptive, active, accept, last, done, keep_step;
};

class Batch {
 public:
  // capacity: packed cells (rounded up to 256); method: a tableau whose last
  // stage is the new state (DormandPrince or BogackiShampine).
  Batch(long long capacity, const std::string& method);
  ~Batch();

  // Admits a ferromagnet (its magnetisation, parameters, geometry and its
  // world's solver settings: dt, adaptive, max_error, headroom, bounds).
  // Returns the member id, or -1 when its cells do not fit now.
  // arg: steps (BATCH_STEPS), duration (BATCH_RUN), torque threshold
  // (BATCH_RELAX); mode may carry BATCH_KEEP_STEP.  The member's world must
  // hold only this magnet.
  int add(Ferromagnet* fm, int mode, double arg);
  // Copies the member's state back into its magnet and its world's solver
  // (time, dt, step counts) and frees its cells.
  void remove(int id);
  // n batch steps (one attempt per active member each).
  void advance(long long n);
  // Active members' ids, and those finished since they were admitted.
  std::vector<int> finished();
  int nmembers() const;
  long long free_cells() const;
  long long capacity() const { return cs_; }
  const BatchMember& member_state(int id);  // synchronises
  // Algorithmic HBM bytes of one batch step (all active members).
  double step_bytes();
  int launches_per_step();

  bool use_graph = true;
  // How a batch step runs: STEP_PASSES, one launch per pass (a captured
  // graph when use_graph); STEP_PERSISTENT, one persistent kernel for all n
  // steps of an advance; STEP_AUTO, whichever of these measures faster for
  // the batch's size class (see tune_report).
  enum StepMode { STEP_PASSES = 0, STEP_PERSISTENT = 1, STEP_AUTO = 2 };
  int step_mode = STEP_PASSES;
  int persist_blocks_per_cu = 0;    // resident blocks per compute unit (0: as many as fit)
  int persist_max_grid = 0;         // cap on the persistent grid (0: none)
  unsigned persist_spin = 1u << 22;  // grid-barrier poll limit (0: fail the first barrier)
  // Whether the persistent kernel can run this batch (line lengths up to 256
  // points), and the grid it launches (0 when it cannot).
  bool persist_supported();
  int persist_grid();
  // The step path the last advance took (0 passes, 1 persistent).
  int last_path() const { return last_path_; }
  // One line per size class the autotuner has seen: the class, the chosen
  // path and the measured ms per step of each candidate.
  std::string tune_report() const;
  // Rebuild the persistent kernel's tables and arguments before its next
  // launch (after a change of persist_spin).
  void invalidate_persist() { pdirty_ = true; }
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr;  // the single layers' FFT chain
  hipEvent_t fork_ = nullptr, join_ = nullptr;
  Profiler prof;

 private:
  struct Slot {
    Ferromagnet* fm = nullptr;
    long long off = 0, ncell = 0, span = 0;
    int mode = 0;
    bool demag = false;
    bool used = false;
    FMDesc desc{};  // the member's parameters as read on admission
  };
  void rebuild();
  void prime(const std::vector<int>& fresh);
  void record_step();
  void advance_passes(long long n);
  void advance_persistent(long long n, int blocks_per_cu, int cap);
  void build_persist();
  int grid_for(int blocks_per_cu, int cap) const;
  void persist_recover(int err);
  std::string size_class();
  void eval_rhs(const float* x, float* out, const BatchNext* next, const BatchErr* err);
  BatchErr step_error();
  bool two_chains() const;
  float* stage_state(int s);
  BatchNext next_stage(int s);
  void drop_graph();
  long long alloc_range(long long span);
  void free_range(long long off, long long span);

  std::string method_;
  long long cs_;
  std::vector<Slot> slots_;
  std::vector<std::pair<long long, long long>> free_;  // (off, span), sorted
  DeviceArray<float> X_, XO_, X0_, HD_;  // state (and stage states), the other stage states, step start, demag
  std::vector<DeviceArray<float>> K_;
  DeviceArray<uint8_t> G_;
  DeviceArray<int> c2m_;  // 256-cell chunk -> member id, -1 when free
  DeviceArray<BatchMember> st_;
  DeviceArray<unsigned> errbits_, tqbits_;
  std::vector<int> c2m_host_;
  std::unique_ptr<DemagBatch> demag_;
  std::unique_ptr<FieldBatch> field_;   // every member
  // the same split by demag group: single layers with demag (field1_) and
  // the others (field0_); the step runs the two groups' whole chains on two
  // streams
  std::unique_ptr<FieldBatch> field0_, field1_;
  double fbytes0_ = 0.0, fbytes1_ = 0.0;  // field launch bytes per group (without the next stage)
  bool dirty_ = true;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
  // persistent step kernel: phase table, arguments, barrier words
  DeviceArray<char> phases_, pargs_;
  DeviceArray<unsigned> pflags_;  // arrival words, then the release word
  DeviceArray<int> perr_;
  int nph_ = 0, pfit_ = 0, pcus_ = 0, pmaxnvb_ = 0;
  size_t pshm_ = 0;
  unsigned pgen_ = 1;
  bool pdirty_ = true;
  bool broken_ = false;
  int last_path_ = 0;
};

// The step-path autotuner's record (process-wide): written to and merged
// from a text file, or cleared.
void batch_tune_save(const std::string& path);
void batch_tune_load(const std::string& path);
void batch_tune_clear();

}  // namespace mpa
