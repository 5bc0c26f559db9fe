// Runge-Kutta stage combination, error norm and reductions.
#pragma once

#include "common.hpp"

namespace mpa {

constexpr int MAX_STAGES = 7;

// x = x0 + sum_j c[j] * k[j] (sequential FP32 adds in j order), normalised when
// `normalize` (rsqrt of |x|^2, as k_normalize in src/core/fieldops.cu:232-247);
// zero outside the geometry.
struct Combo {
  const float* k[MAX_STAGES];
  float c[MAX_STAGES];
  int nk;
};
void launch_combine(float* out, const float* x0, const Combo& co, int ncomp, int N, const uint8_t* geom,
                    bool normalize, hipStream_t s);

// Several 3-vector combinations in one launch (the state variables of one
// Runge-Kutta stage).  out2, when set, receives a second copy of the result.
constexpr int MAX_JOBS = 8;
struct CombineJob {
  float* out;
  float* out2;
  const float* x0;
  const uint8_t* geom;
  Combo co;
  int N;
  int normalize;
};
void launch_combine_batch(const CombineJob* jobs, int njobs, hipStream_t s);

// m = ((4 - h^2 |t|^2) m0 + 4 h t) / (4 + h^2 |t|^2) per cell, the descent step
// of the energy minimiser (src/physics/minimizer.cu:59-74).
void launch_descent(float* m, const float* m0, const float* t, float h, int N, hipStream_t s);

// Reductions write into a small device workspace; results are read back by
// the caller after a stream synchronisation.
class Reducer {
 public:
  Reducer();
  // max over geometry cells of |sum_j c[j] k[j]| (3 components), FP32 norms.
  float max_combo_norm(const Combo& co, int N, const uint8_t* geom, hipStream_t s);
  float max_vec_norm(const float* v, int N, const uint8_t* geom, hipStream_t s);
  // max over geometry cells of |v| (one component; NaN cells are skipped)
  float max_abs(const float* v, int N, const uint8_t* geom, hipStream_t s);
  // sum over geometry cells of each component, accumulated in FP64.
  void component_sums(const float* v, int ncomp, int N, const uint8_t* geom, double* out, hipStream_t s);
  // sum over geometry cells of |v|^2 (3 components), FP64 (dotSum, src/core/reduce.cu:153-195).
  double dot_sum(const float* v, int N, const uint8_t* geom, hipStream_t s);
  // dm = m1 - m0, dt = t0 - t1: out = {dm.dm, dm.dt, dt.dt} summed in FP64, *max_dm = max |dm|.
  void bb_sums(const float* m1, const float* m0, const float* t1, const float* t0, int N, const uint8_t* geom,
               double out[3], float* max_dm, hipStream_t s);

 private:
  DeviceArray<unsigned> maxbits_;
  DeviceArray<double> partials_;
};

}  // namespace mpa
