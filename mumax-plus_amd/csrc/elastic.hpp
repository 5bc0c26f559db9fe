// Elastodynamics and magnetoelastic body force of a magnet.
//
// Replaces the strain (src/physics/straintensor.cu:13-97), stress
// (stresstensor.cu:9-40), elastic force (elasticforce.cu:25-101),
// magnetoelastic force (magnetoelasticforce.cu:10-94), elastic damping
// (elasticdamping.cu:13-28) and acceleration (elastodynamics.cu:23-84)
// kernels and their chain of temporaries: one kernel computes strain and
// stress from the displacement, one computes any of the force terms (or the
// whole acceleration) from the stress, velocity and magnetisations.
#pragma once

#include "common.hpp"

namespace mpa {

struct ElDesc {
  int nx, ny, nz, ncell;
  int ox, oy, oz;
  int mgx, mgy, mgz;
  int mox, moy, moz;
  float cx, cy, cz;
  const uint8_t* geom;
  PS C11, C12, C44, eta, rho;
  PV fext;
  int elastic;  // elastic force enabled (elastodynamics on, some C non-zero)
  int ext;      // external body force non-zero
  int damping;  // -eta v enabled
  int divrho;   // divide by rho (rho not assured zero)
};

// A ferromagnet (or sublattice) whose magnetoelastic force acts on the magnet.
struct MEForce {
  const float* m;  // nullptr: none
  PS B1, B2;
};

enum ElMode {
  EL_FORCE = 0,    // elastic force only
  EL_MEFORCE = 1,  // magnetoelastic force of `a` only
  EL_DAMPING = 2,  // -eta v
  EL_BODY = 3,     // elastic + external + magnetoelastic (a, b)
  EL_ACCEL = 4,    // (body force + damping) / rho
};

// The next Runge-Kutta stage state of (u, v), formed in the acceleration
// launch of stage s (k_combine_batch's arithmetic and term order):
// v' = v0 + sum_{j<nk} c_j a_j + cself a_s with a_s the acceleration just
// computed, u' = u0 + sum_{j<nk} c_j w_j + cself w_s with w_j the stage
// velocities (the displacement's derivatives, ku[0..nk]), both zero outside
// the geometry; v' goes to vout and to kdisp (the displacement's derivative
// of the next stage).
struct ElNext {
  const float* v0;
  const float* u0;
  const float* ka[7];
  const float* ku[8];
  float c[7];
  float cself;
  int nk;
  float* vout;
  float* uout;
  float* kdisp;
};

// strain (6N: xx yy zz xy xz yz) of displacement u; stress (6N) as well when
// stress != nullptr.
void launch_strain(const ElDesc& d, const float* u, float* strain, float* stress, hipStream_t s);
// out (3N) = the selected term.  stress may be nullptr when d.elastic == 0.
void launch_elastic(const ElDesc& d, const float* stress, const float* v, const MEForce& a, const MEForce& b,
                    float* out, ElMode mode, hipStream_t s);

// launch_elastic in EL_ACCEL mode that also writes the next stage state of
// (u, v) (ElNext).
void launch_elastic_next(const ElDesc& d, const float* stress, const float* v, const MEForce& a, const MEForce& b,
                         float* out, const ElNext& nx, hipStream_t s);

enum ElLocal {
  EL_KINETIC_EDENS = 0,  // 0.5 rho |v|^2 (elasticenergies.cu:7-23), 1 component
  EL_ELASTIC_EDENS = 1,  // 0.5 sum sigma_ii e_ii + sum sigma_ij e_ij (:58-78), 1 component
  EL_POYNTING = 2,       // -sigma v (poyntingvector.cu:10-30), 3 components
};
// Cell-local energy densities and energy flux; zero outside the geometry.
// With accumulate, an energy density is added to out instead of stored.
void launch_elastic_local(const ElDesc& d, const float* stress, const float* strain, const float* v, float* out,
                          ElLocal mode, hipStream_t s, bool accumulate = false);

}  // namespace mpa
