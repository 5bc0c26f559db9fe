// Electrical potential of a ferromagnet: the finite-volume system of
// div(sigma grad V) = 0 with applied potentials as fixed cells
// (PoissonSystem, src/physics/poissonsystem.cu:134-307) and its iterative
// solvers (LinSolver, src/linsolver/linsolver.cpp: Jacobi, conjugate
// gradient, minimal residual, steepest descent).
//
// The matrix is stored in ELL order (entry e of row i at e * nrows + i, a
// column of -1 when unused) with every row divided by its diagonal, as the
// reference builds it; the vectors and the matrix are FP64.
#pragma once

#include <string>

#include "common.hpp"

namespace mpa {

// Inputs of the row construction, per cell: the conductivity (one value, or
// the six tensor components xx yy zz xy xz yz with AMR), the applied
// potential (NaN where none) and the geometry.
struct PoissonDesc {
  int nx, ny, nz;
  double cx, cy, cz;
  const uint8_t* geom;  // nullptr: every cell
  PS sigma0;            // conductivity
  PS amr;               // AMR ratio
  const float* m;       // magnetisation (3N), used when aniso
  int aniso;            // the AMR ratio is not assured zero
  PS pot;               // applied potential
};

// The conductivity tensor with AMR (conductivitytensor.cu:11-38), 6N.
void launch_conductivity_tensor(const PoissonDesc& d, float* out, hipStream_t s);

class PoissonSolver {
 public:
  enum Method { JACOBI = 0, CONJUGATEGRADIENT = 1, MINIMALRESIDUAL = 2, STEEPESTDESCENT = 3 };
  int max_iter = -1;  // no limit when negative
  double tol = 1e-5;  // on the max norm of the residual

  // LinSolver::getMethodByName (linsolver.cpp); throws MPA_ERR_INVALID
  void set_method(const std::string& name);
  // PoissonSystem::init: build the rows; x is reset (and the stepper
  // restarted) only when the number of rows changes (LinSolver::setSystem).
  void init(const PoissonDesc& d, hipStream_t s);
  void restart(hipStream_t s);
  void step(hipStream_t s);
  double residual_max_norm(hipStream_t s);
  // LinSolver::solve: step while the residual max norm exceeds tol and the
  // iteration limit allows; the solution is x().
  void solve(hipStream_t s);
  const double* x() const { return x_.get(); }
  int rows() const { return nrows_; }
  bool ready() const { return nrows_ > 0; }

 private:
  void apply(double* y, const double* v, double ka, double kb, hipStream_t s);  // y = ka A v + kb b
  double dot(const double* a, const double* b, hipStream_t s);
  void axpby(double* out, double a, const double* x, double b, const double* y, hipStream_t s);
  Method method_ = JACOBI;
  int nrows_ = 0, nnz_ = 0;
  DeviceArray<int> idx_;
  DeviceArray<double> val_, b_, x_, r_, p_, t_, partial_;
  double rr_ = 0.0;
};

}  // namespace mpa
