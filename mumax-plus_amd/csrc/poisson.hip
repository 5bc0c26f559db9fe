// Electrical potential: row construction and the iterative solvers.
#include "poisson.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace mpa {
namespace {

constexpr int BLOCK = 256;
constexpr int SUM_BLOCKS = 1024;

__device__ __forceinline__ bool in_geom(const PoissonDesc& d, int x, int y, int z) {
  if (x < 0 || y < 0 || z < 0 || x >= d.nx || y >= d.ny || z >= d.nz) return false;
  return !d.geom || d.geom[(z * d.ny + y) * d.nx + x];
}

// Component c of the conductivity of cell i: sigma0 (c = 0, isotropic), or
// the AMR tensor xx yy zz xy xz yz (conductivitytensor.cu:11-32).
__device__ __forceinline__ double sigma_at(const PoissonDesc& d, int i, int c) {
  const double s0 = d.sigma0.at(i);
  if (!d.aniso) return s0;
  const int N = d.nx * d.ny * d.nz;
  const double amr = d.amr.at(i);
  const double f = 6.0 * amr / (6.0 + amr);
  const double m[3] = {d.m[i], d.m[N + i], d.m[2 * N + i]};
  if (c < 3) return s0 * (1.0 - f * (m[c] * m[c] - 1.0 / 3.0));
  const int a = c == 5 ? 1 : 0, b = c == 3 ? 1 : 2;
  return s0 * f * m[a] * m[b];
}

__global__ void k_conductivity_tensor(PoissonDesc d, float* __restrict__ out) {
  const int N = d.nx * d.ny * d.nz;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const bool in = !d.geom || d.geom[i];
  for (int c = 0; c < 6; ++c) out[(size_t)c * N + i] = in ? (float)sigma_at(d, i, c) : 0.f;
}

// One row per cell (poissonsystem.cu:134-258): the finite-volume balance of
// the currents through the faces to in-geometry neighbours, the cross terms
// of an anisotropic tensor spread over the four cells around each face,
// divided by the diagonal; an identity row (b = applied potential, or 0) for
// cells outside the geometry, without conductivity or with an applied
// potential.
__global__ void k_rows(PoissonDesc d, int nnz, int* __restrict__ idx, double* __restrict__ val,
                       double* __restrict__ b) {
  const int N = d.nx * d.ny * d.nz;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int x = i % d.nx, y = (i / d.nx) % d.ny, z = i / (d.nx * d.ny);
  // the 27 cells around (x, y, z), slot (dx+1) + 3 (dy+1) + 9 (dz+1)
  double a[27];
  for (int k = 0; k < 27; ++k) a[k] = 0.0;
  auto slot = [](int dx, int dy, int dz) { return (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1); };
  auto inside = [&](int dx, int dy, int dz) { return in_geom(d, x + dx, y + dy, z + dz); };
  // a difference V(p) - V(q) with weight w, kept when both cells take part
  auto diff = [&](const int* p, const int* q, double w) {
    if (inside(p[0], p[1], p[2]) && inside(q[0], q[1], q[2])) {
      a[slot(p[0], p[1], p[2])] += w;
      a[slot(q[0], q[1], q[2])] -= w;
    }
  };
  const double pot = d.pot.at(i);
  double rhs = 0.0;
  if (!inside(0, 0, 0) || sigma_at(d, i, 0) == 0.0 || !isnan(pot)) {
    a[slot(0, 0, 0)] = 1.0;
    rhs = isnan(pot) ? 0.0 : pot;
  } else {
    const double cs[3] = {d.cx, d.cy, d.cz};
    for (int ax = 0; ax < 3; ++ax) {
      for (int sgn = -1; sgn <= 1; sgn += 2) {
        int e[3] = {0, 0, 0};
        e[ax] = sgn;
        if (!inside(e[0], e[1], e[2])) continue;
        const int j = (z + e[2]) * d.nx * d.ny + (y + e[1]) * d.nx + (x + e[0]);
        const double area = cs[(ax + 1) % 3] * cs[(ax + 2) % 3];
        const int o3[3] = {0, 0, 0};
        // geometric mean of the two cells' conductivities
        const double saa = sqrt(sigma_at(d, i, d.aniso ? ax : 0) * sigma_at(d, j, d.aniso ? ax : 0));
        diff(o3, e, area * saa / cs[ax]);
        if (!d.aniso) continue;
        for (int ot = 0; ot < 3; ++ot) {
          if (ot == ax) continue;
          const int c = (ax + ot == 1) ? 3 : ((ax + ot == 2) ? 4 : 5);  // xy, xz, yz
          const double sao = sqrt(sigma_at(d, i, c) * sigma_at(d, j, c));
          const double w = -sgn * area * sao / cs[ot] / 4.0;
          int mo[3] = {0, 0, 0}, po[3] = {0, 0, 0}, em[3], ep[3];
          mo[ot] = -1;
          po[ot] = 1;
          for (int k = 0; k < 3; ++k) {
            em[k] = e[k] + mo[k];
            ep[k] = e[k] + po[k];
          }
          diff(o3, mo, w);
          diff(e, em, w);
          diff(po, o3, w);
          diff(ep, e, w);
        }
      }
    }
  }
  const double diag = a[slot(0, 0, 0)];
  int c = 0;
  for (int k = 0; k < 27 && c < nnz; ++k) {
    if (a[k] == 0.0) continue;
    const int dx = k % 3 - 1, dy = (k / 3) % 3 - 1, dz = k / 9 - 1;
    idx[(size_t)c * N + i] = (z + dz) * d.nx * d.ny + (y + dy) * d.nx + (x + dx);
    val[(size_t)c * N + i] = a[k] / diag;
    ++c;
  }
  for (; c < nnz; ++c) idx[(size_t)c * N + i] = -1;
  b[i] = rhs;
}

// y = ka A v + kb b (linsystem.cu:14-34)
__global__ void k_apply(int n, int nnz, const int* __restrict__ idx, const double* __restrict__ val,
                        const double* __restrict__ b, const double* __restrict__ v, double ka, double kb,
                        double* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int e = 0; e < nnz; ++e) {
    const int j = idx[(size_t)e * n + i];
    if (j >= 0) s += val[(size_t)e * n + i] * v[j];
  }
  y[i] = ka * s + kb * b[i];
}

// out may be x or y
__global__ void k_axpby(int n, double a, const double* x, double b, const double* y, double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a * x[i] + b * y[i];
}

__device__ __forceinline__ double block_reduce(double v, bool take_max) {
  __shared__ double sh[BLOCK / 64];
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_down(v, off, 64);
    v = take_max ? fmax(v, o) : v + o;
  }
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int k = 1; k < BLOCK / 64; ++k) v = take_max ? fmax(v, sh[k]) : v + sh[k];
  return v;
}

// per block: sum of a b, or the max of |a| (b == nullptr)
__global__ void __launch_bounds__(BLOCK) k_partial(int n, const double* __restrict__ a, const double* __restrict__ b,
                                                   double* __restrict__ out) {
  const bool mx = b == nullptr;
  double v = 0.0;
  for (int i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK)
    v = mx ? fmax(v, fabs(a[i])) : v + a[i] * b[i];
  v = block_reduce(v, mx);
  if (threadIdx.x == 0) out[blockIdx.x] = v;
}

int blocks(int n) { return (n + BLOCK - 1) / BLOCK; }

}  // namespace

void launch_conductivity_tensor(const PoissonDesc& d, float* out, hipStream_t s) {
  const int N = d.nx * d.ny * d.nz;
  k_conductivity_tensor<<<blocks(N), BLOCK, 0, s>>>(d, out);
  checkLaunch("k_conductivity_tensor");
}

void PoissonSolver::set_method(const std::string& name) {
  // LinSolver::getMethodByName
  if (name == "jacobi") method_ = JACOBI;
  else if (name == "conjugategradient") method_ = CONJUGATEGRADIENT;
  else if (name == "minimalresidual") method_ = MINIMALRESIDUAL;
  else if (name == "steepestdescent") method_ = STEEPESTDESCENT;
  else throw Error(MPA_ERR_INVALID, "Linear system solver method '" + name + "' is not implemented");
}

void PoissonSolver::init(const PoissonDesc& d, hipStream_t s) {
  const int N = d.nx * d.ny * d.nz;
  const bool three_d = d.nz > 1;
  // poissonsystem.cu:262-283: 19 / 9 entries with a tensor, 7 / 5 without
  const int nnz = d.aniso ? (three_d ? 19 : 9) : (three_d ? 7 : 5);
  if (nnz_ != nnz || idx_.n != (size_t)nnz * N) {
    idx_.allocate((size_t)nnz * N);
    val_.allocate((size_t)nnz * N);
  }
  nnz_ = nnz;
  if (b_.n != (size_t)N) b_.allocate(N);
  k_rows<<<blocks(N), BLOCK, 0, s>>>(d, nnz, idx_.get(), val_.get(), b_.get());
  checkLaunch("k_rows");
  if (nrows_ != N) {
    nrows_ = N;
    x_.allocate(N);
    r_.allocate(N);
    p_.allocate(N);
    t_.allocate(N);
    partial_.allocate(SUM_BLOCKS);
    // LinSolver::resetState: x = 0 and a fresh stepper
    HIP_CHECK(hipMemsetAsync(x_.get(), 0, (size_t)N * sizeof(double), s));
    restart(s);
  }
}

void PoissonSolver::apply(double* y, const double* v, double ka, double kb, hipStream_t s) {
  k_apply<<<blocks(nrows_), BLOCK, 0, s>>>(nrows_, nnz_, idx_.get(), val_.get(), b_.get(), v, ka, kb, y);
  checkLaunch("k_apply");
}

void PoissonSolver::axpby(double* out, double a, const double* x, double b, const double* y, hipStream_t s) {
  k_axpby<<<blocks(nrows_), BLOCK, 0, s>>>(nrows_, a, x, b, y, out);
  checkLaunch("k_axpby");
}

double PoissonSolver::dot(const double* a, const double* b, hipStream_t s) {
  const int nb = std::min(SUM_BLOCKS, std::max(1, blocks(nrows_)));
  k_partial<<<nb, BLOCK, 0, s>>>(nrows_, a, b, partial_.get());
  checkLaunch("k_partial");
  std::vector<double> h(nb);
  HIP_CHECK(hipMemcpyAsync(h.data(), partial_.get(), nb * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  double r = 0.0;
  if (b) {
    for (double v : h) r += v;
  } else {
    for (double v : h) r = std::fmax(r, v);
  }
  return r;
}

double PoissonSolver::residual_max_norm(hipStream_t s) {
  if (!ready()) return 0.0;
  apply(t_.get(), x_.get(), -1.0, 1.0, s);  // b - A x
  return dot(t_.get(), nullptr, s);
}

// Stepper::restart of each method (linsolver.cpp)
void PoissonSolver::restart(hipStream_t s) {
  if (!ready() || method_ == JACOBI) return;
  apply(r_.get(), x_.get(), -1.0, 1.0, s);  // r = b - A x
  if (method_ == CONJUGATEGRADIENT) {
    HIP_CHECK(hipMemcpyAsync(p_.get(), r_.get(), (size_t)nrows_ * sizeof(double), hipMemcpyDeviceToDevice, s));
  } else {
    apply(p_.get(), r_.get(), 1.0, 0.0, s);  // p = A r
  }
  rr_ = dot(r_.get(), r_.get(), s);
}

// Stepper::step of each method (linsolver.cpp)
void PoissonSolver::step(hipStream_t s) {
  if (!ready()) return;
  double* x = x_.get();
  double* r = r_.get();
  double* p = p_.get();
  double* t = t_.get();
  switch (method_) {
    case JACOBI:
      // x = x + (b - A x): the rows are divided by their diagonal
      apply(t, x, -1.0, 1.0, s);
      axpby(x, 1.0, x, 1.0, t, s);
      break;
    case CONJUGATEGRADIENT: {
      apply(t, p, 1.0, 0.0, s);  // t = A p
      const double alpha = rr_ / dot(p, t, s);
      axpby(x, 1.0, x, alpha, p, s);
      axpby(r, 1.0, r, -alpha, t, s);
      const double rr_old = rr_;
      rr_ = dot(r, r, s);
      axpby(p, 1.0, r, rr_ / rr_old, p, s);
      break;
    }
    case MINIMALRESIDUAL: {
      const double alpha = rr_ / dot(p, p, s);
      axpby(x, 1.0, x, alpha, r, s);
      axpby(r, 1.0, r, -alpha, p, s);
      apply(p, r, 1.0, 0.0, s);
      rr_ = dot(r, r, s);
      break;
    }
    case STEEPESTDESCENT: {
      const double alpha = rr_ / dot(p, r, s);
      axpby(x, 1.0, x, alpha, r, s);
      axpby(r, 1.0, r, -alpha, p, s);
      apply(p, r, 1.0, 0.0, s);
      rr_ = dot(r, r, s);
      break;
    }
  }
}

void PoissonSolver::solve(hipStream_t s) {
  int n = 0;
  while (residual_max_norm(s) > tol) {
    if (max_iter >= 0 && n > max_iter) break;
    step(s);
    ++n;
  }
}

}  // namespace mpa
