// Elastodynamics kernels: strain and stress of the displacement, and the
// elastic, magnetoelastic, damping and total body forces (see elastic.hpp).
#include "elastic.hpp"
#include "elastic_cell.hpp"

namespace mpa {

namespace {

constexpr int BLOCK = 256;

using namespace elc;

__global__ void __launch_bounds__(BLOCK) k_strain(ElDesc d, const float* __restrict__ u, float* __restrict__ strain,
                                                  float* __restrict__ stress) {
  const int idx = xcd_remap(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
  const int N = d.ncell;
  if (idx >= N) return;
  if (!in_geo(d, idx)) {
    for (int c = 0; c < 6; ++c) {
      strain[(size_t)c * N + idx] = 0.f;
      if (stress) stress[(size_t)c * N + idx] = 0.f;
    }
    return;
  }
  const int x = idx % d.nx, y = (idx / d.nx) % d.ny, z = idx / (d.nx * d.ny);
  const float3 u0 = ld3(u, idx, N);
  const float3 dx = deriv(d, u, axis_nb(d, x, y, z, 0), 0, u0);  // d_x u
  const float3 dy = deriv(d, u, axis_nb(d, x, y, z, 1), 1, u0);
  const float3 dz = deriv(d, u, axis_nb(d, x, y, z, 2), 2, u0);
  // straintensor.cu:81-97: e_ii = d_i u_i, e_ij = (d_i u_j + d_j u_i)/2
  const float e[6] = {dx.x, dy.y, dz.z, 0.5f * (dx.y + dy.x), 0.5f * (dx.z + dz.x), 0.5f * (dy.z + dz.y)};
  for (int c = 0; c < 6; ++c) strain[(size_t)c * N + idx] = e[c];
  if (stress) {
    // stresstensor.cu:27-38 (2 C44 on the real shear strain)
    const float c11 = d.C11.at(idx), c12 = d.C12.at(idx), c44 = d.C44.at(idx);
    const float s[6] = {c11 * e[0] + c12 * e[1] + c12 * e[2], c11 * e[1] + c12 * e[2] + c12 * e[0],
                        c11 * e[2] + c12 * e[0] + c12 * e[1], 2.f * c44 * e[3],
                        2.f * c44 * e[4], 2.f * c44 * e[5]};
    for (int c = 0; c < 6; ++c) stress[(size_t)c * N + idx] = s[c];
  }
}

__global__ void __launch_bounds__(BLOCK) k_elastic(ElDesc d, const float* __restrict__ stress,
                                                   const float* __restrict__ v, MEForce a, MEForce b,
                                                   float* __restrict__ out, int mode) {
  const int idx = xcd_remap(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
  if (idx >= d.ncell) return;
  elastic_cell(d, stress, v, a, b, out, mode, idx);
}

template <int NK>
__global__ void __launch_bounds__(BLOCK) k_elastic_next(ElDesc d, const float* __restrict__ stress,
                                                        const float* __restrict__ v, MEForce a, MEForce b,
                                                        float* __restrict__ out, ElNext nx) {
  const int idx = xcd_remap(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
  const int N = d.ncell;
  if (idx >= N) return;
  const float3 f = elastic_term(d, stress, v, a, b, EL_ACCEL, idx);
  out[idx] = f.x;
  out[(size_t)N + idx] = f.y;
  out[(size_t)2 * N + idx] = f.z;
  elastic_next<NK>(d, nx, f, idx);
}

__global__ void __launch_bounds__(BLOCK) k_elastic_local(ElDesc d, const float* __restrict__ stress,
                                                         const float* __restrict__ strain,
                                                         const float* __restrict__ v, float* __restrict__ out,
                                                         int mode, int accumulate) {
  const int idx = blockIdx.x * BLOCK + threadIdx.x;
  const int N = d.ncell;
  if (idx >= N) return;
  const bool in = in_geo(d, idx);
  if (mode == EL_KINETIC_EDENS) {
    const float3 w = ld3(v, idx, N);
    const float e = in ? 0.5f * dot3(w, w) * d.rho.at(idx) : 0.f;
    out[idx] = accumulate ? out[idx] + e : e;
  } else if (mode == EL_ELASTIC_EDENS) {
    float e = 0.f;
    if (in)
      for (int i = 0; i < 3; ++i) {
        e += 0.5f * stress[(size_t)i * N + idx] * strain[(size_t)i * N + idx];
        e += stress[(size_t)(i + 3) * N + idx] * strain[(size_t)(i + 3) * N + idx];
      }
    out[idx] = accumulate ? out[idx] + e : e;
  } else {
    const float3 w = ld3(v, idx, N);
    for (int i = 0; i < 3; ++i) {
      const float3 row = stress_row(stress, idx, N, i);
      out[(size_t)i * N + idx] = in ? -dot3(row, w) : 0.f;
    }
  }
}

unsigned blocks(int n) { return (unsigned)((n + BLOCK - 1) / BLOCK); }

}  // namespace

void launch_strain(const ElDesc& d, const float* u, float* strain, float* stress, hipStream_t s) {
  k_strain<<<blocks(d.ncell), BLOCK, 0, s>>>(d, u, strain, stress);
  checkLaunch("k_strain");
}

void launch_elastic(const ElDesc& d, const float* stress, const float* v, const MEForce& a, const MEForce& b,
                    float* out, ElMode mode, hipStream_t s) {
  if ((mode == EL_FORCE || ((mode == EL_BODY || mode == EL_ACCEL) && d.elastic)) && !stress)
    throw Error(MPA_ERR_RUNTIME, "elastic force needs the stress tensor");
  k_elastic<<<blocks(d.ncell), BLOCK, 0, s>>>(d, stress, v, a, b, out, (int)mode);
  checkLaunch("k_elastic");
}

void launch_elastic_next(const ElDesc& d, const float* stress, const float* v, const MEForce& a, const MEForce& b,
                         float* out, const ElNext& nx, hipStream_t s) {
  if (d.elastic && !stress) throw Error(MPA_ERR_RUNTIME, "elastic force needs the stress tensor");
  const unsigned g = blocks(d.ncell);
  switch (nx.nk) {
    case 0: k_elastic_next<0><<<g, BLOCK, 0, s>>>(d, stress, v, a, b, out, nx); break;
    case 1: k_elastic_next<1><<<g, BLOCK, 0, s>>>(d, stress, v, a, b, out, nx); break;
    case 2: k_elastic_next<2><<<g, BLOCK, 0, s>>>(d, stress, v, a, b, out, nx); break;
    case 3: k_elastic_next<3><<<g, BLOCK, 0, s>>>(d, stress, v, a, b, out, nx); break;
    case 4: k_elastic_next<4><<<g, BLOCK, 0, s>>>(d, stress, v, a, b, out, nx); break;
    case 5: k_elastic_next<5><<<g, BLOCK, 0, s>>>(d, stress, v, a, b, out, nx); break;
    default: throw Error(MPA_ERR_RUNTIME, "launch_elastic_next: 0 to 5 earlier derivatives");
  }
  checkLaunch("k_elastic_next");
}

void launch_elastic_local(const ElDesc& d, const float* stress, const float* strain, const float* v, float* out,
                          ElLocal mode, hipStream_t s, bool accumulate) {
  k_elastic_local<<<blocks(d.ncell), BLOCK, 0, s>>>(d, stress, strain, v, out, (int)mode, accumulate ? 1 : 0);
  checkLaunch("k_elastic_local");
}

}  // namespace mpa
