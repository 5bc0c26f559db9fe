// Cell-level device functions of the elastodynamics kernels: the neighbour
// indices of a cell along an axis, the derivative stencils, the stress
// divergence, the magnetoelastic force and the whole body force / acceleration
// of one cell.  Shared by the elastic kernels and the combined
// sublattice-field and acceleration launch of an antiferromagnet.
#pragma once

#include "elastic.hpp"

namespace mpa {
namespace elc {

__device__ __forceinline__ float3 ld3(const float* p, int i, int N) {
  return f3(p[i], p[(size_t)N + i], p[(size_t)2 * N + i]);
}

// Local index of the cell `k` steps along `axis`, after the mastergrid wrap
// (grid.hpp:75-83); -1 outside the magnet grid.
__device__ __forceinline__ int shifted(const ElDesc& d, int x, int y, int z, int axis, int k) {
  int c[3] = {x, y, z};
  const int o[3] = {d.ox, d.oy, d.oz};
  const int mg[3] = {d.mgx, d.mgy, d.mgz};
  const int mo[3] = {d.mox, d.moy, d.moz};
  const int n[3] = {d.nx, d.ny, d.nz};
  int g = o[axis] + c[axis] + k;
  if (mg[axis] > 0) g = mo[axis] + wrap_near(g - mo[axis], mg[axis]);
  const int l = g - o[axis];
  if (l < 0 || l >= n[axis]) return -1;
  c[axis] = l;
  return (c[2] * d.ny + c[1]) * d.nx + c[0];
}

__device__ __forceinline__ bool in_geo(const ElDesc& d, int i) { return i >= 0 && (!d.geom || d.geom[i]); }

// The cells -2, -1, +1, +2 steps along an axis from a cell (after the
// mastergrid wrap; -1 outside the magnet grid) and whether each is in the
// geometry: computed once per cell and axis, shared by every derivative and
// divergence the cell takes along that axis.
struct AxisNb {
  int i[4];  // -2, -1, +1, +2
  bool g[4];
};

__device__ __forceinline__ AxisNb axis_nb(const ElDesc& d, int x, int y, int z, int axis) {
  AxisNb a;
  const int k[4] = {-2, -1, 1, 2};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    a.i[q] = shifted(d, x, y, z, axis, k[q]);
    a.g[q] = in_geo(d, a.i[q]);
  }
  return a;
}

// d f / d axis of a 3-vector field with the stencil chain of
// straintensor.cu:45-79 (shared by magnetoelasticforce.cu:42-76).  Every
// branch reads only cells it has checked to be in the geometry.
__device__ __forceinline__ float3 deriv(const ElDesc& d, const float* f, const AxisNb& nb, int axis, float3 f0) {
  const int N = d.ncell;
  const int im2 = nb.i[0], im1 = nb.i[1], ip1 = nb.i[2], ip2 = nb.i[3];
  const bool gm2 = nb.g[0], gm1 = nb.g[1], gp1 = nb.g[2], gp2 = nb.g[3];
  float3 r;
  if (!gm1 && !gp1) {
    r = f3(0.f, 0.f, 0.f);                                                   // --1--
  } else if ((!gm2 || !gp2) && gm1 && gp1) {
    r = 0.5f * (ld3(f, ip1, N) - ld3(f, im1, N));                            // -111-, 1111-, -1111
  } else if (!gm2 && !gp1) {
    r = f0 - ld3(f, im1, N);                                                 // -11--
  } else if (!gm1 && !gp2) {
    r = ld3(f, ip1, N) - f0;                                                 // --11-
  } else if (gm2 && !gp1) {
    r = 0.5f * ld3(f, im2, N) - 2.0f * ld3(f, im1, N) + 1.5f * f0;           // 111--
  } else if (!gm1 && gp1) {
    r = -0.5f * ld3(f, ip2, N) + 2.0f * ld3(f, ip1, N) - 1.5f * f0;          // --111
  } else {
    r = (2.0f / 3.0f) * (ld3(f, ip1, N) - ld3(f, im1, N)) +
        (1.0f / 12.0f) * (ld3(f, im2, N) - ld3(f, ip2, N));                 // 11111
  }
  const float c[3] = {d.cx, d.cy, d.cz};
  return (1.0f / c[axis]) * r;
}

// Row i of the stress tensor (sigma_i0, sigma_i1, sigma_i2) at cell j.
__device__ __forceinline__ float3 stress_row(const float* s, int j, int N, int i) {
  const int c0 = i == 0 ? 0 : (i == 1 ? 3 : 4);
  const int c1 = i == 0 ? 3 : (i == 1 ? 1 : 5);
  const int c2 = i == 0 ? 4 : (i == 1 ? 5 : 2);
  return f3(s[(size_t)c0 * N + j], s[(size_t)c1 * N + j], s[(size_t)c2 * N + j]);
}

// f_j = d_i sigma_ij with traction-free boundaries, elasticforce.cu:47-98.
__device__ __forceinline__ float3 stress_divergence(const ElDesc& d, const float* s, int idx, const AxisNb (&nb)[3]) {
  const int N = d.ncell;
  const float c[3] = {d.cx, d.cy, d.cz};
  float3 f = f3(0.f, 0.f, 0.f);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float w = 1.0f / c[i];
    const int im2 = nb[i].i[0], im1 = nb[i].i[1], ip1 = nb[i].i[2], ip2 = nb[i].i[3];
    const bool gm2 = nb[i].g[0], gm1 = nb[i].g[1], gp1 = nb[i].g[2], gp2 = nb[i].g[3];
    if (!gm1 && !gp1) continue;
    if (!gm1) {
      f = f + w * (stress_row(s, idx, N, i) + (1.f / 3.f) * stress_row(s, ip1, N, i));
    } else if (!gp1) {
      f = f + w * ((-1.f / 3.f) * stress_row(s, im1, N, i) - stress_row(s, idx, N, i));
    } else if (!gm2 || !gp2) {
      f = f + (0.5f * w) * (stress_row(s, ip1, N, i) - stress_row(s, im1, N, i));
    } else {
      f = f + w * ((4.f / 6.f) * (stress_row(s, ip1, N, i) - stress_row(s, im1, N, i)) +
                   (1.f / 12.f) * (stress_row(s, im2, N, i) - stress_row(s, ip2, N, i)));
    }
  }
  return f;
}

// magnetoelasticforce.cu:82-91 with der[k] = d_k m.
__device__ __forceinline__ float3 me_force(const ElDesc& d, const MEForce& a, int idx, const AxisNb (&nb)[3]) {
  const int N = d.ncell;
  const float3 m = ld3(a.m, idx, N);
  const float3 d0 = deriv(d, a.m, nb[0], 0, m), d1 = deriv(d, a.m, nb[1], 1, m), d2 = deriv(d, a.m, nb[2], 2, m);
  const float b1 = a.B1.at(idx), b2 = a.B2.at(idx);
  return f3(2.f * b1 * m.x * d0.x + b2 * m.x * (d1.y + d2.z) + b2 * (m.y * d1.x + m.z * d2.x),
            2.f * b1 * m.y * d1.y + b2 * m.y * (d2.z + d0.x) + b2 * (m.z * d2.y + m.x * d0.y),
            2.f * b1 * m.z * d2.z + b2 * m.z * (d0.x + d1.y) + b2 * (m.x * d0.z + m.y * d1.z));
}

// The selected elastic term at cell idx (< ncell; see ElMode).
__device__ __forceinline__ float3 elastic_term(const ElDesc& d, const float* __restrict__ stress,
                                               const float* __restrict__ v, const MEForce& a, const MEForce& b,
                                               int mode, int idx) {
  const int N = d.ncell;
  float3 f = f3(0.f, 0.f, 0.f);
  if (in_geo(d, idx)) {
    const int x = idx % d.nx, y = (idx / d.nx) % d.ny, z = idx / (d.nx * d.ny);
    AxisNb nb[3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) nb[ax] = axis_nb(d, x, y, z, ax);
    const bool body = mode == EL_BODY || mode == EL_ACCEL;
    if (mode == EL_FORCE || (body && d.elastic)) f = f + stress_divergence(d, stress, idx, nb);
    if (body && d.ext) f = f + d.fext.at(idx, N);
    if ((mode == EL_MEFORCE || body) && a.m) f = f + me_force(d, a, idx, nb);
    if (body && b.m) f = f + me_force(d, b, idx, nb);
    if (mode == EL_DAMPING || (mode == EL_ACCEL && d.damping)) f = f + (-d.eta.at(idx)) * ld3(v, idx, N);
    if (mode == EL_ACCEL && d.divrho) {
      const float p = d.rho.at(idx);
      f = p != 0.f ? (1.f / p) * f : f3(0.f, 0.f, 0.f);
    }
  }
  return f;
}

// The selected elastic term at cell idx (< ncell), written to out
// (k_elastic's cell).
__device__ __forceinline__ void elastic_cell(const ElDesc& d, const float* __restrict__ stress,
                                             const float* __restrict__ v, const MEForce& a, const MEForce& b,
                                             float* __restrict__ out, int mode, int idx) {
  const int N = d.ncell;
  const float3 f = elastic_term(d, stress, v, a, b, mode, idx);
  out[idx] = f.x;
  out[(size_t)N + idx] = f.y;
  out[(size_t)2 * N + idx] = f.z;
}

// The next stage state of (u, v) at cell idx from the acceleration f just
// computed there (ElNext); NK = nx.nk, known at compile time so that every
// load is issued before the sums.
template <int NK>
__device__ __forceinline__ void elastic_next(const ElDesc& d, const ElNext& nx, float3 f, int idx) {
  const int N = d.ncell;
  const bool in = in_geo(d, idx);
  const float fc[3] = {f.x, f.y, f.z};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const size_t i = (size_t)c * N + idx;
    float av = nx.v0[i];
#pragma unroll
    for (int j = 0; j < NK; ++j) av += nx.c[j] * nx.ka[j][i];
    av += nx.cself * fc[c];
    float au = nx.u0[i];
#pragma unroll
    for (int j = 0; j < NK; ++j) au += nx.c[j] * nx.ku[j][i];
    au += nx.cself * nx.ku[NK][i];
    const float rv = in ? av : 0.f, ru = in ? au : 0.f;
    nx.vout[i] = rv;
    nx.kdisp[i] = rv;
    nx.uout[i] = ru;
  }
}

}  // namespace elc
}  // namespace mpa
