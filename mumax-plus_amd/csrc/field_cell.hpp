// Device code shared by the field kernels and the fused x-inverse FFT pass:
// the uniform-parameter field and torque of one cell, and its helpers.
#pragma once

#include "batch.hpp"
#include "fields.hpp"

namespace mpa {

constexpr float GAMMALL_F = 1.7595e11f;  // src/physics/constants.hpp:7

__device__ __forceinline__ float3 load3(const float* __restrict__ v, int i, int n) {
  return f3(v[i], v[n + i], v[2 * n + i]);
}

// getGamma (dmi.hpp:53-69) with the absolute normal n along `axis`.
__device__ __forceinline__ float3 dmi_gamma(const float D[D_COUNT], int axis, float3 m) {
  const float nx = axis == 0, ny = axis == 1, nz = axis == 2;
  return f3(-D[D_XXY] * nx * m.y - D[D_XXZ] * nx * m.z - D[D_YXZ] * ny * m.z - D[D_ZXY] * nz * m.y -
                D[D_YXY] * ny * m.y - D[D_ZXZ] * nz * m.z,
            D[D_XXY] * nx * m.x - D[D_ZYZ] * nz * m.z + D[D_YXY] * ny * m.x - D[D_XYZ] * nx * m.z +
                D[D_ZXY] * nz * m.x - D[D_YYZ] * ny * m.z,
            D[D_XXZ] * nx * m.x + D[D_YYZ] * ny * m.y + D[D_XYZ] * nx * m.y + D[D_YXZ] * ny * m.x +
                D[D_ZXZ] * nz * m.x + D[D_ZYZ] * nz * m.y);
}

// Local index of the neighbour in direction dir (0:-x 1:+x 2:-y 3:+y 4:-z 5:+z)
// after the mastergrid wrap, or -1 when it falls outside the magnet grid.
__device__ __forceinline__ int neighbour(const FMDesc& d, int x, int y, int z, int dir) {
  const int axis = dir >> 1;
  const int step = (dir & 1) ? 1 : -1;
  int c[3] = {x, y, z};
  const int o[3] = {d.ox, d.oy, d.oz};
  const int mg[3] = {d.mgx, d.mgy, d.mgz};
  const int mo[3] = {d.mox, d.moy, d.moz};
  const int n[3] = {d.nx, d.ny, d.nz};
  int g = o[axis] + c[axis] + step;
  if (mg[axis] > 0) g = mo[axis] + wrap_near(g - mo[axis], mg[axis]);
  const int l = g - o[axis];
  if (l < 0 || l >= n[axis]) return -1;
  c[axis] = l;
  return (c[2] * d.ny + c[1]) * d.nx + c[0];
}

// ---------------------------------------------------------------------------
// Uniform-parameter path.  When every material parameter is one value for
// the whole magnet, the per-neighbour harmonic means reduce to that value
// and all divisions fold into host-side coefficients:
//   exchange   2A/(Ms c^2) (m_n - m);  Neumann boundary: s Gamma/(Ms c)
//   DMI        s (D/(Ms c)) . m_n      with m_n = m + Gamma s c/(2A) at a
//                                      Neumann boundary (s = +-1 the side)
//   anisotropy (2K1/Ms mu + 4K2/Ms mu^3) u,  cubic with -2Kc1/Ms, ...
// The branch structure (open boundaries, geometry holes, msat = 0 never
// occurs since msat is uniform and non-zero here) is the same as above.
struct FMUniform {
  float kex[3];      // 2A/(Ms c^2) per axis
  float inv_msc[3];  // 1/(Ms c)
  float dxz[3], dxy[3], dyz[3];  // D_(axis,xz|xy|yz)/(Ms c)
  float D[D_COUNT];  // raw tensor for the boundary term Gamma
  float ncoef[3];    // c/(2A) (Neumann displacement), 0 when A == 0
  float u[3], ku1, ku2;
  float c1[3], c2[3], c3[3], kc1, kc2, kc3;
  float hext[3];
  float llg, alpha;  // -gamma/(1+alpha^2)
  int has_ex, has_dmi, uni, cubic, demag, frozen, openbc_dmi, openbc;
  // one layer (Nz == 1) and no DMI entries along z: the exchange and DMI
  // terms of the two missing (or self-wrapped) z neighbours cancel
  int zflat;
};

__device__ __forceinline__ float3 gamma_u(const FMUniform& u, int axis, float3 m) { return dmi_gamma(u.D, axis, m); }

// What a field kernel does with a cell's result besides storing it: nothing,
// or (NextStage) the next Runge-Kutta stage state of the cell.
struct NoEpilogue {
  __device__ __forceinline__ void operator()(int, float3, float3, bool) const {}
};

// The field (mode MODE_FIELD) or torque of one cell on the uniform path from
// its own value mc and its six neighbour values mn[dir] (dir 0:-x 1:+x 2:-y
// 3:+y 4:-z 5:+z).  ok[dir] is false for a missing neighbour (outside the
// grid or the geometry), whose mn is not read.  Only the first ndir
// directions are visited (4 skips the z axis of a flat layer).  hs: the stray
// field of the other magnets (added when has_stray); hd: the demag field
// (added when u.demag).  Term order: exchange, anisotropy, external, stray,
// DMI, demag.
// FAST: the configuration uniform_fast() accepts, with its switches fixed at
// compile time (exchange and DMI on, Neumann boundaries, uniaxial
// anisotropy, external field, demag, no stray field, LLG torque, not
// frozen); zero DMI or anisotropy constants then add exact zeros, and the
// boundary terms are taken through ghost neighbours (same terms, rounded in
// another order).
template <bool FAST = false>
__device__ __forceinline__ float3 uniform_result(const FMUniform& u, int mode, float3 mc, const float3 (&mn)[6],
                                                 const bool (&ok)[6], int ndir, bool has_stray, float3 hs,
                                                 float3 hd) {
  const bool has_ex = FAST || u.has_ex, has_dmi = FAST || u.has_dmi;
  const bool openbc = !FAST && u.openbc, openbc_dmi = !FAST && u.openbc_dmi;
  const bool uni = FAST || u.uni, cubic = !FAST && u.cubic, demag = FAST || u.demag;
  float3 hex = f3(0.f, 0.f, 0.f), hdm = f3(0.f, 0.f, 0.f);
  // DMI field of neighbour value v along `axis`, before the direction sign
  auto dmi_of = [&](int axis, float3 v) {
    return f3(u.dxy[axis] * v.y + u.dxz[axis] * v.z, -u.dxy[axis] * v.x + u.dyz[axis] * v.z,
              -u.dxz[axis] * v.x - u.dyz[axis] * v.y);
  };
  if constexpr (FAST) {
    // Neumann boundaries as ghost values: the boundary terms below (exchange
    // s Gamma/(Ms c), DMI s D.(m + s c Gamma/(2A))) are the interior formula
    // with the missing neighbour replaced by m + s (c/(2A)) Gamma(m).  Only
    // the lanes at a boundary compute a ghost; then every cell takes the
    // same branch-free two-sided formula.
    float3 nb[6];
#pragma unroll
    for (int dir = 0; dir < 6; ++dir) {
      nb[dir] = mn[dir];
      if (dir < ndir && !ok[dir]) {
        const int axis = dir >> 1;
        const float sg = (dir & 1) ? 1.f : -1.f;
        nb[dir] = mc + (sg * u.ncoef[axis]) * gamma_u(u, axis, mc);
      }
    }
#pragma unroll
    for (int axis = 0; axis < 3; ++axis) {
      if (2 * axis >= ndir) break;
      const float3 mm = nb[2 * axis], mp = nb[2 * axis + 1];
      hex = hex + u.kex[axis] * (mm + mp - 2.f * mc);
      hdm = hdm + dmi_of(axis, mp - mm);
    }
  }
  // Missing or outside-geometry neighbour: Neumann with the DMI correction,
  // or the open boundary.
  auto boundary = [&](int axis, float s) {
    const float3 gam = gamma_u(u, axis, mc);
    if (has_ex && !openbc) hex = hex + (s * u.inv_msc[axis]) * gam;
    if (has_dmi && !openbc_dmi) hdm = hdm + s * dmi_of(axis, mc + (s * u.ncoef[axis]) * gam);
  };
  if (!FAST && (has_ex || has_dmi)) {
#pragma unroll
    for (int axis = 0; axis < 3; ++axis) {
      if (2 * axis >= ndir) break;
      const float3 mm = mn[2 * axis], mp = mn[2 * axis + 1];
      if (ok[2 * axis] && ok[2 * axis + 1]) {
        // both neighbours plain: kex (m+ + m- - 2m) and D (m+ - m-)
        if (has_ex) hex = hex + u.kex[axis] * (mm + mp - 2.f * mc);
        if (has_dmi) hdm = hdm + dmi_of(axis, mp - mm);
      } else {
        if (ok[2 * axis]) {
          if (has_ex) hex = hex + u.kex[axis] * (mm - mc);
          if (has_dmi) hdm = hdm - dmi_of(axis, mm);
        } else {
          boundary(axis, -1.f);
        }
        if (ok[2 * axis + 1]) {
          if (has_ex) hex = hex + u.kex[axis] * (mp - mc);
          if (has_dmi) hdm = hdm + dmi_of(axis, mp);
        } else {
          boundary(axis, 1.f);
        }
      }
    }
  }
  float3 han = f3(0.f, 0.f, 0.f);
  if (uni) {
    const float3 ua = f3(u.u[0], u.u[1], u.u[2]);
    const float mu = dot3(mc, ua);
    han = (u.ku1 * mu + u.ku2 * mu * mu * mu) * ua;
  } else if (cubic) {
    const float3 c1 = f3(u.c1[0], u.c1[1], u.c1[2]), c2 = f3(u.c2[0], u.c2[1], u.c2[2]),
                 c3 = f3(u.c3[0], u.c3[1], u.c3[2]);
    const float a1 = dot3(c1, mc), a2 = dot3(c2, mc), a3 = dot3(c3, mc);
    const float s1 = a1 * a1, s2 = a2 * a2, s3 = a3 * a3;
    han = u.kc1 * ((s2 + s3) * a1 * c1 + (s1 + s3) * a2 * c2 + (s1 + s2) * a3 * c3) +
          u.kc2 * ((s2 * s3) * a1 * c1 + (s1 * s3) * a2 * c2 + (s1 * s2) * a3 * c3) +
          u.kc3 * ((s2 * s2 + s3 * s3) * (s1 * a1) * c1 + (s1 * s1 + s3 * s3) * (s2 * a2) * c2 +
                   (s1 * s1 + s2 * s2) * (s3 * a3) * c3);
  }
  float3 h = hex;
  h = h + han;
  h = h + f3(u.hext[0], u.hext[1], u.hext[2]);
  if (!FAST && has_stray) h = h + hs;
  h = h + hdm;
  if (demag) h = h + hd;
  if (!FAST && mode == MODE_FIELD) return h;
  if (!FAST && u.frozen) return f3(0.f, 0.f, 0.f);
  if (FAST || mode == MODE_LLG) {
    const float3 mxh = cross3(mc, h);
    return u.llg * (mxh + u.alpha * cross3(mc, mxh));
  }
  return -GAMMALL_F * cross3(mc, cross3(mc, h));
}

// One cell (x, y, z) of the uniform path; m and out have component stride N
// (the magnet's cell count, or the packed length of a batch); hdem(idx) gives
// the demag field of the cell (read from memory, or the x-inverse FFT pass's
// own registers when the field is computed inside that pass).
template <class HDem, class Epi = NoEpilogue>
__device__ __forceinline__ void field_uniform_cell(const FMDesc& d, const FMUniform& u, const float* __restrict__ m,
                                                   const HDem& hdem, float* __restrict__ out, int mode, int x, int y,
                                                   int z, int N, const Epi& epi = Epi()) {
  const int idx = (z * d.ny + y) * d.nx + x;
  if (d.geom && !d.geom[idx]) {
    out[idx] = 0.f;
    out[N + idx] = 0.f;
    out[2 * N + idx] = 0.f;
    epi(idx, f3(0.f, 0.f, 0.f), f3(0.f, 0.f, 0.f), false);
    return;
  }
  const float3 mc = load3(m, idx, N);
  // Every neighbour value is loaded before any is used (a missing neighbour
  // reads the cell itself and is not used), so the six loads are in flight
  // together; the wrap, geometry and boundary rules then act on registers.
  // A single layer without z-axis DMI entries skips z: its two terms cancel.
  const int ndir = u.zflat ? 4 : 6;
  bool ok[6];
  float3 mn[6];
  if (u.has_ex || u.has_dmi) {
    int nbi[6];
#pragma unroll
    for (int dir = 0; dir < 6; ++dir) {
      if (dir < ndir) {
        const int nb = neighbour(d, x, y, z, dir);
        ok[dir] = nb >= 0 && (!d.geom || d.geom[nb]);
        nbi[dir] = nb >= 0 ? nb : idx;
      } else {
        ok[dir] = false;
        nbi[dir] = idx;
      }
    }
#pragma unroll
    for (int dir = 0; dir < 6; ++dir) mn[dir] = dir < ndir ? load3(m, nbi[dir], N) : mc;
  } else {
#pragma unroll
    for (int dir = 0; dir < 6; ++dir) {
      ok[dir] = false;
      mn[dir] = mc;
    }
  }
  const float3 hs = d.hstray ? load3(d.hstray, idx, d.ncell) : f3(0.f, 0.f, 0.f);
  const float3 hd = u.demag ? hdem(idx) : f3(0.f, 0.f, 0.f);
  const float3 r = uniform_result(u, mode, mc, mn, ok, ndir, d.hstray != nullptr, hs, hd);
  out[idx] = r.x;
  out[N + idx] = r.y;
  out[2 * N + idx] = r.z;
  epi(idx, mc, r, true);
}


// A member of a batched field launch: its description, coefficients, first
// cell in the packed arrays and first block of the launch.
struct FBInst {
  FMDesc d;
  FMUniform u;
  int off;
  int first;
  int mode;
  int slot;
};

// Epilogue of the batched field kernel: the member's next stage state at
// the cell (BatchNext), from the derivative r in registers.  c[j] = dt a_j.
// NK > 0: n.nk == NK known at compile time, so every load is issued before
// the sums (NK < 0: n.nk read at run time).
template <int NK>
struct BatchNextCell {
  const BatchNext& n;
  const float* c;
  int off, cs;
  __device__ __forceinline__ void operator()(int idx, float3 mc, float3 r, bool in) const {
    const size_t p = (size_t)off + idx;
    const float rr[3] = {r.x, r.y, r.z};
    const int nk = NK > 0 ? NK : n.nk;
    float v[3];
#pragma unroll
    for (int comp = 0; comp < 3; ++comp) {
      const size_t i = (size_t)comp * cs + p;
      float a = n.x0[i];
#pragma unroll
      for (int j = 0; j < (NK > 0 ? NK : 7); ++j) {
        if (NK < 0 && j >= nk) break;
        a += c[j] * n.k[j][i];
      }
      a += c[nk] * rr[comp];
      v[comp] = a;
    }
    const float inv = rsqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
#pragma unroll
    for (int comp = 0; comp < 3; ++comp) n.xout[(size_t)comp * cs + p] = in ? v[comp] * inv : 0.f;
  }
};

// Block b of a batched field launch: 256 consecutive cells of the member
// b2i[b] (every member starts on a multiple of 256 cells of the packed
// arrays X, HD and OUT, whose component stride is cs).  NK != 0: also the
// next stage state of an active member (BatchNext; NK > 0 its nk, -1 any).
template <int NK>
__device__ __forceinline__ void field_uniform_b_block(const FBInst* __restrict__ tab, const int* __restrict__ b2i,
                                                      const float* __restrict__ X, const float* __restrict__ HD,
                                                      float* __restrict__ OUT, int cs, int b, const BatchNext& nx) {
  const FBInst& I = tab[b2i[b]];
  const int j = (b - I.first) * 256 + threadIdx.x;
  const int nxc = I.d.nx, ny = I.d.ny;
  if (j >= I.d.ncell) return;
  const int x = j % nxc, yz = j / nxc, y = yz % ny, z = yz / ny;
  const float* hd = HD + I.off;
  auto hdem = [&](int i) { return load3(hd, i, cs); };
  if constexpr (NK != 0) {
    const BatchMember& S = nx.st[I.slot];
    if (S.active) {
      const int nk = NK > 0 ? NK : nx.nk;
      float c[7];
#pragma unroll
      for (int q = 0; q < 7; ++q) c[q] = q <= nk ? (float)(S.dt * nx.a[q]) : 0.f;
      field_uniform_cell(I.d, I.u, X + I.off, hdem, OUT + I.off, I.mode, x, y, z, cs,
                         BatchNextCell<NK>{nx, c, I.off, cs});
      return;
    }
  }
  field_uniform_cell(I.d, I.u, X + I.off, hdem, OUT + I.off, I.mode, x, y, z, cs);
}


// Epilogue of the last stage's batched field launch: the cell's error norm
// and torque norm (BatchErr) into err and tq.  c[j] = dt e_j.
struct BatchErrCell {
  const BatchErr& e;
  const float* c;
  int off, cs;
  bool want_err, want_tq;
  float& err;
  float& tq;
  __device__ __forceinline__ void operator()(int idx, float3 mc, float3 r, bool in) const {
    if (!in) return;
    const size_t p = (size_t)off + idx;
    if (want_err) {
      const float rr[3] = {r.x, r.y, r.z};
      float ev[3];
#pragma unroll
      for (int comp = 0; comp < 3; ++comp) {
        const size_t i = (size_t)comp * cs + p;
        float a = 0.f;
        for (int j = 0; j < e.nk; ++j) a += c[j] * e.k[j][i];
        a += c[e.nk] * rr[comp];
        ev[comp] = a;
      }
      err = fmaxf(0.f, sqrtf(ev[0] * ev[0] + ev[1] * ev[1] + ev[2] * ev[2]));
    }
    if (want_tq) tq = fmaxf(0.f, sqrtf(r.x * r.x + r.y * r.y + r.z * r.z));
  }
};

// Block maximum of v over the 256 threads (every thread calls it; red: 4
// floats of LDS; the result in thread 0).
__device__ __forceinline__ float field_block_max(float v, float* red) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  float r = 0.f;
  if (threadIdx.x == 0)
    for (int i = 0; i < 4; ++i) r = fmaxf(r, red[i]);
  return r;
}

// Block b of the last stage's batched field launch, with the error norms of
// the member's cells merged into errbits / tqbits (every thread of the block
// takes part in the reductions).
__device__ __forceinline__ void field_uniform_b_block_err(const FBInst* __restrict__ tab, const int* __restrict__ b2i,
                                                          const float* __restrict__ X, const float* __restrict__ HD,
                                                          float* __restrict__ OUT, int cs, int b, const BatchErr& be) {
  __shared__ float red[2][4];
  const FBInst& I = tab[b2i[b]];
  const int j = (b - I.first) * 256 + threadIdx.x;
  const bool valid = j < I.d.ncell;
  const int nxc = I.d.nx, ny = I.d.ny;
  const int x = j % nxc, yz = j / nxc, y = yz % ny, z = yz / ny;
  const float* hd = HD + I.off;
  auto hdem = [&](int i) { return load3(hd, i, cs); };
  const BatchMember& S = be.st[I.slot];
  // uniform over the block: one member
  const bool want_err = S.active && S.adaptive && be.nk > 0;
  const bool want_tq = S.active && (S.mode == BATCH_RELAX || S.accept == 2);
  if (!want_err && !want_tq) {
    if (valid) field_uniform_cell(I.d, I.u, X + I.off, hdem, OUT + I.off, I.mode, x, y, z, cs);
    return;
  }
  float c[7];
#pragma unroll
  for (int q = 0; q < 7; ++q) c[q] = q <= be.nk ? (float)(S.dt * be.e[q]) : 0.f;
  float err = 0.f, tq = 0.f;
  if (valid)
    field_uniform_cell(I.d, I.u, X + I.off, hdem, OUT + I.off, I.mode, x, y, z, cs,
                       BatchErrCell{be, c, I.off, cs, want_err, want_tq, err, tq});
  const float bm = field_block_max(err, red[0]);
  const float bt = field_block_max(tq, red[1]);
  if (threadIdx.x == 0) {
    if (want_err) atomicMax(be.errbits + I.slot, __float_as_uint(bm));
    if (want_tq) atomicMax(be.tqbits + I.slot, __float_as_uint(bt));
  }
}

// Every parameter uniform and only the terms the uniform path has.
bool all_uniform(const FMDesc& d);
// Host-side coefficients of the uniform path.
FMUniform make_uniform(const FMDesc& d);

}  // namespace mpa
