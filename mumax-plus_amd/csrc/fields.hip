// Fused local effective field + torque stencil.
//
// One thread per cell computes, in this order (the summation order of
// evalEffectiveField, src/physics/effectivefield.cpp:15-28):
//   exchange  (k_exchangeField FM variant, src/physics/exchange.cu:20-104)
//   anisotropy (uniaxial, else cubic: src/physics/anisotropy.cu:25-140)
//   external  (world bias + magnet bias, src/physics/zeeman.cu:53-91)
//   DMI       (k_dmiFieldFM, src/physics/dmi.cu:87-178 with getGamma, dmi.hpp:53-69)
//   demag     (added from the FFT passes)
// and then the LLG torque (k_llgtorque, src/physics/torque.cu:21-45), the
// damping-only torque (:58-77) or an energy density (src/physics/energy.cu:18-34,
// anisotropy.cu:142-230).  Exchange and DMI share one pass over the six
// neighbours; boundary handling (Neumann with the DMI boundary term, open
// boundaries, msat = 0 neighbours, mastergrid wrap) follows the reference
// branch by branch.
#include "fields.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <mutex>
#include <sstream>
#include <string>

#include "field_cell.hpp"

#include "philox.hpp"

namespace mpa {

namespace {

constexpr float GAMMALL = 1.7595e11f;  // src/physics/constants.hpp:7
constexpr float QE = 1.60217646e-19f;     // constants.hpp:13
constexpr float MUB = 9.274009152e-24f;   // constants.hpp:15
constexpr float HBAR = 1.054571817e-34f;  // constants.hpp:17
constexpr int BLOCK = 256;
// Widest x extent of a uniform-path tile (the rest of the 256 threads go
// along y), chosen by timing on MI355X.
#ifndef MPA_FIELD_TILE_X
#define MPA_FIELD_TILE_X 256
#endif
#ifndef MPA_FIELD_ZPER
#define MPA_FIELD_ZPER 1
#endif
#ifndef MPA_FIELD_NT_STORE
#define MPA_FIELD_NT_STORE 1
#endif

__device__ __forceinline__ float harmonic_mean(float a, float b) {
  if (a + b == 0.0f) return 0.0f;
  if (a == b) return a;
  return 2 * a * b / (a + b);
}

__device__ __forceinline__ float3 normalized3(float3 v) {
  const float n = sqrtf(dot3(v, v));
  return v / n;
}

// UNI: every parameter is uniform, so each lookup is the scalar and the
// harmonic mean of a cell and its neighbour is the value itself.  NK >= 0:
// the cell also writes the next Runge-Kutta stage state (nx, NK derivatives
// from memory and the one just computed; x0 == nullptr: the cell's state).
template <bool UNI, int NK = -2>
__device__ __forceinline__ void field_body(const FMDesc& d, const float* __restrict__ m,
                                           const float* __restrict__ hdemag, float* __restrict__ out, int mode,
                                           int bid, const NextStage* nx = nullptr) {
  const int idx = bid * BLOCK + threadIdx.x;
  const int N = d.ncell;
  if (idx >= N) return;
  auto S = [&](const PS& p, int i) { return UNI ? p.v : p.at(i); };
  auto V = [&](const PV& p, int i) { return UNI ? make_float3(p.x, p.y, p.z) : p.at(i, N); };
  auto HM = [&](float u, float v) { return UNI ? u : harmonic_mean(u, v); };
  // Exchange constant of the link to an in-geometry neighbour nb: hm, or
  // the inter-region value when the regions differ and it is non-zero,
  // times the inter-region scale (exchange.cu:79-99, afmexchange.cu:116-143).
  auto link = [&](float hm, int nb, const IP& inter, const IP& scale) {
    if (!d.regions) return hm;
    const unsigned r = d.regions[idx], r_ = d.regions[nb];
    if (r == r_) return hm;
    const float in = inter.between(r, r_);
    return (in != 0.f ? in : hm) * scale.between(r, r_);
  };
  const bool ingeom = !d.geom || d.geom[idx];
  if (!ingeom) {
    if (mode == MODE_EDENS) out[idx] = 0.f;
    else { out[idx] = 0.f; out[N + idx] = 0.f; out[2 * N + idx] = 0.f; }
    if constexpr (NK >= 0) nx->template apply<NK>(idx, f3(0.f, 0.f, 0.f), f3(0.f, 0.f, 0.f), false);
    return;
  }
  const int x = idx % d.nx;
  const int y = (idx / d.nx) % d.ny;
  const int z = idx / (d.nx * d.ny);
  const float3 mc = load3(m, idx, N);
  const float ms = S(d.msat, idx);
  // one reciprocal each for the divisions by msat and the cell sizes (the
  // field is VALU-bound and an IEEE division is about ten instructions)
  const float ims = 1.f / ms;
  float edens = 0.f;

  float3 hex = f3(0.f, 0.f, 0.f), hdm = f3(0.f, 0.f, 0.f);
  const bool do_ex = (d.terms & TERM_EXCH) && ms != 0.f;
  const bool do_dmi = (d.terms & TERM_DMI) && ms != 0.f;
  const bool sub = d.sublattice != 0;
  const float3 m2c = sub ? load3(d.m2, idx, N) : f3(0.f, 0.f, 0.f);
  const float csz[3] = {d.cx, d.cy, d.cz};
  const float icsz[3] = {1.f / d.cx, 1.f / d.cy, 1.f / d.cz};
  // Neighbour table, and every neighbour value loaded before any is used: the
  // index after the wrap (the cell itself when missing), whether it is an
  // in-geometry neighbour, and m (and the sister's m) there.  The opposite
  // neighbour of direction dir is entry dir ^ 1.
  int nbtab[6];
  bool nin[6];
  float3 mn_[6], m2n_[6];
  const bool stencil = do_ex || do_dmi || (sub && (d.terms & TERM_AFM_NN));
  if (stencil) {
#pragma unroll
    for (int dir = 0; dir < 6; ++dir) {
      const int nb = neighbour(d, x, y, z, dir);
      nin[dir] = nb >= 0 && (!d.geom || d.geom[nb]);
      nbtab[dir] = nin[dir] ? nb : idx;
    }
#pragma unroll
    for (int dir = 0; dir < 6; ++dir) {
      mn_[dir] = load3(m, nbtab[dir], N);
      m2n_[dir] = sub ? load3(d.m2, nbtab[dir], N) : f3(0.f, 0.f, 0.f);
    }
  }
  if (do_ex || do_dmi) {
    const float a = S(d.aex, idx);
    const float an = sub ? S(d.afmex_nn, idx) : 0.f;
    float D[D_COUNT];
#pragma unroll
    for (int k = 0; k < D_COUNT; ++k) D[k] = S(d.dmi[k], idx);
    const float w[3] = {icsz[0] * icsz[0], icsz[1] * icsz[1], icsz[2] * icsz[2]};
    const float h2a = 0.5f / a;  // delta / (2 a) = delta h2a at a Neumann boundary
    const bool openbc = d.openbc != 0;
    const bool open_dmi = openbc || a == 0.f;
#pragma unroll
    for (int dir = 0; dir < 6; ++dir) {
      const int axis = dir >> 1;
      const float delta = (dir & 1) ? csz[axis] : -csz[axis];
      const float idelta = (dir & 1) ? icsz[axis] : -icsz[axis];
      const int nb = nbtab[dir];
      const bool nb_in = nin[dir];
      float3 gamma = f3(0.f, 0.f, 0.f);
      // sister's normal derivative taken on the bulk side of a boundary
      // (exchange.cu:176-183, dmi.cu:200-207)
      float3 dm2 = f3(0.f, 0.f, 0.f);
      if (!nb_in) {
        gamma = dmi_gamma(D, axis, mc);
        if (sub && nin[dir ^ 1]) dm2 = idelta * (m2c - m2n_[dir ^ 1]);
      }
      if (do_ex) {
        bool use = true;
        float3 mn;
        float an_;
        bool linked = false;  // an in-geometry neighbour: the region rules apply
        if (sub && !openbc) {
          // sublattice variant, exchange.cu:106-202 (no open boundaries)
          if (!nb_in) {
            if (a == 0.f) use = false;
            mn = mc + (delta * h2a) * (an * cross3(cross3(dm2, mc), mc) + gamma);
            an_ = a;
          } else {
            if (S(d.msat, nb) == 0.f) use = false;
            mn = mn_[dir];
            an_ = S(d.aex, nb);
            linked = true;
          }
        } else if (!nb_in) {
          // exchange.cu:61-101
          if (openbc || a == 0.f) use = false;
          mn = mc + (delta * h2a) * gamma;
          an_ = a;
        } else {
          if (openbc && S(d.msat, nb) == 0.f) use = false;
          mn = mn_[dir];
          an_ = S(d.aex, nb);
          linked = true;
        }
        if (use) {
          float A = HM(a, an_);
          if (linked) A = link(A, nb, d.inter_ex, d.scale_ex);
          hex = hex + (2 * A * w[axis]) * (mn - mc);
        }
      }
      if (do_dmi) {
        const int nbi = nb;  // the cell itself when the neighbour is missing
        const int kxz = axis == 0 ? D_XXZ : (axis == 1 ? D_YXZ : D_ZXZ);
        const int kxy = axis == 0 ? D_XXY : (axis == 1 ? D_YXY : D_ZXY);
        const int kyz = axis == 0 ? D_XYZ : (axis == 1 ? D_YYZ : D_ZYZ);
        float Dxz = D[kxz], Dxy = D[kxy], Dyz = D[kyz];
        if (nbi != idx) {
          Dxz = HM(Dxz, S(d.dmi[kxz], nbi));
          Dxy = HM(Dxy, S(d.dmi[kxy], nbi));
          Dyz = HM(Dyz, S(d.dmi[kyz], nbi));
        }
        if (!sub) {
          // dmi.cu:114-173
          const bool skip = open_dmi && (!nb_in || S(d.msat, nbi) == 0.f);
          if (!skip) {
            float3 mn;
            if (!nb_in && !open_dmi) mn = mc + (delta * h2a) * gamma;
            else mn = mn_[dir];
            hdm.x += (Dxy * mn.y + Dxz * mn.z) * idelta;
            hdm.y += (-Dxy * mn.x + Dyz * mn.z) * idelta;
            hdm.z += (-Dxz * mn.x - Dyz * mn.y) * idelta;
          }
        } else {
          // sublattice with the inter-sublattice tensor, dmi.cu:112-240
          const bool skip = open_dmi && (!nb_in || (S(d.msat, nbi) == 0.f && S(d.msat2, nbi) == 0.f));
          if (!skip) {
            const float Dixz = HM(S(d.hdmi[kxz], idx), S(d.hdmi[kxz], nbi));
            const float Dixy = HM(S(d.hdmi[kxy], idx), S(d.hdmi[kxy], nbi));
            const float Diyz = HM(S(d.hdmi[kyz], idx), S(d.hdmi[kyz], nbi));
            float3 m1n, m2n;
            if (!nb_in) {
              // Neumann boundary; both sublattices use this sublattice's own tensor
              const float3 g2 = dmi_gamma(D, axis, m2c);
              m1n = mc + (delta * h2a) * (an * cross3(cross3(dm2, mc), mc) + gamma);
              m2n = m2c + (delta * h2a) * (an * cross3(cross3(idelta * (m1n - mc), m2c), m2c) + g2);
            } else {
              m1n = mn_[dir];
              m2n = m2n_[dir];
            }
            const float s1 = S(d.msat, nbi) != 0.f ? 1.f : 0.f;
            const float s2 = S(d.msat2, nbi) != 0.f ? 1.f : 0.f;
            hdm.x += (s1 * (Dxy * m1n.y + Dxz * m1n.z) - s2 * (Dixy * m2n.y - Dixz * m2n.z)) * idelta;
            hdm.y += (s1 * (-Dxy * m1n.x + Dyz * m1n.z) - s2 * (-Dixy * m2n.x - Diyz * m2n.z)) * idelta;
            hdm.z += (s1 * (-Dxz * m1n.x - Dyz * m1n.y) - s2 * (-Dixz * m2n.x + Diyz * m2n.y)) * idelta;
          }
        }
      }
    }
    if (do_ex) hex = ims * hex;
    if (do_dmi) hdm = ims * hdm;
  }

  // magnetoelastic field, magnetoelasticfield.cu:44-103 (no msat check, as there)
  float3 hme = f3(0.f, 0.f, 0.f);
  if (d.terms & TERM_ME) {
    float e[6];
    if (d.me_rigid) {
      const float3 n = V(d.rnorm, idx), s = V(d.rshear, idx);
      e[0] = n.x; e[1] = n.y; e[2] = n.z; e[3] = s.x; e[4] = s.y; e[5] = s.z;
    } else {
#pragma unroll
      for (int c = 0; c < 6; ++c) e[c] = d.strain[(size_t)c * N + idx];
    }
    const float b1 = S(d.B1, idx), b2 = S(d.B2, idx), f = -2.f * ims;
    hme = f3(f * (b1 * e[0] * mc.x + b2 * (e[3] * mc.y + e[4] * mc.z)),
             f * (b1 * e[1] * mc.y + b2 * (e[5] * mc.z + e[3] * mc.x)),
             f * (b1 * e[2] * mc.z + b2 * (e[4] * mc.x + e[5] * mc.y)));
  }

  // antiferromagnetic exchange with the sister sublattice, divided by the
  // sister's msat (afmexchange.cu:28-149)
  float3 hafm_nn = f3(0.f, 0.f, 0.f), hafm_site = f3(0.f, 0.f, 0.f);
  if (sub && (d.terms & (TERM_AFM_NN | TERM_AFM_SITE))) {
    const float ms2 = S(d.msat2, idx);
    if (ms2 != 0.f && (d.terms & TERM_AFM_SITE)) {
      const float l = S(d.latcon, idx);
      hafm_site = (4.f * S(d.afmex_cell, idx) / (l * l * ms2)) * m2c;
    }
    if (ms2 != 0.f && (d.terms & TERM_AFM_NN)) {
      const float a = S(d.aex, idx);
      const float h2a = 0.5f / a;
      const float ann = S(d.afmex_nn, idx);
      const bool open_nn = d.openbc != 0 || a == 0.f;
      float D[D_COUNT];
#pragma unroll
      for (int k = 0; k < D_COUNT; ++k) D[k] = S(d.dmi[k], idx);
#pragma unroll
      for (int dir = 0; dir < 6; ++dir) {
        const int axis = dir >> 1;
        const float delta = (dir & 1) ? csz[axis] : -csz[axis];
        const float idelta = (dir & 1) ? icsz[axis] : -icsz[axis];
        const int nb = nbtab[dir];
        const bool nb_in = nin[dir];
        if (!nb_in && open_nn) continue;
        if (open_nn && S(d.msat2, nb) == 0.f) continue;  // nb_in here
        float3 m2n;
        float ann_;
        if (nb_in) {
          m2n = m2n_[dir];
          ann_ = S(d.afmex_nn, nb);
        } else {
          const float3 g2 = dmi_gamma(D, axis, m2c);
          float3 dm1 = f3(0.f, 0.f, 0.f);
          if (nin[dir ^ 1]) dm1 = idelta * (mc - mn_[dir ^ 1]);
          m2n = m2c + (delta * h2a) * (ann * cross3(cross3(dm1, m2c), m2c) + g2);
          ann_ = ann;
        }
        float A = HM(ann, ann_);
        if (nb_in) A = link(A, nb, d.inter_afm, d.scale_afm);
        hafm_nn = hafm_nn + (A * (idelta * idelta)) * (m2n - m2c);
      }
      hafm_nn = (1.f / ms2) * hafm_nn;
    }
  }

  float3 han = f3(0.f, 0.f, 0.f);
  if (ms != 0.f) {
    if (d.terms & TERM_UNI) {
      const float3 u = normalized3(V(d.anisU, idx));
      const float k1 = S(d.ku1, idx), k2 = S(d.ku2, idx);
      const float mu = dot3(mc, u);
      han = ((2 * k1 * mu + 4 * k2 * mu * mu * mu) * ims) * u;
      if (mode == MODE_EDENS) edens += -k1 * mu * mu - k2 * mu * mu * mu * mu;
    } else if (d.terms & TERM_CUBIC) {
      const float3 c1 = normalized3(V(d.anisC1, idx));
      const float3 c2 = normalized3(V(d.anisC2, idx));
      const float3 c3 = cross3(c1, c2);
      const float kc1 = S(d.kc1, idx), kc2 = S(d.kc2, idx), kc3 = S(d.kc3, idx);
      const float a1 = dot3(c1, mc), a2 = dot3(c2, mc), a3 = dot3(c3, mc);
      const float s1 = a1 * a1, s2 = a2 * a2, s3 = a3 * a3;
      han = (-2 * kc1 * ims) * ((s2 + s3) * a1 * c1 + (s1 + s3) * a2 * c2 + (s1 + s2) * a3 * c3) +
            (-2 * kc2 * ims) * ((s2 * s3) * a1 * c1 + (s1 * s3) * a2 * c2 + (s1 * s2) * a3 * c3) +
            (-4 * kc3 * ims) * ((s2 * s2 + s3 * s3) * (s1 * a1) * c1 + (s1 * s1 + s3 * s3) * (s2 * a2) * c2 +
                               (s1 * s1 + s2 * s2) * (s3 * a3) * c3);
      if (mode == MODE_EDENS)
        edens += kc1 * (s1 * s2 + s1 * s3 + s2 * s3) + kc2 * s1 * s2 * s3 +
                 kc3 * (s1 * s1 * s2 * s2 + s1 * s1 * s3 * s3 + s2 * s2 * s3 * s3);
    }
  }
  float3 hext = f3(0.f, 0.f, 0.f);
  if (d.terms & TERM_EXT) {
    hext = f3(d.wb[0], d.wb[1], d.wb[2]) + V(d.bias, idx);
    if (d.hstray) hext = hext + load3(d.hstray, idx, N);
  }
  float3 hdem = f3(0.f, 0.f, 0.f);
  if (d.terms & TERM_DEMAG) hdem = load3(hdemag, idx, N);

  if (mode == MODE_EDENS) {
    edens += -0.5f * ms * dot3(mc, hex);
    edens += -1.0f * ms * dot3(mc, hext);
    edens += -0.5f * ms * dot3(mc, hdm);
    edens += -0.5f * ms * dot3(mc, hdem);
    edens += -0.5f * ms * dot3(mc, hme + hafm_nn + hafm_site);
    out[idx] = edens;
    return;
  }
  // the terms of effectivefield.cpp:15-28, the demag field added last (so a
  // field evaluated without demag plus the demag field is this sum bit for
  // bit: k_torque_pair)
  float3 h = hex;
  h = h + han;
  h = h + hext;
  h = h + hdm;
  h = h + hme;
  h = h + hafm_nn;
  h = h + hafm_site;
  h = h + hdem;

  // spin-transfer torque (evalSpinTransferTorque, stt.cu:152-195): Slonczewski
  // when it is not assured zero, else Zhang-Li; zero for frozen spins
  float3 tstt = f3(0.f, 0.f, 0.f);
  if ((d.terms & TERM_STT) && S(d.frozen, idx) == 0.f && (mode == MODE_LLG || mode == MODE_STT)) {
    const float3 j = V(d.jcur, idx);
    const float al = S(d.alpha, idx), pol = S(d.pol, idx);
    if (d.terms & TERM_SLONC) {
      // k_Slonczewski, stt.cu:104-150
      const float jz = j.z;
      const float3 p = V(d.fixed, idx);
      const float lam = S(d.lambda, idx), epsp = S(d.epsp, idx);
      float th = S(d.fthick, idx);
      if (!d.fixed_on_top) th = -th;
      const bool zero = ms == 0.f || jz == 0.f || th == 0.f || (p.x == 0.f && p.y == 0.f && p.z == 0.f) ||
                        (epsp == 0.f && (lam == 0.f || pol == 0.f));
      if (!zero) {
        const float B = (HBAR / QE) * jz / (ms * th);
        const float l2 = lam * lam;
        const float eps = pol * l2 / ((l2 + 1) + (l2 - 1) * dot3(mc, p));
        const float3 pxm = cross3(p, mc);
        const float3 mxpxm = cross3(mc, pxm);
        tstt = ((B / (1 + al * al)) * GAMMALL) * ((eps + epsp * al) * mxpxm + (epsp - eps * al) * pxm);
      }
    } else {
      // k_ZhangLi, stt.cu:27-102: central differences over in-grid neighbours
      // with non-zero msat (after the mastergrid wrap)
      const float xi = S(d.xi, idx);
      if (!(ms == 0.f || pol == 0.f || (j.x == 0.f && j.y == 0.f && j.z == 0.f))) {
        const float3 u = (MUB * pol / (QE * ms * (1 + xi * xi))) * j;
        const float uc[3] = {u.x, u.y, u.z};
        float3 hs = f3(0.f, 0.f, 0.f);
#pragma unroll
        for (int dir = 0; dir < 6; ++dir) {
          const int axis = dir >> 1;
          const float sign = (dir & 1) ? 1.f : -1.f;
          const int nb = neighbour(d, x, y, z, dir);
          if (nb >= 0 && S(d.msat, nb) != 0.f) hs = hs + (sign * uc[axis] / (2 * csz[axis])) * load3(m, nb, N);
        }
        const float3 mxh = cross3(mc, hs);
        const float3 mxmxh = cross3(mc, mxh);
        tstt = (-1 / (1 + al * al)) * ((1 + xi * al) * mxmxh + (xi - al) * mxh);
      }
    }
  }
  // thermal field of this cell (k_thermalNoise, thermalnoise.cu:21-44)
  float3 tn = f3(0.f, 0.f, 0.f);
  if ((d.terms & TERM_THERMAL) && (d.noise || mode == MODE_NOISE) && ms != 0.f) {
    float zn[3];
    normal3(d.noise_seed, (unsigned)idx, d.noise_purpose, d.noise_draw, zn);
    const float a = S(d.alpha, idx), T = S(d.temperature, idx);
    const float amp = sqrtf(d.noise_pref * a * T / ((1 + a * a) * ms));
    tn = f3(amp * zn[0], amp * zn[1], amp * zn[2]);
  }

  float3 r;
  if (mode == MODE_FIELD) {
    r = h;
  } else if (mode == MODE_STT) {
    r = tstt;
  } else if (mode == MODE_NOISE) {
    r = tn;
  } else if (S(d.frozen, idx) != 0.f) {
    r = f3(0.f, 0.f, 0.f);
  } else if (mode == MODE_LLG) {
    const float al = S(d.alpha, idx);
    const float3 mxh = cross3(mc, h);
    const float3 mxmxh = cross3(mc, mxh);
    r = (-GAMMALL / (1 + al * al)) * (mxh + al * mxmxh);
    r = r + tstt;  // evalTorque, torque.cu:14-19
  } else {
    r = -GAMMALL * cross3(mc, cross3(mc, h));
  }
  // the stepper's noise term, added to every stage derivative (also of frozen spins)
  if (d.noise && (mode == MODE_LLG || mode == MODE_DAMPING)) r = r + d.noise_scale * tn;
  out[idx] = r.x;
  out[N + idx] = r.y;
  out[2 * N + idx] = r.z;
  if constexpr (NK >= 0) nx->template apply<NK>(idx, mc, r, true);
}

template <bool UNI>
__global__ void __launch_bounds__(BLOCK) k_field(FMDesc d, const float* __restrict__ m,
                                                 const float* __restrict__ hdemag, float* __restrict__ out,
                                                 int mode) {
  field_body<UNI>(d, m, hdemag, out, mode, xcd_remap(blockIdx.x, gridDim.x));
}

// Both sublattices of an antiferromagnet in one launch: blockIdx.y selects
// one.  Each branch reads its own kernel argument at fixed offsets (indexing
// a descriptor array by blockIdx.y would copy it to scratch).
template <bool UNI>
__global__ void __launch_bounds__(BLOCK) k_field_pair(FMDesc d0, const float* __restrict__ m0,
                                                      const float* __restrict__ h0, float* __restrict__ o0, FMDesc d1,
                                                      const float* __restrict__ m1, const float* __restrict__ h1,
                                                      float* __restrict__ o1, int mode) {
  const int bid = xcd_remap(blockIdx.x, gridDim.x);
  if (blockIdx.y == 0) field_body<UNI>(d0, m0, h0, o0, mode, bid);
  else field_body<UNI>(d1, m1, h1, o1, mode, bid);
}

// k_field_pair that also writes both sublattices' next stage states (n0, n1
// with NK derivatives from memory each).
template <bool UNI, int NK>
__global__ void __launch_bounds__(BLOCK)
    k_field_pair_next(FMDesc d0, const float* __restrict__ m0, const float* __restrict__ h0, float* __restrict__ o0,
                      NextStage n0, FMDesc d1, const float* __restrict__ m1, const float* __restrict__ h1,
                      float* __restrict__ o1, NextStage n1, int mode) {
  const int bid = xcd_remap(blockIdx.x, gridDim.x);
  if (blockIdx.y == 0) field_body<UNI, NK>(d0, m0, h0, o0, mode, bid, &n0);
  else field_body<UNI, NK>(d1, m1, h1, o1, mode, bid, &n1);
}

// The torque of a sublattice cell from its field without demag (hl) and the
// demag field (hd), as field_body forms it for uniform parameters in MODE_LLG
// or MODE_DAMPING (no spin-transfer torque, no thermal noise, not frozen).
__device__ __forceinline__ void torque_cell(const FMDesc& d, const float* __restrict__ m,
                                            const float* __restrict__ hl, const float* __restrict__ hd,
                                            float* __restrict__ out, int mode, int idx) {
  const int N = d.ncell;
  if (idx >= N) return;
  if (d.geom && !d.geom[idx]) {
    out[idx] = 0.f;
    out[N + idx] = 0.f;
    out[2 * N + idx] = 0.f;
    return;
  }
  const float3 mc = load3(m, idx, N);
  const float3 h = load3(hl, idx, N) + load3(hd, idx, N);
  float3 r;
  if (mode == MODE_LLG) {
    const float al = d.alpha.v;
    const float3 mxh = cross3(mc, h);
    const float3 mxmxh = cross3(mc, mxh);
    r = (-GAMMALL / (1 + al * al)) * (mxh + al * mxmxh);
  } else {
    r = -GAMMALL * cross3(mc, cross3(mc, h));
  }
  out[idx] = r.x;
  out[N + idx] = r.y;
  out[2 * N + idx] = r.z;
}

__global__ void __launch_bounds__(BLOCK) k_torque_pair(FMDesc d0, const float* __restrict__ m0,
                                                       const float* __restrict__ hl0, const float* __restrict__ hd0,
                                                       float* __restrict__ o0, FMDesc d1, const float* __restrict__ m1,
                                                       const float* __restrict__ hl1, const float* __restrict__ hd1,
                                                       float* __restrict__ o1, int mode) {
  const int idx = xcd_remap(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
  if (blockIdx.y == 0) torque_cell(d0, m0, hl0, hd0, o0, mode, idx);
  else torque_cell(d1, m1, hl1, hd1, o1, mode, idx);
}

template <bool NEXT>
__global__ void __launch_bounds__(BLOCK) k_field_uniform(FMDesc d, FMUniform u, const float* __restrict__ m,
                                                         const float* __restrict__ hdemag, float* __restrict__ out,
                                                         int mode, NextStage next) {
  // 2-D blocks: threadIdx.x along x, threadIdx.y along y, one z plane per
  // block row.  Workgroups are dealt round-robin to the 8 XCDs in dispatch
  // order; the remap gives each XCD a contiguous run of tiles, so the y and z
  // neighbour rows a tile reads are mostly in that XCD's L2 already.
  const int nbx = gridDim.x, nby = gridDim.y;
  const int lin = blockIdx.x + nbx * (blockIdx.y + nby * blockIdx.z);
  const int tile = xcd_remap(lin, nbx * nby * gridDim.z);
  const int bxi = tile % nbx, rest = tile / nbx;
  const int x = bxi * blockDim.x + threadIdx.x;
  const int y = (rest % nby) * blockDim.y + threadIdx.y;
  if (x >= d.nx || y >= d.ny) return;
  const int N = d.ncell;
  // MPA_FIELD_ZPER consecutive z planes per thread
#pragma unroll
  for (int k = 0; k < MPA_FIELD_ZPER; ++k) {
    const int z = (rest / nby) * MPA_FIELD_ZPER + k;
    if (z < d.nz) {
      auto hd = [&](int i) { return load3(hdemag, i, N); };
      if constexpr (NEXT) field_uniform_cell(d, u, m, hd, out, mode, x, y, z, N, next);
      else field_uniform_cell(d, u, m, hd, out, mode, x, y, z, N);
    }
  }
}

// LDS halo-tile variant of the uniform path, for a magnet with no geometry
// mask and no periodic wrap.  A block owns a TILE_X x TILE_Y column of cells
// and marches it through `zc` consecutive z planes.  Per plane, the block's
// centre values and a one-cell halo ring go into LDS (double-buffered, one
// barrier per plane), the x and y neighbours are LDS reads at immediate
// offsets, and the z neighbours are the thread's own centre values of the
// previous and next planes, kept in registers.  Each plane's centre is thus
// read from global memory once (plus the two planes beyond the chunk), and
// the next plane's loads are in flight while the current plane is computed.
// The arithmetic is uniform_result's, so the field equals k_field_uniform's
// bit for bit.
constexpr int TILE_X = 64, TILE_Y = 4;
constexpr int TILE_PX = TILE_X + 2;  // LDS row pitch (halo column on each side)

template <bool NEXT, bool FAST, int NK = -1>
__global__ void __launch_bounds__(TILE_X* TILE_Y) k_field_tile(FMDesc d, FMUniform u, const float* __restrict__ m,
                                                                 const float* __restrict__ hdemag,
                                                                 float* __restrict__ out, int mode, NextStage next,
                                                                 int zc) {
  __shared__ float s[2][3][(TILE_Y + 2) * TILE_PX];
  const int nbx = gridDim.x, nby = gridDim.y;
  const int tile = xcd_remap(blockIdx.x + nbx * (blockIdx.y + nby * blockIdx.z), nbx * nby * gridDim.z);
  const int bxi = tile % nbx, rest = tile / nbx;
  const int byi = rest % nby, bzi = rest / nby;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int nx = d.nx, ny = d.ny, nz = d.nz, N = d.ncell, P = nx * ny;
  const int x0 = bxi * TILE_X, y0 = byi * TILE_Y;
  const int x = x0 + tx, y = y0 + ty;
  const int z0 = bzi * zc, z1 = min(z0 + zc, nz);
  const bool inb = x < nx && y < ny;
  const int col = y * nx + x;  // index within a plane
  // Halo cell of this thread: rows y0-1 (ty 0) and y0+TILE_Y (ty 1) over the
  // tile's x range, columns x0-1 and x0+TILE_X (ty 2, tx < 2 TILE_Y) over its
  // y range; ty 3 has none.  hl: LDS slot, hc: plane index, hv: inside grid.
  int hl = -1, hc = 0;
  bool hv = false;
  if (ty < 2) {
    const int hy = ty == 0 ? y0 - 1 : y0 + TILE_Y;
    hl = (ty == 0 ? 0 : TILE_Y + 1) * TILE_PX + tx + 1;
    hv = x < nx && hy >= 0 && hy < ny;
    hc = hy * nx + x;
  } else if (ty == 2 && tx < 2 * TILE_Y) {
    const int r = tx % TILE_Y, hx = tx < TILE_Y ? x0 - 1 : x0 + TILE_X, hy = y0 + r;
    hl = (r + 1) * TILE_PX + (tx < TILE_Y ? 0 : TILE_X + 1);
    hv = hx >= 0 && hx < nx && hy < ny;
    hc = hy * nx + hx;
  }
  const int cl = (ty + 1) * TILE_PX + tx + 1;  // LDS slot of the thread's cell
  auto ld = [&](int plane, int c) -> float3 { return load3(m, plane * P + c, N); };
  auto put = [&](int b, int slot, float3 v) {
    s[b][0][slot] = v.x;
    s[b][1][slot] = v.y;
    s[b][2][slot] = v.z;
  };
  const float3 zero = f3(0.f, 0.f, 0.f);
  // planes z0 - 1 and z0 into registers, plane z0 into LDS buffer 0
  float3 mzm = (inb && z0 > 0) ? ld(z0 - 1, col) : zero;
  float3 mc = inb ? ld(z0, col) : zero;
  put(0, cl, mc);
  if (hl >= 0) put(0, hl, hv ? ld(z0, hc) : zero);
  __syncthreads();
  const int ndir = u.zflat ? 4 : 6;
  const bool xm_ok = x > 0, xp_ok = x + 1 < nx, ym_ok = y > 0, yp_ok = y + 1 < ny;
  for (int z = z0; z < z1; ++z) {
    const int b = (z - z0) & 1;
    // next plane's centre (the +z neighbour) and halo cell
    const bool more = z + 1 < nz;
    const float3 mzp = (inb && more) ? ld(z + 1, col) : zero;
    const float3 hnext = (hl >= 0 && hv && z + 1 < z1) ? ld(z + 1, hc) : zero;
    if (inb) {
      const int idx = z * P + col;
      const float3 hs = (!FAST && d.hstray) ? load3(d.hstray, idx, N) : zero;
      const float3 hd = (FAST || u.demag) ? load3(hdemag, idx, N) : zero;
      float3 mn[6];
      bool ok[6];
      auto at = [&](int slot) { return f3(s[b][0][slot], s[b][1][slot], s[b][2][slot]); };
      mn[0] = at(cl - 1);
      mn[1] = at(cl + 1);
      mn[2] = at(cl - TILE_PX);
      mn[3] = at(cl + TILE_PX);
      mn[4] = mzm;
      mn[5] = mzp;
      ok[0] = xm_ok;
      ok[1] = xp_ok;
      ok[2] = ym_ok;
      ok[3] = yp_ok;
      ok[4] = z > 0;
      ok[5] = more;
      const float3 r = uniform_result<FAST>(u, mode, mc, mn, ok, ndir, d.hstray != nullptr, hs, hd);
#if MPA_FIELD_NT_STORE
      if constexpr (NEXT) {
        // the derivative is read again only by later stages: stream it past
        // the caches, which keep the stage state the next demag pass reads
        __builtin_nontemporal_store(r.x, out + idx);
        __builtin_nontemporal_store(r.y, out + N + idx);
        __builtin_nontemporal_store(r.z, out + 2 * N + idx);
      } else
#endif
      {
        out[idx] = r.x;
        out[N + idx] = r.y;
        out[2 * N + idx] = r.z;
      }
      if constexpr (NEXT) next.template apply<NK>(idx, mc, r, true);
    }
    if (z + 1 < z1) {
      put(b ^ 1, cl, mzp);
      if (hl >= 0) put(b ^ 1, hl, hnext);
      __syncthreads();
    }
    mzm = mc;
    mc = mzp;
  }
}

// Batched uniform path: 256 consecutive cells of one magnet per block (every
// magnet starts on a multiple of 256 cells of the packed arrays).
template <int NK>
__global__ void __launch_bounds__(BLOCK) k_field_uniform_b(const FBInst* __restrict__ tab, const int* __restrict__ b2i,
                                                           const float* __restrict__ X, const float* __restrict__ HD,
                                                           float* __restrict__ OUT, int cs, BatchNext nx) {
  field_uniform_b_block<NK>(tab, b2i, X, HD, OUT, cs, xcd_remap(blockIdx.x, gridDim.x), nx);
}

// Every parameter the per-cell kernel reads is uniform.
bool params_uniform(const FMDesc& d) {
  bool u = !d.msat.p && !d.aex.p && !d.alpha.p && !d.ku1.p && !d.ku2.p && !d.kc1.p && !d.kc2.p && !d.kc3.p &&
           !d.frozen.p && !d.anisU.p && !d.anisC1.p && !d.anisC2.p && !d.bias.p && !d.B1.p && !d.B2.p &&
           !d.rnorm.p && !d.rshear.p && !d.msat2.p && !d.afmex_cell.p && !d.afmex_nn.p && !d.latcon.p &&
           !d.pol.p && !d.xi.p && !d.lambda.p && !d.epsp.p && !d.fthick.p && !d.jcur.p && !d.fixed.p &&
           !d.temperature.p;
  for (int k = 0; k < D_COUNT; ++k) u = u && !d.dmi[k].p && !d.hdmi[k].p;
  return u;
}


}  // namespace

bool all_uniform(const FMDesc& d) {
  bool u = !d.msat.p && !d.aex.p && !d.alpha.p && !d.ku1.p && !d.ku2.p && !d.kc1.p && !d.kc2.p && !d.kc3.p &&
           !d.frozen.p && !d.anisU.p && !d.anisC1.p && !d.anisC2.p && !d.bias.p;
  for (int k = 0; k < D_COUNT; ++k) u = u && !d.dmi[k].p;
  // sublattices, inter-region exchange, magnetoelastic coupling,
  // spin-transfer torques and the thermal field use the per-cell kernel only
  return u && !d.sublattice && !d.regions &&
         !(d.terms & (TERM_ME | TERM_AFM_NN | TERM_AFM_SITE | TERM_STT | TERM_THERMAL));
}

// Host-side coefficients of the uniform path, in double then rounded once.
FMUniform make_uniform(const FMDesc& d) {
  FMUniform u{};
  const double ms = d.msat.v, A = d.aex.v;
  const double c[3] = {d.cx, d.cy, d.cz};
  const bool msok = ms != 0.0;
  u.has_ex = (d.terms & TERM_EXCH) && msok;
  u.has_dmi = (d.terms & TERM_DMI) && msok;
  u.openbc = d.openbc != 0;
  u.openbc_dmi = d.openbc != 0 || A == 0.0;
  const int kxz[3] = {D_XXZ, D_YXZ, D_ZXZ}, kxy[3] = {D_XXY, D_YXY, D_ZXY}, kyz[3] = {D_XYZ, D_YYZ, D_ZYZ};
  for (int k = 0; k < D_COUNT; ++k) u.D[k] = d.dmi[k].v;
  for (int a = 0; a < 3; ++a) {
    u.kex[a] = msok ? (float)(2.0 * A / (ms * c[a] * c[a])) : 0.f;
    u.inv_msc[a] = msok ? (float)(1.0 / (ms * c[a])) : 0.f;
    u.dxz[a] = msok ? (float)(d.dmi[kxz[a]].v / (ms * c[a])) : 0.f;
    u.dxy[a] = msok ? (float)(d.dmi[kxy[a]].v / (ms * c[a])) : 0.f;
    u.dyz[a] = msok ? (float)(d.dmi[kyz[a]].v / (ms * c[a])) : 0.f;
    u.ncoef[a] = A != 0.0 ? (float)(c[a] / (2.0 * A)) : 0.f;
  }
  u.uni = (d.terms & TERM_UNI) && msok;
  u.cubic = (d.terms & TERM_CUBIC) && msok;
  if (u.uni) {
    const double ux = d.anisU.x, uy = d.anisU.y, uz = d.anisU.z, n = std::sqrt(ux * ux + uy * uy + uz * uz);
    u.u[0] = (float)(ux / n); u.u[1] = (float)(uy / n); u.u[2] = (float)(uz / n);
    u.ku1 = (float)(2.0 * d.ku1.v / ms);
    u.ku2 = (float)(4.0 * d.ku2.v / ms);
  }
  if (u.cubic) {
    double a1[3] = {d.anisC1.x, d.anisC1.y, d.anisC1.z}, a2[3] = {d.anisC2.x, d.anisC2.y, d.anisC2.z};
    const double n1 = std::sqrt(a1[0] * a1[0] + a1[1] * a1[1] + a1[2] * a1[2]);
    const double n2 = std::sqrt(a2[0] * a2[0] + a2[1] * a2[1] + a2[2] * a2[2]);
    for (int i = 0; i < 3; ++i) { a1[i] /= n1; a2[i] /= n2; }
    // c3 = c1 x c2 of the FP32-normalised axes, as in the per-cell kernel
    float f1[3], f2[3];
    for (int i = 0; i < 3; ++i) { f1[i] = (float)a1[i]; f2[i] = (float)a2[i]; }
    const float3 c3 = cross3(f3(f1[0], f1[1], f1[2]), f3(f2[0], f2[1], f2[2]));
    for (int i = 0; i < 3; ++i) { u.c1[i] = f1[i]; u.c2[i] = f2[i]; }
    u.c3[0] = c3.x; u.c3[1] = c3.y; u.c3[2] = c3.z;
    u.kc1 = (float)(-2.0 * d.kc1.v / ms);
    u.kc2 = (float)(-2.0 * d.kc2.v / ms);
    u.kc3 = (float)(-4.0 * d.kc3.v / ms);
  }
  if (d.terms & TERM_EXT) {
    u.hext[0] = (float)((double)d.wb[0] + d.bias.x);
    u.hext[1] = (float)((double)d.wb[1] + d.bias.y);
    u.hext[2] = (float)((double)d.wb[2] + d.bias.z);
  }
  u.demag = (d.terms & TERM_DEMAG) != 0;
  u.zflat = d.nz == 1 && d.dmi[D_ZXY].v == 0.f && d.dmi[D_ZYZ].v == 0.f && d.dmi[D_ZXZ].v == 0.f;
  const double al = d.alpha.v;
  u.alpha = (float)al;
  u.llg = (float)(-1.7595e11 / (1.0 + al * al));
  u.frozen = d.frozen.v != 0.f;
  return u;
}


bool field_batchable(const FMDesc& d) { return all_uniform(d) && !d.hstray; }

FieldBatch::FieldBatch(const std::vector<FieldBatchMember>& members, int cs) : cs_(cs) {
  std::vector<FBInst> tab;
  std::vector<int> b2i;
  for (const FieldBatchMember& mb : members) {
    if (!field_batchable(mb.d)) throw Error(MPA_ERR_INVALID, "batched field: every parameter must be uniform");
    if (mb.off % BLOCK != 0) throw Error(MPA_ERR_RUNTIME, "batched field: magnets must start on 256-cell boundaries");
    if (mb.mode != MODE_FIELD && mb.mode != MODE_LLG && mb.mode != MODE_DAMPING)
      throw Error(MPA_ERR_INVALID, "batched field: fields and torques only");
    FBInst I{};
    I.d = mb.d;
    I.u = make_uniform(mb.d);
    I.off = mb.off;
    I.first = (int)b2i.size();
    I.mode = mb.mode;
    I.slot = mb.slot;
    const int nb = (mb.d.ncell + BLOCK - 1) / BLOCK;
    b2i.insert(b2i.end(), nb, (int)tab.size());
    tab.push_back(I);
  }
  nblocks_ = (unsigned)b2i.size();
  if (tab.empty()) return;
  tab_.allocate(tab.size() * sizeof(FBInst));
  HIP_CHECK(hipMemcpy(tab_.get(), tab.data(), tab.size() * sizeof(FBInst), hipMemcpyHostToDevice));
  b2i_.allocate(b2i.size());
  HIP_CHECK(hipMemcpy(b2i_.get(), b2i.data(), b2i.size() * sizeof(int), hipMemcpyHostToDevice));
}

__global__ void __launch_bounds__(BLOCK) k_field_uniform_b_err(const FBInst* __restrict__ tab,
                                                               const int* __restrict__ b2i,
                                                               const float* __restrict__ X,
                                                               const float* __restrict__ HD, float* __restrict__ OUT,
                                                               int cs, BatchErr be) {
  field_uniform_b_block_err(tab, b2i, X, HD, OUT, cs, xcd_remap(blockIdx.x, gridDim.x), be);
}

void FieldBatch::launch_err(const float* X, const float* HD, float* OUT, hipStream_t s, const BatchErr& err) const {
  if (nblocks_ == 0) return;
  if (err.nk < 0 || err.nk > 6) throw Error(MPA_ERR_RUNTIME, "batched field: bad stage count");
  k_field_uniform_b_err<<<nblocks_, BLOCK, 0, s>>>(reinterpret_cast<const FBInst*>(tab_.get()), b2i_.get(), X, HD,
                                                    OUT, cs_, err);
  checkLaunch("k_field_uniform_b_err");
}

void FieldBatch::launch(const float* X, const float* HD, float* OUT, hipStream_t s, const BatchNext* next) const {
  if (nblocks_ == 0) return;
  const FBInst* tab = reinterpret_cast<const FBInst*>(tab_.get());
  if (next) {
    if (next->nk < 0 || next->nk > 6) throw Error(MPA_ERR_RUNTIME, "batched field: bad stage count");
    const BatchNext& n = *next;
    switch (n.nk) {
      case 1: k_field_uniform_b<1><<<nblocks_, BLOCK, 0, s>>>(tab, b2i_.get(), X, HD, OUT, cs_, n); break;
      case 2: k_field_uniform_b<2><<<nblocks_, BLOCK, 0, s>>>(tab, b2i_.get(), X, HD, OUT, cs_, n); break;
      case 3: k_field_uniform_b<3><<<nblocks_, BLOCK, 0, s>>>(tab, b2i_.get(), X, HD, OUT, cs_, n); break;
      case 4: k_field_uniform_b<4><<<nblocks_, BLOCK, 0, s>>>(tab, b2i_.get(), X, HD, OUT, cs_, n); break;
      case 5: k_field_uniform_b<5><<<nblocks_, BLOCK, 0, s>>>(tab, b2i_.get(), X, HD, OUT, cs_, n); break;
      default: k_field_uniform_b<-1><<<nblocks_, BLOCK, 0, s>>>(tab, b2i_.get(), X, HD, OUT, cs_, n); break;
    }
  } else {
    k_field_uniform_b<0><<<nblocks_, BLOCK, 0, s>>>(tab, b2i_.get(), X, HD, OUT, cs_, BatchNext{});
  }
  checkLaunch("k_field_uniform_b");
}

__global__ void k_max_angle(FMDesc d, const float* __restrict__ m, float* __restrict__ out) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int N = d.ncell;
  if (idx >= N) return;
  if ((d.geom && !d.geom[idx]) || d.msat.at(idx) == 0.f) {
    out[idx] = 0.f;
    return;
  }
  const int x = idx % d.nx, y = (idx / d.nx) % d.ny, z = idx / (d.nx * d.ny);
  const float3 mc = load3(m, idx, N);
  const float a = d.aex.at(idx);
  float best = 0.f;
  for (int dir = 0; dir < 6; ++dir) {
    const int nb = neighbour(d, x, y, z, dir);
    if (nb < 0 || (d.geom && !d.geom[nb]) || d.msat.at(nb) == 0.f) continue;
    const float3 mn = load3(m, nb, N);
    const bool same = mn.x == mc.x && mn.y == mc.y && mn.z == mc.z;
    // acos of a dot product rounded above 1 is NaN and never the maximum
    const float ang = same ? 0.f : acosf(dot3(mc, mn));
    if (harmonic_mean(a, d.aex.at(nb)) != 0.f && ang > best) best = ang;
  }
  out[idx] = best;
}

void launch_max_angle(const FMDesc& d, const float* m, float* out, hipStream_t s) {
  k_max_angle<<<(d.ncell + BLOCK - 1) / BLOCK, BLOCK, 0, s>>>(d, m, out);
  checkLaunch("k_max_angle");
}

__global__ void k_param_fill(float* out, const float* base, const ParamFill f, int ncomp,
                             int N) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  for (int c = 0; c < ncomp; ++c) {
    const size_t j = (size_t)c * N + i;
    float v = (base ? base[j] : 0.f) + f.add[c];
    for (int k = 0; k < f.nmask; ++k) v += f.coef[k][c] * f.mask[k][j];
    out[j] = v;
  }
}

void launch_param_fill(float* out, const float* base, const ParamFill& f, int ncomp, int N, hipStream_t s) {
  if (N <= 0) return;
  if (f.nmask < 0 || f.nmask > ParamFill::MAX || ncomp < 1 || ncomp > 3)
    throw Error(MPA_ERR_RUNTIME, "launch_param_fill: bad arguments");
  k_param_fill<<<(N + BLOCK - 1) / BLOCK, BLOCK, 0, s>>>(out, base, f, ncomp, N);
  checkLaunch("k_param_fill");
}

void launch_field_pair(const FMDesc& d0, const float* m0, const float* h0, float* o0, const FMDesc& d1,
                       const float* m1, const float* h1, float* o1, FieldMode mode, hipStream_t s) {
  const bool u0 = params_uniform(d0), u1 = params_uniform(d1);
  if (d0.ncell != d1.ncell || u0 != u1 || (mode != MODE_EDENS && (all_uniform(d0) || all_uniform(d1)))) {
    launch_field(d0, m0, h0, o0, mode, s);
    launch_field(d1, m1, h1, o1, mode, s);
    return;
  }
  const dim3 grid((d0.ncell + BLOCK - 1) / BLOCK, 2);
  if (u0) k_field_pair<true><<<grid, BLOCK, 0, s>>>(d0, m0, h0, o0, d1, m1, h1, o1, (int)mode);
  else k_field_pair<false><<<grid, BLOCK, 0, s>>>(d0, m0, h0, o0, d1, m1, h1, o1, (int)mode);
  checkLaunch("k_field_pair");
}

bool field_pair_next_fusable(const FMDesc& d0, const FMDesc& d1, FieldMode mode) {
  const bool u0 = params_uniform(d0), u1 = params_uniform(d1);
  return mode == MODE_LLG && d0.ncell == d1.ncell && u0 == u1 && !all_uniform(d0) && !all_uniform(d1) &&
         !d0.noise && !d1.noise;
}

void launch_field_pair_next(const FMDesc& d0, const float* m0, const float* h0, float* o0, const NextStage& n0,
                            const FMDesc& d1, const float* m1, const float* h1, float* o1, const NextStage& n1,
                            FieldMode mode, hipStream_t s) {
  if (!field_pair_next_fusable(d0, d1, mode) || n0.nk != n1.nk || (n0.nk > 0 && (!n0.x0 || !n1.x0)))
    throw Error(MPA_ERR_RUNTIME, "launch_field_pair_next: not a fusable sublattice pair");
  const dim3 grid((d0.ncell + BLOCK - 1) / BLOCK, 2);
  const bool u = params_uniform(d0);
#define MPA_PAIR_NEXT(K)                                                                                        \
  if (u) k_field_pair_next<true, K><<<grid, BLOCK, 0, s>>>(d0, m0, h0, o0, n0, d1, m1, h1, o1, n1, (int)mode);  \
  else k_field_pair_next<false, K><<<grid, BLOCK, 0, s>>>(d0, m0, h0, o0, n0, d1, m1, h1, o1, n1, (int)mode);
  switch (n0.nk) {
    case 0: MPA_PAIR_NEXT(0) break;
    case 1: MPA_PAIR_NEXT(1) break;
    case 2: MPA_PAIR_NEXT(2) break;
    case 3: MPA_PAIR_NEXT(3) break;
    case 4: MPA_PAIR_NEXT(4) break;
    case 5: MPA_PAIR_NEXT(5) break;
    default: throw Error(MPA_ERR_RUNTIME, "launch_field_pair_next: 0 to 5 earlier derivatives");
  }
#undef MPA_PAIR_NEXT
  checkLaunch("k_field_pair_next");
}

bool torque_pair_splittable(const FMDesc& d0, const FMDesc& d1, FieldMode mode) {
  auto one = [&](const FMDesc& d) {
    return params_uniform(d) && !all_uniform(d) && (d.terms & TERM_DEMAG) && !(d.terms & TERM_STT) && !d.noise &&
           d.frozen.v == 0.f && !d.hstray;
  };
  return (mode == MODE_LLG || mode == MODE_DAMPING) && d0.ncell == d1.ncell && one(d0) && one(d1);
}

void launch_torque_pair(const FMDesc& d0, const float* m0, const float* hl0, const float* hd0, float* o0,
                        const FMDesc& d1, const float* m1, const float* hl1, const float* hd1, float* o1,
                        FieldMode mode, hipStream_t s) {
  if (!torque_pair_splittable(d0, d1, mode)) throw Error(MPA_ERR_RUNTIME, "launch_torque_pair: not splittable");
  const dim3 grid((d0.ncell + BLOCK - 1) / BLOCK, 2);
  k_torque_pair<<<grid, BLOCK, 0, s>>>(d0, m0, hl0, hd0, o0, d1, m1, hl1, hd1, o1, (int)mode);
  checkLaunch("k_torque_pair");
}

namespace {
// The configuration uniform_result<true> fixes at compile time: exchange on
// with Neumann boundaries (A and Ms non-zero, no open boundaries), DMI on or
// all nine constants zero, no cubic anisotropy, demag on, no stray field
// from other magnets, not frozen, LLG torque.
bool uniform_fast(const FMDesc& d, const FMUniform& u, FieldMode mode) {
  bool dmi_ok = u.has_dmi != 0;
  if (!dmi_ok) {
    dmi_ok = true;
    for (int k = 0; k < D_COUNT; ++k) dmi_ok = dmi_ok && u.D[k] == 0.f;
    for (int a = 0; a < 3; ++a) dmi_ok = dmi_ok && u.dxz[a] == 0.f && u.dxy[a] == 0.f && u.dyz[a] == 0.f;
  }
  return mode == MODE_LLG && u.has_ex && !u.openbc && !u.openbc_dmi && !u.cubic && u.demag && !u.frozen &&
         !d.hstray && dmi_ok;
}

// The halo-tile kernel covers magnets without a geometry mask or periodic
// wrap that are at least one tile wide, unless the world switched it off.
bool tile_eligible(const FMDesc& d) {
  return d.tile_zc != 0 && !d.geom && d.mgx == 0 && d.mgy == 0 && d.mgz == 0 && d.nx >= TILE_X && d.ny >= TILE_Y;
}

// Tile-schedule autotuner of the halo-tile field kernel: per grid size and
// kernel variant, the z planes per workgroup (TILE_ZC) are timed on the
// first launch (three launches each, the kernel's own outputs, which every
// candidate writes with the same values) and the fastest is kept for the
// process.
constexpr int TILE_NZC = 4;
constexpr int TILE_ZC[TILE_NZC] = {1, 2, 4, 8};
constexpr int TILE_ZC_DEFAULT = 4;

struct TileTune {
  int zc = 0;
  float us[TILE_NZC] = {0.f, 0.f, 0.f, 0.f};  // per launch; 0: not timed (beyond nz)
};

std::mutex tile_mu;
std::map<std::array<int, 4>, TileTune>& tile_table() {
  static std::map<std::array<int, 4>, TileTune> t;
  return t;
}

void launch_field_uniform(const FMDesc& d, const float* m, const float* hdemag, float* out, FieldMode mode,
                          const NextStage* next, hipStream_t s) {
  const FMUniform u = make_uniform(d);
  if (tile_eligible(d)) {
    const NextStage ns = next ? *next : NextStage{};
    const bool fast = uniform_fast(d, u, mode);
    auto launch_tile = [&](int zc_req) {
    const int zc = std::min(zc_req, d.nz);
    dim3 grid((d.nx + TILE_X - 1) / TILE_X, (d.ny + TILE_Y - 1) / TILE_Y, (d.nz + zc - 1) / zc);
    const dim3 blk(TILE_X, TILE_Y, 1);
    if (next && fast) {
      // derivative counts of the first-same-as-last tableaus' stages
      switch (ns.x0 ? ns.nk : -2) {
        case 1: k_field_tile<true, true, 1><<<grid, blk, 0, s>>>(d, u, m, hdemag, out, (int)mode, ns, zc); break;
        case 2: k_field_tile<true, true, 2><<<grid, blk, 0, s>>>(d, u, m, hdemag, out, (int)mode, ns, zc); break;
        case 3: k_field_tile<true, true, 3><<<grid, blk, 0, s>>>(d, u, m, hdemag, out, (int)mode, ns, zc); break;
        case 4: k_field_tile<true, true, 4><<<grid, blk, 0, s>>>(d, u, m, hdemag, out, (int)mode, ns, zc); break;
        case 5: k_field_tile<true, true, 5><<<grid, blk, 0, s>>>(d, u, m, hdemag, out, (int)mode, ns, zc); break;
        default: k_field_tile<true, true><<<grid, blk, 0, s>>>(d, u, m, hdemag, out, (int)mode, ns, zc); break;
      }
    } else if (next) k_field_tile<true, false><<<grid, blk, 0, s>>>(d, u, m, hdemag, out, (int)mode, ns, zc);
    else if (fast) k_field_tile<false, true><<<grid, blk, 0, s>>>(d, u, m, hdemag, out, (int)mode, ns, zc);
    else k_field_tile<false, false><<<grid, blk, 0, s>>>(d, u, m, hdemag, out, (int)mode, ns, zc);
    checkLaunch("k_field_tile");
    };
    if (d.tile_zc > 0) {
      launch_tile(d.tile_zc);
      return;
    }
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    HIP_CHECK(hipStreamIsCapturing(s, &cap));
    const std::array<int, 4> key{d.nx, d.ny, d.nz, (next ? 2 : 0) + (fast ? 1 : 0) + (next && ns.x0 ? 4 * ns.nk : 0)};
    int zc = 0;
    {
      std::lock_guard<std::mutex> lk(tile_mu);
      auto it = tile_table().find(key);
      if (it != tile_table().end()) zc = it->second.zc;
    }
    if (zc == 0 && cap != hipStreamCaptureStatusNone) zc = TILE_ZC_DEFAULT;  // no timing inside a capture
    if (zc == 0) {
      TileTune tt;
      hipEvent_t e0, e1;
      HIP_CHECK(hipEventCreate(&e0));
      HIP_CHECK(hipEventCreate(&e1));
      float best = 0.f;
      for (int c = 0; c < TILE_NZC; ++c) {
        if (c > 0 && TILE_ZC[c - 1] >= d.nz) break;  // the same schedule as the previous one
        launch_tile(TILE_ZC[c]);  // warm-up
        HIP_CHECK(hipEventRecord(e0, s));
        for (int r = 0; r < 3; ++r) launch_tile(TILE_ZC[c]);
        HIP_CHECK(hipEventRecord(e1, s));
        HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        tt.us[c] = ms * 1e3f / 3.f;
        if (tt.zc == 0 || tt.us[c] < best) {
          tt.zc = TILE_ZC[c];
          best = tt.us[c];
        }
      }
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      std::lock_guard<std::mutex> lk(tile_mu);
      tile_table()[key] = tt;
      zc = tt.zc;
    }
    launch_tile(zc);
    return;
  }
  int bx = 1;
  while (bx < d.nx && bx < MPA_FIELD_TILE_X) bx <<= 1;
  const int by = BLOCK / bx;
  dim3 grid((d.nx + bx - 1) / bx, (d.ny + by - 1) / by, (d.nz + MPA_FIELD_ZPER - 1) / MPA_FIELD_ZPER);
  if (next) {
    k_field_uniform<true><<<grid, dim3(bx, by, 1), 0, s>>>(d, u, m, hdemag, out, (int)mode, *next);
    checkLaunch("k_field_uniform<next>");
  } else {
    k_field_uniform<false><<<grid, dim3(bx, by, 1), 0, s>>>(d, u, m, hdemag, out, (int)mode, NextStage{});
    checkLaunch("k_field_uniform");
  }
}
}  // namespace

std::string field_tile_report() {
  std::lock_guard<std::mutex> lk(tile_mu);
  std::ostringstream o;
  for (const auto& kv : tile_table()) {
    o << kv.first[0] << "x" << kv.first[1] << "x" << kv.first[2] << " variant " << kv.first[3] << " | "
      << kv.second.zc;
    for (int c = 0; c < TILE_NZC; ++c) {
      o << " | " << TILE_ZC[c] << " ";
      if (kv.second.us[c] > 0.f) o << kv.second.us[c] << " us";
      else o << "-";
    }
    o << "\n";
  }
  return o.str();
}

void field_tile_clear() {
  std::lock_guard<std::mutex> lk(tile_mu);
  tile_table().clear();
}

bool field_next_fusable(const FMDesc& d, FieldMode mode) {
  return mode == MODE_LLG && all_uniform(d) && d.nz <= 65535;
}

void launch_field_next(const FMDesc& d, const float* m, const float* hdemag, float* out, FieldMode mode,
                       const NextStage& next, hipStream_t s) {
  if (!field_next_fusable(d, mode)) throw Error(MPA_ERR_RUNTIME, "launch_field_next: not the uniform LLG path");
  if (next.nk < 0 || next.nk > 7 || !next.out || next.out == m || next.out == out || next.out == next.x0)
    throw Error(MPA_ERR_RUNTIME, "launch_field_next: bad stage combination");
  launch_field_uniform(d, m, hdemag, out, mode, &next, s);
}

void launch_field(const FMDesc& d, const float* m, const float* hdemag, float* out, FieldMode mode, hipStream_t s) {
  if (mode != MODE_EDENS && all_uniform(d) && d.nz <= 65535) {
    launch_field_uniform(d, m, hdemag, out, mode, nullptr, s);
    return;
  }
  const int nblk = (d.ncell + BLOCK - 1) / BLOCK;
  if (params_uniform(d)) k_field<true><<<nblk, BLOCK, 0, s>>>(d, m, hdemag, out, (int)mode);
  else k_field<false><<<nblk, BLOCK, 0, s>>>(d, m, hdemag, out, (int)mode);
  checkLaunch("k_field");
}

}  // namespace mpa
