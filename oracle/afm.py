"""Antiferromagnet sublattice terms in FP64.

Test infrastructure only: imported by the tests, the smoke check and the
benchmark's CPU baseline, never by the product.

A sublattice `sub` is an oracle.model.Ferromagnet whose `host` is an
oracle.model.Antiferromagnet (host parameters afmex_cell, afmex_nn, latcon
and the inter-sublattice DMI tensor `host.dmi`).  `m` is the sublattice's
magnetisation, `m2` its sister's; msat2 is the sister's msat.

Restated functions (reference file:line):

* exchange_field_sublattice  k_exchangeField sublattice variant (Neumann
                             boundary with the sister's normal derivative),
                             src/physics/exchange.cu:106-202; chosen by
                             evalExchangeField only without open boundaries
                             (:224-236)
* dmi_field_afm              k_dmiFieldAFM, src/physics/dmi.cu:112-240
* afm_site_field             k_afmExchangeFieldSite, afmexchange.cu:28-48:
                             4 A_cell m2 / (latcon^2 Ms2)
* afm_nn_field               k_afmExchangeFieldNN, afmexchange.cu:51-149:
                             A_nn Laplacian of m2 / Ms2
* assured-zero rules         afmexchange.cu:13-25, dmi.cu:12-17

Pinned by tests/test_oracle_afm.py: the numpy formulations of
test/test_exchange_afm.py (sublattice, homogeneous and inhomogeneous
exchange), the reduction to the ferromagnet fields without AFM coupling
(the premise of test/test_DMI_*_afm.py), linearity in the inter-sublattice
tensor, and the AFM stray field of the total magnetisation.

Pinned by tests/test_oracle_afm_pins.py, with planted mistakes that fail:
* the sister-derivative terms of the Neumann ghost cells (exchange.cu:181-186,
  dmi.cu:205-207, afmexchange.cu:122-132): locked antiparallel sublattices
  equal a ferromagnet with A_eff = A - A_nn/2, edges included, for a state
  whose edge difference obeys the boundary condition;
* the yz elements of the inter-sublattice term (dmi.cu:229-231): locked
  parallel sublattices see them as the same elements of an own tensor;
* its structure: reciprocal for the xy and yz elements, anti-reciprocal for
  the xz elements.
Parity unpinned: the overall sign of the xy and xz elements of the
inter-sublattice term.  They are restated from dmi.cu:229-231; the xy
elements act with the sign opposite to the energy-derived one and the xz
elements follow no energy, so no closed form fixes them, and no reference
test or example has an observable that depends on them.
"""

import numpy as np

from .fields import DIRS, _DMI_AXIS, cross, get_gamma, harmonic_mean


def _sister(sub):
    h = sub.host
    return h.sub2 if sub is h.sub1 else h.sub1


def homo_assured_zero(sub):
    return sub.host is None or sub.host.assured_zero("afmex_cell") or _sister(sub).assured_zero("msat")


def inhomo_assured_zero(sub):
    return sub.host is None or sub.host.assured_zero("afmex_nn") or _sister(sub).assured_zero("msat")


def dmi_assured_zero(sub):
    if sub.assured_zero("msat"):
        return True
    own = all(sub.assured_zero(k) for k in sub.dmi)
    return own and all(sub.host.assured_zero(k) for k in sub.host.dmi)


def _other_side(sub, m, d):
    """(in-geometry mask, values) of the cell at -d (the bulk side of a boundary)."""
    o_in, o = sub.neighbor(tuple(-v for v in d))
    return o_in, m[:, o[0], o[1], o[2]]


def exchange_field_sublattice(sub, m, m2):
    """exchange.cu:106-202 (no open boundaries in this variant)."""
    nz, ny, nx = sub.shape
    h = np.zeros((3, nz, ny, nx))
    if sub.exchange_assured_zero():
        return h
    ms = sub.val("msat")
    a = sub.val("aex")
    an = sub.host.val("afmex_nn")
    D = sub.dmi_arrays()
    w = [1.0 / (c * c) for c in sub.cellsize]
    for d in DIRS:
        axis = int(np.flatnonzero(d)[0])
        nb_in, nb = sub.neighbor(d)
        delta = d[axis] * sub.cellsize[axis]
        n = tuple(abs(v) for v in d)
        o_in, m2o = _other_side(sub, m2, d)
        d_m2 = np.where(o_in[None], (m2 - m2o) / delta, 0.0)
        gamma1 = get_gamma(D, n, m)
        with np.errstate(divide="ignore", invalid="ignore"):
            m_neu = m + (an * cross(cross(d_m2, m), m) + gamma1) * delta / (2 * a)
        m_nb = m[:, nb[0], nb[1], nb[2]]
        include = np.where(nb_in, ms[nb] != 0, a != 0)
        m_ = np.where(nb_in[None], m_nb, m_neu)
        a_ = np.where(nb_in, a[nb], a)
        A = sub.region_link("inter_exchange", "scale_exchange", nb_in, nb, harmonic_mean(a, a_))
        h += np.where(include[None], 2 * A * w[axis] * (m_ - m), 0.0)
    act = sub.geom_mask() & (ms != 0)
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(act[None], h / ms, 0.0)


def dmi_field_afm(sub, m1, m2):
    """dmi.cu:112-240."""
    nz, ny, nx = sub.shape
    h = np.zeros((3, nz, ny, nx))
    if dmi_assured_zero(sub):
        return h
    sis = _sister(sub)
    ms, ms2 = sub.val("msat"), sis.val("msat")
    a = sub.val("aex")
    an = sub.host.val("afmex_nn")
    D = sub.dmi_arrays()
    Di = {k: sub.host.val(k) for k in sub.host.dmi}
    open_eff = np.logical_or(sub.openbc, a == 0)
    self_idx = np.indices((nz, ny, nx))
    for d in DIRS:
        axis = int(np.flatnonzero(d)[0])
        nb_in, nb = sub.neighbor(d)
        nbi = tuple(np.where(nb_in, nb[i], self_idx[i]) for i in range(3))
        skip = open_eff & (~nb_in | ((ms[nbi] == 0) & (ms2[nbi] == 0)))
        kxz, kxy, kyz = _DMI_AXIS[axis]
        Dxz, Dxy, Dyz = (harmonic_mean(D[k], D[k][nbi]) for k in (kxz, kxy, kyz))
        Dixz, Dixy, Diyz = (harmonic_mean(Di[k], Di[k][nbi]) for k in (kxz, kxy, kyz))
        Dzx, Dyx, Dzy = -Dxz, -Dxy, -Dyz
        Dizx, Diyx, Dizy = -Dixz, -Dixy, -Diyz
        delta = d[axis] * sub.cellsize[axis]
        n = tuple(abs(v) for v in d)
        # Neumann boundary (only reached without open boundaries)
        g1 = get_gamma(D, n, m1)
        g2 = get_gamma(D, n, m2)  # the sublattice's own tensor, as in dmi.cu:197-198
        o_in, m2o = _other_side(sub, m2, d)
        d_m2 = np.where(o_in[None], (m2 - m2o) / delta, 0.0)
        with np.errstate(divide="ignore", invalid="ignore"):
            m1_neu = m1 + (an * cross(cross(d_m2, m1), m1) + g1) * delta / (2 * a)
            m2_neu = m2 + (an * cross(cross((m1_neu - m1) / delta, m2), m2) + g2) * delta / (2 * a)
        m1_ = np.where(nb_in[None], m1[:, nbi[0], nbi[1], nbi[2]], m1_neu)
        m2_ = np.where(nb_in[None], m2[:, nbi[0], nbi[1], nbi[2]], m2_neu)
        s1 = (ms[nbi] != 0).astype(float)
        s2 = (ms2[nbi] != 0).astype(float)
        c = np.stack([
            (s1 * (Dxy * m1_[1] + Dxz * m1_[2]) - s2 * (Dixy * m2_[1] - Dixz * m2_[2])) / delta,
            (s1 * (Dyx * m1_[0] + Dyz * m1_[2]) - s2 * (Diyx * m2_[0] - Diyz * m2_[2])) / delta,
            (s1 * (Dzx * m1_[0] + Dzy * m1_[1]) - s2 * (Dizx * m2_[0] - Dizy * m2_[1])) / delta])
        h += np.where(skip[None], 0.0, np.nan_to_num(c))
    act = sub.geom_mask() & (ms != 0)
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(act[None], h / ms, 0.0)


def afm_site_field(sub, m2):
    """afmexchange.cu:28-48; divided by the sister's msat."""
    nz, ny, nx = sub.shape
    if homo_assured_zero(sub):
        return np.zeros((3, nz, ny, nx))
    ms2 = _sister(sub).val("msat")
    l = sub.host.val("latcon")
    with np.errstate(divide="ignore", invalid="ignore"):
        h = 4 * sub.host.val("afmex_cell") * m2 / (l * l) / ms2
    ok = sub.geom_mask() & (ms2 != 0)
    return np.where(ok[None], h, 0.0)


def afm_nn_field(sub, m1, m2):
    """afmexchange.cu:51-149; divided by the sister's msat."""
    nz, ny, nx = sub.shape
    h = np.zeros((3, nz, ny, nx))
    if inhomo_assured_zero(sub):
        return h
    ms2 = _sister(sub).val("msat")
    a = sub.val("aex")
    ann = sub.host.val("afmex_nn")
    D = sub.dmi_arrays()
    open_eff = np.logical_or(sub.openbc, a == 0)
    for d in DIRS:
        axis = int(np.flatnonzero(d)[0])
        nb_in, nb = sub.neighbor(d)
        delta = d[axis] * sub.cellsize[axis]
        n = tuple(abs(v) for v in d)
        include = ~(~nb_in & open_eff) & ((ms2[nb] != 0) | ~open_eff)
        o_in, m1o = _other_side(sub, m1, d)
        d_m1 = np.where(o_in[None], (m1 - m1o) / delta, 0.0)
        g2 = get_gamma(D, n, m2)
        with np.errstate(divide="ignore", invalid="ignore"):
            m2_neu = m2 + (ann * cross(cross(d_m1, m2), m2) + g2) * delta / (2 * a)
        m2_ = np.where(nb_in[None], m2[:, nb[0], nb[1], nb[2]], m2_neu)
        ann_ = np.where(nb_in, ann[nb], ann)
        A = sub.host.region_link("inter_afmex_nn", "scale_afmex_nn", nb_in, nb, harmonic_mean(ann, ann_))
        h += np.where(include[None], np.nan_to_num(A * (m2_ - m2) / (delta * delta)), 0.0)
    ok = sub.geom_mask() & (ms2 != 0)
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(ok[None], h / ms2, 0.0)
