"""Pins of the antiferromagnet's boundary and inter-sublattice terms in oracle/afm.py.

Three properties that the mathematics fixes, each checked to FP64 round-off:

* Locked antiparallel sublattices (m2 = -m1) behave as one ferromagnet with
  A_eff = A - A_nn/2.  The energy A(grad m1)^2 + A(grad m2)^2 + A_nn grad m1 .
  grad m2 becomes 2 A_eff (grad m1)^2, and the free-boundary condition
  2A dm1/dn + A_nn (dm2/dn)_perp = Gamma becomes 2 A_eff dm1/dn = Gamma.  The
  reference's Neumann ghost cells use the bulk difference next to the edge for
  dm2/dn (exchange.cu:181-186, dmi.cu:205-207, afmexchange.cu:122-132).  For a
  state whose edge difference already satisfies the boundary condition, the
  sublattice field (exchange + inhomogeneous AFM exchange + DMI) therefore
  equals the ferromagnet field cell for cell, edges included.
* Locked parallel sublattices (m2 = m1): the yz elements of the
  inter-sublattice tensor (xyz, yyz, zyz) act exactly as the same elements of
  a ferromagnet's own tensor, the field of the energy of dmi.hpp:6-16 summed
  over both sublattice orderings (m1_j d m2_k and m2_j d m1_k).
* Reciprocity.  A coupling that derives from an energy satisfies
  sum Ms1 u.h1[m2=w] = sum Ms2 w.h2[m1=u].  The reference's inter-sublattice
  term (dmi.cu:229-231) is reciprocal for the xy and yz elements and
  anti-reciprocal for the xz elements (it couples m_x and m_z symmetrically,
  which no energy does).  Both are fixed here.

What stays unpinned: the overall sign of the xy and xz elements of the
inter-sublattice term.  The xy elements act with the sign opposite to the
energy-derived one and the xz elements follow no energy, so no closed form
fixes them, and no test or example of the reference sets them with an
observable that depends on them (the Neel canting of test_neel_canting.py moves
by 1e-4 when the whole term changes sign, because a 1-D DMI energy is a total
derivative).

The mutation tests at the end plant each mistake in a copy of oracle/afm.py
and show that one of the pins fails.
"""

import inspect
import types

import numpy as np
import pytest

from oracle import afm as oa
from oracle import fields as of
from oracle import model as om


def _unit(v):
    return v / np.linalg.norm(v, axis=0)


def _rand_unit(rng, shape):
    return _unit(rng.normal(size=(3,) + shape))


# ---------------------------------------------------------------- pin 1

def _chain_state(rng, n, cell, gamma_of, a_eff):
    """Unit vectors on an x chain whose edge differences obey 2 A_eff dm/dn = Gamma.

    At x = 0 the outward normal is -x and the reference's one-sided difference is
    (m(1) - m(0))/c; at x = n-1 it is (m(n-1) - m(n-2))/c.  Both must have the
    perpendicular part Gamma(m_edge)/(2 A_eff)."""
    m = _unit(np.array([0.3, 0.4, 1.0])[:, None] + 0.5 * rng.normal(size=(3, n)))
    g0 = gamma_of(m[:, 0]) / (2 * a_eff)
    m[:, 1] = np.sqrt(1 - cell * cell * g0 @ g0) * m[:, 0] + cell * g0
    gn = gamma_of(m[:, n - 1]) / (2 * a_eff)
    m[:, n - 2] = np.sqrt(1 - cell * cell * gn @ gn) * m[:, n - 1] - cell * gn
    return m.reshape(3, 1, 1, n)


def locked_antiparallel_error(oa_mod, inter_scale=0.7, seed=0):
    """max |h_afm - h_fm| / max |h_fm| on both sublattices of an x chain (periodic in y, z)."""
    rng = np.random.default_rng(seed)
    n, c = 12, 1e-9
    A, Ann, Ms, D = 1.0e-11, -0.6e-11, 5e5, 2e-3
    a_eff = A - Ann / 2
    cs = (c, c, c)

    w = om.World(cs, mastergrid=(0, 1, 1), pbc=(0, 1, 1))
    afm = w.add_antiferromagnet((n, 1, 1), defer_timestep=True)
    for s in afm.sublattices:
        s.set("msat", Ms)
        s.set("aex", A)
        s.set("xyz", D)
        s.enable_demag = False
    afm.set("afmex_nn", Ann)
    afm.set("xyz", inter_scale * D)

    fw = om.World(cs, mastergrid=(0, 1, 1), pbc=(0, 1, 1))
    fm = fw.add_ferromagnet((n, 1, 1), defer_timestep=True)
    fm.set("msat", Ms)
    fm.set("aex", a_eff)
    fm.set("xyz", D)
    # the inter term on m2 = -m1 is minus a ferromagnet DMI field of tensor
    # inter_scale * D whose ghost cell is the same as fm's: aex = inter_scale * A_eff
    fi = fw.add_ferromagnet((n, 1, 1), defer_timestep=True)
    fi.set("msat", Ms)
    fi.set("aex", inter_scale * a_eff if inter_scale else a_eff)
    fi.set("xyz", inter_scale * D)

    gamma_of = lambda v: of.get_gamma(
        {k: np.full((1, 1, 1), D if k == "xyz" else 0.0) for k in fm.dmi}, (1, 0, 0),
        v.reshape(3, 1, 1, 1)).reshape(3)
    m1 = _chain_state(rng, n, c, gamma_of, a_eff)
    m2 = -m1
    h_fm = of.exchange_field(fm, m1) + of.dmi_field(fm, m1)
    if inter_scale:
        h_fm = h_fm - of.dmi_field(fi, m1)
    err = 0.0
    for sub, ma, mb, sign in ((afm.sub1, m1, m2, 1.0), (afm.sub2, m2, m1, -1.0)):
        h = (oa_mod.exchange_field_sublattice(sub, ma, mb) + oa_mod.dmi_field_afm(sub, ma, mb)
             + oa_mod.afm_nn_field(sub, ma, mb))
        err = max(err, float(np.abs(h - sign * h_fm).max() / np.abs(h_fm).max()))
    return err


# ---------------------------------------------------------------- pin 2

YZ = ("xyz", "yyz", "zyz")
XY = ("xxy", "yxy", "zxy")
XZ = ("xxz", "yxz", "zxz")


def _periodic_afm(rng, elements, size=(7, 6, 5), cs=(1.0e-9, 1.3e-9, 0.8e-9), ms=(5e5, 3e5)):
    w = om.World(cs, mastergrid=size, pbc=(1, 1, 1))
    afm = w.add_antiferromagnet(size, defer_timestep=True)
    for s, m in zip(afm.sublattices, ms):
        s.set("msat", m)
        s.set("aex", 1e-11)
        s.enable_demag = False
    vals = {k: float(rng.uniform(0.5e-3, 2e-3) * rng.choice([-1, 1])) for k in elements}
    for k, v in vals.items():
        afm.set(k, v)
    return w, afm, vals


def locked_parallel_error(oa_mod, seed=1):
    rng = np.random.default_rng(seed)
    w, afm, vals = _periodic_afm(rng, YZ, ms=(5e5, 5e5))
    fw = om.World(afm.cellsize, mastergrid=afm.shape[::-1], pbc=(1, 1, 1))
    fm = fw.add_ferromagnet(afm.shape[::-1], defer_timestep=True)
    fm.set("msat", 5e5)
    fm.set("aex", 1e-11)
    for k, v in vals.items():
        fm.set(k, v)
    m = _rand_unit(rng, afm.shape)
    want = of.dmi_field(fm, m)
    err = 0.0
    for sub in afm.sublattices:
        got = oa_mod.dmi_field_afm(sub, m, m)
        err = max(err, float(np.abs(got - want).max() / np.abs(want).max()))
    return err


# ---------------------------------------------------------------- pin 3

def reciprocity(oa_mod, elements, seed=2):
    """(S1, S2) with S1 = sum Ms1 u.h1[w], S2 = sum Ms2 w.h2[u] (inter term only)."""
    rng = np.random.default_rng(seed)
    w, afm, _ = _periodic_afm(rng, elements)
    u, v = _rand_unit(rng, afm.shape), _rand_unit(rng, afm.shape)
    ms1, ms2 = afm.sub1.val("msat"), afm.sub2.val("msat")
    h1 = oa_mod.dmi_field_afm(afm.sub1, u, v)  # own tensor zero: only the term in m2 = v
    h2 = oa_mod.dmi_field_afm(afm.sub2, v, u)
    return float(np.sum(ms1 * of.dot(u, h1))), float(np.sum(ms2 * of.dot(v, h2)))


def reciprocity_error(oa_mod):
    """Largest violation of: xy and yz reciprocal, xz anti-reciprocal."""
    err = 0.0
    for elements, sign in ((XY, 1.0), (YZ, 1.0), (XZ, -1.0)):
        s1, s2 = reciprocity(oa_mod, elements)
        err = max(err, abs(s1 - sign * s2) / max(abs(s1), abs(s2)))
    return err


# ---------------------------------------------------------------- tests

def test_locked_antiparallel_equals_ferromagnet_with_effective_exchange():
    assert locked_antiparallel_error(oa, inter_scale=0.0) < 1e-11
    assert locked_antiparallel_error(oa, inter_scale=0.7) < 1e-11


def test_locked_antiparallel_state_is_not_trivial():
    """The edge cells carry a boundary term comparable to the bulk one, so the
    identity above is not satisfied by the bulk alone."""
    rng = np.random.default_rng(0)
    n, c, a_eff, D = 12, 1e-9, 1.3e-11, 2e-3
    gamma_of = lambda v: np.array([0.0, -D * v[2], D * v[1]])
    m = _chain_state(rng, n, c, gamma_of, a_eff).reshape(3, n)
    # the edge difference is as large as the boundary condition demands
    g = gamma_of(m[:, 0]) / (2 * a_eff)
    assert np.linalg.norm(g) * c > 0.01


def test_inter_yz_elements_act_as_own_tensor_when_parallel():
    assert locked_parallel_error(oa) < 1e-12


@pytest.mark.parametrize("elements,sign", [(XY, 1.0), (YZ, 1.0), (XZ, -1.0)])
def test_inter_term_reciprocity(elements, sign):
    s1, s2 = reciprocity(oa, elements)
    assert abs(s1) > 0
    assert abs(s1 - sign * s2) <= 1e-12 * abs(s1)


# ---------------------------------------------------------------- mutations

_SRC = inspect.getsource(oa)

MUTATIONS = {
    # the whole inter-sublattice term of dmi.cu:229-231 with the opposite sign
    "inter_sign": [("- s2 * (Dixy", "+ s2 * (Dixy"), ("- s2 * (Diyx", "+ s2 * (Diyx"),
                   ("- s2 * (Dizx", "+ s2 * (Dizx")],
    # one line of it
    "inter_x_line": [("- s2 * (Dixy", "+ s2 * (Dixy")],
    "inter_z_line": [("- s2 * (Dizx", "+ s2 * (Dizx")],
    # an element swapped within a line
    "inter_xz_sign": [("Dixy * m2_[1] - Dixz", "Dixy * m2_[1] + Dixz")],
    # the sister-derivative terms of the Neumann ghost cells
    "exchange_sister": [("m_neu = m + (an * cross(cross(d_m2, m), m) + gamma1)", "m_neu = m + (gamma1)")],
    "exchange_sister_sign": [("m_neu = m + (an * cross(cross(d_m2, m), m) + gamma1)",
                              "m_neu = m + (-an * cross(cross(d_m2, m), m) + gamma1)")],
    "dmi_sister_m1": [("m1_neu = m1 + (an * cross(cross(d_m2, m1), m1) + g1)", "m1_neu = m1 + (g1)")],
    "dmi_sister_m2": [("m2_neu = m2 + (an * cross(cross((m1_neu - m1) / delta, m2), m2) + g2)",
                       "m2_neu = m2 + (g2)")],
    "nn_sister": [("m2_neu = m2 + (ann * cross(cross(d_m1, m2), m2) + g2)", "m2_neu = m2 + (g2)")],
}


def _mutant(name):
    src = _SRC
    for old, new in MUTATIONS[name]:
        assert old in src, f"mutation {name}: text not found in oracle/afm.py"
        src = src.replace(old, new)
    mod = types.ModuleType("oracle.afm_mutant")
    mod.__package__ = "oracle"
    exec(compile(src, f"<oracle/afm.py:{name}>", "exec"), mod.__dict__)
    return mod


@pytest.mark.parametrize("name", sorted(MUTATIONS))
def test_planted_mistake_fails_a_pin(name):
    mod = _mutant(name)
    with np.errstate(all="ignore"):
        errs = {
            "antiparallel": max(locked_antiparallel_error(mod, 0.0), locked_antiparallel_error(mod, 0.7)),
            "parallel": locked_parallel_error(mod),
            "reciprocity": reciprocity_error(mod),
        }
    assert max(errs.values()) > 1e-6, errs
