"""Pins of the oracle's antiferromagnet terms (oracle/afm.py) and of the
elastodynamic equations of oracle/model.World.

* the reference tests' numpy formulations of test/test_exchange_afm.py:7-107
  (sublattice exchange, homogeneous and inhomogeneous AFM exchange);
* reductions: with afmex_nn = 0 and no inter-sublattice DMI, the sublattice
  exchange and DMI fields equal the ferromagnet ones (the premise of
  test/test_DMI_1D_afm.py:1-4);
* a longitudinal standing elastic wave under fixed-step Dormand-Prince
  integration against u(t) = A sin(kx) cos(w t), w = k sqrt(C11/rho) with
  the fourth-order stencil's dispersion.
"""

import math

import numpy as np
import pytest

from oracle import afm as oa
from oracle import fields as of
from oracle import model as om


def rand_unit(rng, shape):
    v = rng.normal(size=(3,) + shape)
    return v / np.linalg.norm(v, axis=0)


def make_afm(size=(16, 16, 4), cs=(1e3, 2e3, 3e3), **host):
    w = om.World(cs)
    afm = w.add_antiferromagnet(size, defer_timestep=True)
    for k, v in host.items():
        afm.set(k, v)
    return w, afm


def numpy_laplacian_open(m, cs):
    """test/test_exchange_afm.py:7-33 / 41-64 (open edges)."""
    h = np.zeros(m.shape)
    for axis, c in ((1, cs[2]), (2, cs[1]), (3, cs[0])):
        for shift, keep in ((1, slice(1, None)), (-1, slice(0, -1))):
            m_ = np.roll(m, shift, axis=axis)
            sl = [slice(None)] * 4
            sl[axis] = keep
            h[tuple(sl)] += (m_ - m)[tuple(sl)] / c ** 2
    return h


def test_sublattice_exchange_numpy_formulation():
    rng = np.random.default_rng(0)
    w, afm = make_afm()
    for sub in afm.sublattices:
        sub.set("aex", 3.2e7)
        sub.set("msat", 5.4)
    m1, m2 = rand_unit(rng, afm.shape), rand_unit(rng, afm.shape)
    for sub, m, ms in ((afm.sub1, m1, m2), (afm.sub2, m2, m1)):
        got = oa.exchange_field_sublattice(sub, m, ms)
        want = 2 * 3.2e7 * numpy_laplacian_open(m, afm.cellsize) / 5.4
        assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


def test_homogeneous_exchange_numpy_formulation():
    rng = np.random.default_rng(1)
    w, afm = make_afm(afmex_cell=-10e4)
    afm.sub1.set("msat", 5.4)
    afm.sub2.set("msat", 3.1)  # the field is divided by the sister's msat
    m1, m2 = rand_unit(rng, afm.shape), rand_unit(rng, afm.shape)
    l = 0.35e-9
    got1 = oa.afm_site_field(afm.sub1, m2)
    assert np.allclose(got1, 4 * -10e4 * m2 / (l * l * 3.1), rtol=1e-13, atol=0)
    got2 = oa.afm_site_field(afm.sub2, m1)
    assert np.allclose(got2, 4 * -10e4 * m1 / (l * l * 5.4), rtol=1e-13, atol=0)


def test_inhomogeneous_exchange_numpy_formulation():
    rng = np.random.default_rng(2)
    w, afm = make_afm(afmex_nn=-10e4)
    for sub in afm.sublattices:
        sub.set("msat", 5.4)  # aex = 0: open boundaries
    m1, m2 = rand_unit(rng, afm.shape), rand_unit(rng, afm.shape)
    got = oa.afm_nn_field(afm.sub1, m1, m2)
    want = -10e4 * numpy_laplacian_open(m2, afm.cellsize) / 5.4
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("kind", ["interfacial", "bulk"])
@pytest.mark.parametrize("openbc", [False, True])
def test_reduces_to_ferromagnet_without_afm_coupling(kind, openbc):
    rng = np.random.default_rng(3)
    size, cs = (9, 7, 3), (2e-9, 3e-9, 2.5e-9)
    w, afm = make_afm(size, cs)
    fw = om.World(cs)
    fm = fw.add_ferromagnet(size, defer_timestep=True)
    dmi = {"interfacial": {"xxz": 3e-3, "yyz": 3e-3}, "bulk": {"xyz": 3e-3, "yxz": -3e-3, "zxy": 3e-3}}[kind]
    for mag in (afm.sub1, afm.sub2, fm):
        mag.set("msat", 8.6e5)
        mag.set("aex", 1.3e-11)
        mag.openbc = openbc
        mag.enable_demag = False  # the AFM demag field is that of both sublattices
        for k, v in dmi.items():
            mag.set(k, v)
    m1, m2 = rand_unit(rng, afm.shape), rand_unit(rng, afm.shape)
    assert np.allclose(oa.dmi_field_afm(afm.sub1, m1, m2), of.dmi_field(fm, m1), rtol=1e-12, atol=1e-6)
    # the sister-aware exchange is only used without open boundaries (exchange.cu:224-236)
    if not openbc:
        assert np.allclose(oa.exchange_field_sublattice(afm.sub1, m1, m2), of.exchange_field(fm, m1),
                           rtol=1e-12, atol=1e-6)
    h_world = w.effective_field(afm.sub1, m1, m2, afm.u)
    h_fm = fw.effective_field(fm, m1, None, fm.u)
    assert np.allclose(h_world, h_fm, rtol=1e-12, atol=1e-6)


def test_inter_sublattice_dmi_contributes():
    """The inter-sublattice tensor adds a term linear in m2 (dmi.cu:231-233)
    and leaves the field linear in D (doubling D doubles the field)."""
    rng = np.random.default_rng(4)
    w, afm = make_afm((6, 5, 2), (2e-9, 2e-9, 2e-9))
    for sub in afm.sublattices:
        sub.set("msat", 8e5)
        sub.set("aex", 1e-11)
    afm.set("xyz", 1e-3)
    afm.set("yxz", -1e-3)
    afm.set("zxy", 1e-3)
    m1, m2 = rand_unit(rng, afm.shape), rand_unit(rng, afm.shape)
    h1 = oa.dmi_field_afm(afm.sub1, m1, m2)
    assert np.abs(h1).max() > 0
    for k in ("xyz", "yxz", "zxy"):
        afm.set(k, 2 * afm.dmi[k])
    assert np.allclose(oa.dmi_field_afm(afm.sub1, m1, m2), 2 * h1, rtol=1e-12, atol=0)


def test_standing_elastic_wave():
    n, cs = 64, (2e-9, 2e-9, 2e-9)
    w = om.World(cs, mastergrid=(n, 0, 0), pbc=(1, 0, 0))
    mag = w.add_ferromagnet((n, 1, 1), defer_timestep=True)
    mag.enable_demag = False
    mag.enable_elastodynamics = True
    C11, rho = 283e9, 8900.0
    mag.set("C11", C11)
    mag.set("rho", rho)
    k = 2 * math.pi * 2 / (n * cs[0])
    x = np.arange(n) * cs[0]
    A = 1e-12
    mag.u[0] = A * np.sin(k * x)[None, None, :]
    # the discrete 5-point second derivative of sin(kx), twice the 5-point first derivative
    kd = (4 / 3 * math.sin(k * cs[0]) - 1 / 6 * math.sin(2 * k * cs[0])) / cs[0]
    omega = kd * math.sqrt(C11 / rho)
    w.solver.set_method("DormandPrince")
    w.solver.adaptive = False
    w.solver.dt = 2e-14
    nsteps = 400
    w.solver.steps(nsteps)
    t = w.solver.time
    want = A * np.sin(k * x) * math.cos(omega * t)
    assert np.abs(mag.u[0, 0, 0] - want).max() < 1e-4 * A


def test_afm_demag_is_that_of_the_total_magnetisation():
    """strayfieldfft.cu:198-213: the host's stray field of Ms1 m1 + Ms2 m2.
    Parallel sublattices act as one magnet of twice the moment; antiparallel
    ones with equal msat have no stray field."""
    rng = np.random.default_rng(5)
    size, cs = (8, 6, 2), (2e-9, 2e-9, 2e-9)
    w, afm = make_afm(size, cs)
    fw = om.World(cs)
    fm = fw.add_ferromagnet(size, defer_timestep=True)
    fm.set("msat", 2 * 6e5)
    fm.enable_demag = True
    for sub in afm.sublattices:
        sub.set("msat", 6e5)
    m = rand_unit(rng, afm.shape)
    h = w.effective_field(afm.sub1, m, m, afm.u)
    assert np.allclose(h, fw.effective_field(fm, m, None, fm.u), rtol=1e-12, atol=1e-9)
    h0 = w.effective_field(afm.sub1, m, -m, afm.u)
    assert np.abs(h0).max() < 1e-9 * np.abs(h).max()
