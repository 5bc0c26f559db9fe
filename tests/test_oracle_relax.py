"""Pins of oracle/relax.py (FP64 relaxation and energy minimisation).

* minimize: the standard problem 2 coercivity in the small-particle limit
  (test/test_standardproblem2_coercivity.py: H_c/Ms = 0.057069478 within
  1e-5) and the Stoner-Wohlfarth equilibrium angle of a single spin;
* relax: the surface canting of the Neel vector of an antiferromagnet with
  bulk DMI (test/test_neel_canting.py, analytic tanh profile of
  PhysRevB.103.134413, 2 %), on a chain that is periodic in y and z, and the
  Stoner-Wohlfarth angle;
* the solver state restored by Relaxer::exec (relaxer.cpp:167-173).
"""

import math

import numpy as np
import pytest

from oracle import model as om
from oracle import relax as orl

MU0 = 1.25663706212e-6
GAMMALL = 1.7595e11
HBAR = 1.054571817e-34


def _sp2_cell(field):
    msat, aex = 800e3, 13e-12
    l_ex = math.sqrt(2 * aex / (MU0 * msat ** 2))
    d = 0.1
    w = om.World((5.0 * d * l_ex, d * l_ex, 0.1 * d * l_ex))
    f = w.add_ferromagnet((1, 1, 1), defer_timestep=True)
    f.set("msat", msat)
    f.set("aex", aex)
    return w, f, msat


def test_sp2_coercivity_minimize():
    w, f, msat = _sp2_cell(0.0)
    f.set_m(np.array([1.0, 1.0, 1.0]))
    orl.minimize(w, f)
    hdir = -np.ones(3) / math.sqrt(3)

    def m_along():
        return float(np.sum(f.m.reshape(3, -1).mean(axis=1)))

    h0, m0 = 0.0, m_along()
    h1 = 0.0570
    w.bias = tuple(MU0 * h1 * msat * hdir)
    orl.minimize(w, f)
    m1 = m_along()
    while m1 > 0:
        h0, m0 = h1, m1
        h1 += 0.00001
        w.bias = tuple(MU0 * h1 * msat * hdir)
        orl.minimize(w, f)
        m1 = m_along()
        assert h1 < 0.0575
    hc = (m0 * h1 - m1 * h0) / (m0 - m1)
    assert abs(hc - 0.057069478) < 1e-5


def _stoner_wohlfarth(psi, h):
    """Equilibrium angle theta of m from the easy axis x for a field h (in
    units of 2K/Ms) at angle psi: minimise sin^2(theta) - 2h cos(theta - psi),
    starting from theta = 0."""
    theta = 0.0
    for _ in range(200):
        g = 2 * math.sin(theta) * math.cos(theta) + 2 * h * math.sin(theta - psi)
        gp = 2 * math.cos(2 * theta) + 2 * h * math.cos(theta - psi)
        theta -= g / gp
    return theta


@pytest.mark.parametrize("how", ["minimize", "relax"])
def test_stoner_wohlfarth_single_spin(how):
    K, msat = 5e5, 8e5
    w = om.World((2e-9, 2e-9, 2e-9))
    f = w.add_ferromagnet((1, 1, 1), defer_timestep=True)
    f.set("msat", msat)
    f.set("alpha", 0.3)
    f.set("ku1", K)
    f.set("anisU", (1.0, 0.0, 0.0))
    f.enable_demag = False
    psi, h = 2.0, 0.3  # below the switching field at this angle
    b = h * 2 * K / msat
    w.bias = (b * math.cos(psi), b * math.sin(psi), 0.0)
    f.set_m(np.array([1.0, 0.01, 0.0]))
    w.solver.equations_changed()
    if how == "minimize":
        orl.minimize(w, f, tol=1e-9, nsamples=10)
    else:
        orl.relax(w, f, tol=1e-9)
    theta = math.atan2(f.m[1].item(), f.m[0].item())
    assert abs(theta - _stoner_wohlfarth(psi, h)) < 1e-4


def test_neel_canting_relax_chain():
    a = 0.35e-9
    J, H = 2.34e-22, 9.36e-23
    A = J / (2 * a)
    Ku = H / (2 * a ** 3)
    Ms = GAMMALL * HBAR / (2 * a ** 3)
    N = 32
    w = om.World((2 * a, 2 * a, 2 * a), mastergrid=(0, 1, 1), pbc=(0, 1, 1))
    afm = w.add_antiferromagnet((N, 1, 1), defer_timestep=True)
    for s in afm.sublattices:
        for k, v in {"alpha": 0.25, "msat": Ms, "aex": A, "anisU": (0, 0, 1), "ku1": Ku}.items():
            s.set(k, v)
        s.enable_demag = False
    afm.set("afmex_nn", -A)
    afm.sub1.set_m(np.array([0.0, 0.0, 1.0]))
    afm.sub2.set_m(np.array([0.0, 0.0, -1.0]))
    Dc = 4 * math.sqrt(A / 2 * Ku) / math.pi
    D = 0.55 * Dc
    for t in (afm, afm.sub1, afm.sub2):
        for k, v in {"xyz": D, "yxz": -D, "zxy": D}.items():
            t.set(k, v)
    w.solver.equations_changed()
    orl.relax(w, afm, 1e-9)
    n = afm.sub1.val("msat") * afm.sub1.m - afm.sub2.val("msat") * afm.sub2.m
    cant = n[2] / np.linalg.norm(n, axis=0)
    k = math.sqrt(1 - (2 * (D / 2) / (math.pi * Dc)) ** 2)
    # analytical(x, D) of test_neel_canting.py:19-21 at the surface cell x = 0
    x = 0
    want = math.tanh(math.atanh(k) + N // 2 - abs(x - N // 2))
    assert abs(want - k) < 1e-12
    assert abs(cant[0, 0, 0] - want) / want < 2e-2
    # the canting decays into the bulk
    assert cant[0, 0, 0] < cant[0, 0, 4] < 1.0


def test_relax_restores_solver_state():
    w = om.World((3e-9, 3e-9, 3e-9))
    f = w.add_ferromagnet((8, 4, 1), defer_timestep=True)
    f.set("msat", 8e5)
    f.set("aex", 1.3e-11)
    f.set("alpha", 0.2)
    rng = np.random.default_rng(1)
    f.set_m(rng.normal(size=(3, 1, 4, 8)))
    ts = w.solver
    ts.set_method("DormandPrince")
    ts.adaptive = False
    ts.dt = 2e-14
    ts.time = 3e-12
    ts.max_error = 2e-5
    e0 = orl.total_energy(w, f)
    orl.relax(w, f, 1e-7)
    assert orl.total_energy(w, f) < e0
    assert (ts.time, ts.dt, ts.adaptive, ts.max_error, ts.method) == (3e-12, 2e-14, False, 2e-5, "DormandPrince")
    assert w.relax_fms is None
