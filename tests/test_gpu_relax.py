"""relax() and minimize() on the GPU, against the closed forms the reference's
tests use: the standard problem 2 coercivity in the small-particle limit
(test/test_standardproblem2_coercivity.py, H_c/Ms = 0.057069478 within
1e-5, minimize) and the Neel-vector canting at an antiferromagnet's surface
(test/test_neel_canting.py, 2 %, relax); plus the solver bookkeeping of
Relaxer::exec (relaxer.cpp:87-174) and the torque-threshold mode.
"""

import warnings

import numpy as np
import pytest

import mumaxplus as mp
from mumaxplus.util.constants import GAMMALL, HBAR, MU0
from oracle import model as omodel
from oracle import relax as orelax

pytestmark = pytest.mark.gpu


def test_standard_problem2_coercivity_by_minimize():
    """Single cell of aspect 5 : 1 : 0.1 (in exchange lengths, d = 0.1):
    raise a field along -(1,1,1) until <m>.(1,1,1) changes sign."""
    msat, aex = 800e3, 13e-12
    l_ex = np.sqrt(2 * aex / (MU0 * msat ** 2))
    d = 0.1
    world = mp.World(cellsize=(5.0 * d * l_ex, d * l_ex, 0.1 * d * l_ex))
    magnet = mp.Ferromagnet(world, mp.Grid((1, 1, 1)))
    magnet.msat = msat
    magnet.aex = aex
    magnet.magnetization = (1, 1, 1)
    magnet.minimize()
    h_dir = -np.ones(3) / np.sqrt(3)
    h0, m0 = 0.0, float(np.sum(magnet.magnetization.average()))
    h1 = 0.05
    world.bias_magnetic_field = tuple(MU0 * h1 * msat * h_dir)
    magnet.minimize()
    m1 = float(np.sum(magnet.magnetization.average()))
    while m1 > 0:
        h0, m0 = h1, m1
        h1 += 0.00005
        world.bias_magnetic_field = tuple(MU0 * h1 * msat * h_dir)
        magnet.minimize()
        m1 = float(np.sum(magnet.magnetization.average()))
        assert h1 < 0.06, "no switching below 0.06 Ms"
    hc = (m0 * h1 - m1 * h0) / (m0 - m1)
    assert abs(hc - 0.057069478) < 1e-5


def _neel_canting_magnet(N):
    a = 0.35e-9
    J, H = 2.34e-22, 9.36e-23
    A = J / (2 * a)
    Ku = H / (2 * a ** 3)
    Ms = GAMMALL * HBAR / (2 * a ** 3)
    world = mp.World((2 * a, 2 * a, 2 * a))
    magnet = mp.Antiferromagnet(world, mp.Grid((N, N, N)))
    magnet.alpha = 0.25
    magnet.msat = Ms
    magnet.aex = A
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", UserWarning)
        magnet.afmex_nn = -A
    magnet.anisU = (0, 0, 1)
    magnet.ku1 = Ku
    magnet.sub1.magnetization = (0, 0, 1)
    magnet.sub2.magnetization = (0, 0, -1)
    magnet.enable_demag = False
    magnet.enable_openbc = False
    Dc = 4 * np.sqrt(A / 2 * Ku) / np.pi
    return world, magnet, Dc


@pytest.mark.parametrize("frac", [0.1, 0.55])
def test_neel_canting_by_relax(frac):
    """Bulk DMI D on a 32^3 antiferromagnet: the surface canting of the Neel
    vector follows tanh(artanh(k) + N/2 - |x - N/2|), k = sqrt(1 - (2 D' /
    (pi Dc))^2) with D' = D / 2 (PhysRevB.103.134413)."""
    N = 32
    world, magnet, Dc = _neel_canting_magnet(N)
    D = frac * Dc
    magnet.dmi_tensors.set_bulk_dmi(D)
    magnet.relax()
    neel = magnet.neel_vector.eval()
    cant = (neel[2] / np.linalg.norm(neel, axis=0))[N // 2, N // 2, :]
    k = np.sqrt(1 - (2 * (D / 2) / (np.pi * Dc)) ** 2)
    want = np.tanh(np.arctanh(k) + (N // 2 - abs(0 - N // 2)))
    assert abs(cant[0] - want) / abs(want) < 2e-2


def test_relax_restores_solver_state_and_lowers_energy():
    rng = np.random.default_rng(8)
    world = mp.World((3e-9, 3e-9, 3e-9))
    magnet = mp.Ferromagnet(world, mp.Grid((24, 12, 2)))
    magnet.msat = 8e5
    magnet.aex = 1.3e-11
    magnet.alpha = 0.2
    magnet.ku1 = 4e5
    magnet.anisU = (1, 0, 0)
    m = rng.normal(size=(3, 2, 12, 24))
    magnet.magnetization = m / np.linalg.norm(m, axis=0)
    ts = world.timesolver
    ts.set_method("DormandPrince")
    ts.max_error = 3e-5
    ts.time = 1.5e-12
    e0 = magnet.total_energy.eval()
    magnet.relax()
    assert magnet.total_energy.eval() < e0
    assert ts.time == 1.5e-12
    assert ts._impl.method == "DormandPrince"
    assert ts.max_error == pytest.approx(3e-5)
    assert ts.adaptive_timestep
    assert ts.timestep == pytest.approx(ts.sensible_timestep, rel=1e-6)
    # fixed-step solvers keep their step
    ts.adaptive_timestep = False
    ts.timestep = 1e-14
    magnet.relax(1e-6)
    assert ts.timestep == 1e-14 and not ts.adaptive_timestep
    # the relaxed state is close to a minimum: small damping torque
    tq = magnet.damping_torque.eval()
    assert np.linalg.norm(tq, axis=0).max() < 1e-3 * GAMMALL * 1.0


def test_relax_torque_threshold_and_minimize_agree():
    """With a positive RelaxTorqueThreshold the tolerance is lowered only
    while the max damping torque is below it (at the default tolerance the
    adaptive damping dynamics of this film settle near 2e7 /s, so the
    threshold is set above that); minimize reaches the same state, and so
    does the FP64 oracle's minimize."""
    def build():
        world = mp.World((2e-9, 2e-9, 2e-9))
        magnet = mp.Ferromagnet(world, mp.Grid((32, 8, 1)))
        magnet.msat = 8e5
        magnet.aex = 1.3e-11
        magnet.alpha = 0.5
        magnet.magnetization = (1, 0.3, 0.1)
        return world, magnet

    world, a = build()
    a.RelaxTorqueThreshold = 1e8
    assert a.RelaxTorqueThreshold == 1e8
    a.relax(1e-7)
    assert np.linalg.norm(a.damping_torque.eval(), axis=0).max() <= 1e8
    world2, b = build()
    b.minimize()
    ma, mb = np.array(a.magnetization.average()), np.array(b.magnetization.average())
    assert np.abs(ma - mb).max() < 2e-3
    ow = omodel.World((2e-9, 2e-9, 2e-9))
    of_ = ow.add_ferromagnet((32, 8, 1), defer_timestep=True)
    for k, v in {"msat": 8e5, "aex": 1.3e-11, "alpha": 0.5}.items():
        of_.set(k, v)
    of_.set_m(np.array([1.0, 0.3, 0.1]))
    orelax.minimize(ow, of_)
    assert np.abs(mb - of_.m.reshape(3, -1).mean(axis=1)).max() < 1e-4
    assert np.abs(b.magnetization.get() - of_.m).max() < 1e-3
    with pytest.raises(AssertionError):
        a.RelaxTorqueThreshold = 0


@pytest.mark.parametrize("how", ["minimize", "relax"])
def test_stoner_wohlfarth_single_spin(how):
    """Uniaxial single spin in a field below the switching field at 2 rad
    from the easy axis: the equilibrium angle of the Stoner-Wohlfarth
    energy sin^2(theta) - 2 h cos(theta - psi), and the oracle's result."""
    K, msat, psi, h = 5e5, 8e5, 2.0, 0.3
    world = mp.World((2e-9, 2e-9, 2e-9))
    mag = mp.Ferromagnet(world, mp.Grid((1, 1, 1)))
    mag.msat = msat
    mag.alpha = 0.3
    mag.ku1 = K
    mag.anisU = (1, 0, 0)
    mag.enable_demag = False
    b = h * 2 * K / msat
    world.bias_magnetic_field = (b * np.cos(psi), b * np.sin(psi), 0.0)
    mag.magnetization = (1.0, 0.01, 0.0)
    if how == "minimize":
        mag.minimize(tol=1e-7)
    else:
        mag.relax()
    m = mag.magnetization.average()
    theta = np.arctan2(m[1], m[0])
    want = 0.0
    for _ in range(200):  # Newton on the energy derivative, from theta = 0
        g = np.sin(2 * want) + 2 * h * np.sin(want - psi)
        want -= g / (2 * np.cos(2 * want) + 2 * h * np.cos(want - psi))
    assert abs(theta - want) < 1e-4


def test_world_relax_and_minimize_cover_every_magnet():
    world = mp.World((2e-9, 2e-9, 2e-9))
    fm = mp.Ferromagnet(world, mp.Grid((16, 8, 1)), name="fm")
    afm = mp.Antiferromagnet(world, mp.Grid((8, 8, 1), origin=(40, 0, 0)), name="afm")
    for mag in (fm, afm):
        mag.enable_as_stray_field_source = False
        mag.enable_as_stray_field_destination = False
    fm.msat = 8e5
    fm.aex = 1.3e-11
    fm.alpha = 0.3
    fm.magnetization = (1, 1, 0)
    afm.msat = 5.66e5
    afm.aex = 2.48e-12
    afm.alpha = 0.3
    afm.ku1 = 6.4e4
    afm.anisU = (0, 0, 1)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", UserWarning)
        afm.afmex_cell = -4e-12
    afm.sub1.magnetization = (0.3, 0, 1)
    afm.sub2.magnetization = (0, 0.2, -1)
    e_fm, e_afm = fm.total_energy.eval(), afm.total_energy.eval()
    world.relax()
    assert fm.total_energy.eval() <= e_fm and afm.total_energy.eval() < e_afm
    afm.sub1.magnetization = (0.3, 0, 1)
    world.minimize()
    n = np.array(afm.neel_vector.average())
    assert abs(abs(n[2]) - 1) < 1e-3
