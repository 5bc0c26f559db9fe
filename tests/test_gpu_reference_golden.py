"""Full-path GPU runs against golden numbers that the reference's own tests
hold (results of mumax3 runs of the same problems).

* test/test_PBC.py:14-104: a 64x32 rectangle periodically repeated on a
  128x64 mastergrid (2 image repetitions per axis), Ms=1e6, A=1e-11,
  alpha=1, m0=(1, .1, .01); <m> after 1 ns of adaptive Fehlberg time
  stepping (demag with periodic images, exchange, geometry masks).
* test/test_PBC.py:64-76: the demag field of the periodic magnet equals the
  central part of a 5x larger open magnet holding the same pattern.
* test/test_boundaries.py:17-20: edge canting of a 1-D wire with interfacial
  DMI at 90 % of the critical strength (mumax3 Neumann / open / periodic
  values).  The reference relaxes with its minimiser; here the state is
  relaxed by damped time stepping, which has the same fixed point.
"""

import numpy as np
import pytest

import mumaxplus as mp
import mumaxplus.util.shape as shape

pytestmark = pytest.mark.gpu

PBC_RTOL = 1e-3  # test/test_PBC.py:7
PBC_WANT = np.array([0.89947968, 0.23352228, -0.00010287])  # test/test_PBC.py:89,98
C = 5e-9
NX, NY = 128, 64


def _max_relative_error(result, wanted):
    """test/test_PBC.py:9-12."""
    err = np.linalg.norm(np.asarray(result) - wanted, axis=0)
    return np.max(err / np.linalg.norm(wanted, axis=0))


def _rectangle():
    r = shape.Rectangle(NX / 2 * C, NY / 2 * C)
    r.translate(NX / 2 * C, NY / 2 * C, 0)
    return r


def _setup(magnet):
    magnet.msat = 1000e3
    magnet.aex = 10e-12
    magnet.alpha = 1
    magnet.magnetization = (1.0, 0.1, 0.01)


def _periodic(corners=False):
    world = mp.World(cellsize=(C, C, C), mastergrid=mp.Grid((NX, NY, 0)), pbc_repetitions=(2, 2, 0))
    geo = _rectangle().repeat((0, 0, None), (NX * C, NY * C, None))
    if corners:
        geo = geo.translate(NX / 2 * C, NY / 2 * C, 0)
    magnet = mp.Ferromagnet(world, mp.Grid((NX, NY, 1)), geometry=geo)
    _setup(magnet)
    return world, magnet


def test_pbc_demag_matches_large_open_magnet():
    world_p, mag_p = _periodic()
    world = mp.World(cellsize=(C, C, C))
    mag = mp.Ferromagnet(world, mp.Grid((NX * 5, NY * 5, 1)),
                         geometry=_rectangle().repeat((0, 0, None), (NX * C, NY * C, None)))
    _setup(mag)
    hp = np.array(mag_p.demag_field.eval())
    h = np.array(mag.demag_field.eval())
    got = hp[:, :, NY // 4 + 1:3 * NY // 4 + 1, NX // 4 + 1:3 * NX // 4 + 1]
    want = h[:, :, int(NY * 5 / 2 - NY / 4) + 1:int(NY * 5 / 2 + NY / 4) + 1,
             int(NX * 5 / 2 - NX / 4) + 1:int(NX * 5 / 2 + NX / 4) + 1]
    assert _max_relative_error(got, want) < PBC_RTOL


@pytest.mark.parametrize("corners", [False, True])
def test_pbc_relaxation_matches_mumax3(corners):
    world, magnet = _periodic(corners)
    world.timesolver.run(1e-9)
    assert world.timesolver.time == pytest.approx(1e-9, rel=1e-6)
    assert _max_relative_error(magnet.magnetization.average(), PBC_WANT) < PBC_RTOL


# ------------------------------------------------------------ edge canting
CANT_RTOL = 2e-4  # test/test_boundaries.py:28
CANT = {"neumann": 0.44716486657643906, "open": 0.5958252040658198, "periodic": 0.5958256832017506}
NCELL, CS = 1024, 0.05
DMI = 0.9 * 4 / np.pi


def _wire(kind, x_dir):
    if kind == "periodic":
        mg = mp.Grid((0, 1, 0)) if x_dir else mp.Grid((1, 0, 0))
        reps = (0, 1, 0) if x_dir else (1, 0, 0)
        world = mp.World(cellsize=(CS, CS, CS), mastergrid=mg, pbc_repetitions=reps)
    else:
        world = mp.World(cellsize=(CS, CS, CS))
    magnet = mp.Ferromagnet(world, mp.Grid((NCELL, 1, 1) if x_dir else (1, NCELL, 1)))
    magnet.dmi_tensor.set_interfacial_dmi(DMI)
    magnet.enable_demag = False
    magnet.enable_openbc = kind == "open"
    magnet.anisU = (0, 0, 1)
    magnet.aex = 1.0
    magnet.ku1 = 1.0
    magnet.msat = 1.0
    magnet.alpha = 1.0
    # start from the analytic edge profile so damped dynamics converges fast
    n = np.arange(NCELL) * CS
    th = 0.6 * (np.exp(-n) + np.exp(-(n[-1] - n)))
    m = np.zeros((3, 1, 1, NCELL)) if x_dir else np.zeros((3, 1, NCELL, 1))
    comp = 0 if x_dir else 1
    sign = 1.0 if x_dir else -1.0
    m[comp] = (sign * np.sin(th)).reshape(m[comp].shape)
    m[2] = np.cos(th).reshape(m[2].shape)
    # the right edge cants the other way
    half = NCELL // 2
    if x_dir:
        m[comp, 0, 0, half:] *= -1
    else:
        m[comp, 0, half:, 0] *= -1
    magnet.magnetization = m
    return world, magnet


def _analytic_canting():
    """test/test_boundaries.py:52-55."""
    theta0 = np.arcsin(DMI / 2)
    return 2 * np.arctan(np.exp(-CS / 2) * np.tan(theta0 / 2))


def _edge_angle(magnet, x_dir):
    m = np.array(magnet.magnetization.eval())
    return abs(np.arctan2(m[0 if x_dir else 1].ravel()[0], m[2].ravel()[0]))


@pytest.mark.parametrize("kind", ["neumann", "open", "periodic"])
@pytest.mark.parametrize("x_dir", [True, False])
def test_edge_canting_matches_mumax3(kind, x_dir):
    world, magnet = _wire(kind, x_dir)
    ts = world.timesolver
    ts.set_method("DormandPrince")
    # Damped relaxation: the slowest edge mode decays at gamma*2K/Ms*alpha/(1+alpha^2)
    # ~ 1.8e11 /s.  Run in 20 ps chunks until the edge angle is stationary.
    prev = _edge_angle(magnet, x_dir)
    for _ in range(40):
        ts.run(2e-11)
        cur = _edge_angle(magnet, x_dir)
        if abs(cur - prev) < 2e-7:
            break
        prev = cur
    assert abs(cur - prev) < 2e-7, "edge canting did not become stationary"
    # test/test_boundaries.py:88-140: Neumann against mumax3, open and periodic
    # against the analytic canting (and the mumax3 values)
    assert abs(cur - CANT[kind]) / CANT[kind] < CANT_RTOL
    if kind != "neumann":
        want = _analytic_canting()
        assert abs(cur - want) / want < CANT_RTOL


# ---------------------------------------------------------------- frozen spins
# test/test_frozenspins.py:7-53: the standard problem 4 film (128 x 32 x 2
# cells of 500/128 x 125/32 x 1.5 nm, Ms 800e3, A 13e-12, alpha 0.02).
SP4_CS = (500e-9 / 128, 125e-9 / 32, 3e-9 / 2)
SP4_FIELD = (-24.6e-3, 4.3e-3, 0.0)


def _sp4_film():
    world = mp.World(SP4_CS)
    magnet = mp.Ferromagnet(world, mp.Grid((128, 32, 2)))
    magnet.msat = 800e3
    magnet.aex = 13e-12
    magnet.alpha = 0.02
    return world, magnet


def test_frozen_spins_do_not_move():
    """All spins frozen: 1 ns in the switching field leaves <m> unchanged
    (test/test_frozenspins.py:7-23, tolerance 1e-7)."""
    world, magnet = _sp4_film()
    magnet.bias_magnetic_field = SP4_FIELD
    magnet.frozen_spins = 1
    magnet.magnetization = (1, 0.1, 0)
    avg0 = np.array(magnet.magnetization.average())
    world.timesolver.run(1e-9)
    assert np.all(np.abs(np.array(magnet.magnetization.average()) - avg0) < 1e-7)


def test_half_frozen_minimize_and_reversal_golden():
    """Right half frozen: <m> after minimize and after 1 ns of reversal
    equal the values the reference records (test/test_frozenspins.py:26-53,
    tolerance 1e-4)."""
    world, magnet = _sp4_film()
    magnet.magnetization = (1, 0.1, 0)
    cx = magnet.center[0]
    magnet.frozen_spins = lambda x, y, z: 1 if x > cx else 0
    magnet.magnetization = (1, 0.1, 0)
    magnet.minimize()
    got = np.array(magnet.magnetization.average())
    assert np.all(np.abs(got - (0.9809795618057251, 0.11742201447486877, 0.0)) < 1e-4), got
    magnet.bias_magnetic_field = SP4_FIELD
    world.timesolver.run(1e-9)
    got = np.array(magnet.magnetization.average())
    want = (0.14499470591545105, 0.24905619025230408, 0.0021978835575282574)
    assert np.all(np.abs(got - want) < 1e-4), got
