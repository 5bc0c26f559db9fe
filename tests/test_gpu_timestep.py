"""Time stepping on the GPU vs the FP64 oracle and closed forms."""

import numpy as np
import pytest

from helpers import make_pair, random_unit
from oracle import rk as ork

import mumaxplus as mp
from mumaxplus.util.constants import GAMMALL

pytestmark = pytest.mark.gpu

METHODS = list(ork.METHOD_NAMES)


def precession(t, m0, hz, damping):
    """test/test_timesolver.py:19-29 of the reference."""
    mx, my, mz = m0
    theta0 = np.arccos(mz / np.sqrt(mx ** 2 + my ** 2 + mz ** 2))
    phi0 = np.arctan(my / mx)
    freq = GAMMALL * hz / (1 + damping ** 2)
    phi = phi0 + freq * t
    theta = np.pi - 2 * np.arctan(np.exp(damping * freq * t) * np.tan(np.pi / 2 - theta0 / 2))
    return np.array([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)])


@pytest.mark.parametrize("method", METHODS)
def test_single_spin_precession(method):
    world = mp.World(cellsize=(1, 1, 1))
    magnet = mp.Ferromagnet(world, grid=mp.Grid((1, 1, 1)))
    magnet.enable_demag = False
    magnet.magnetization = (1, 0, 0)
    magnet.alpha = 0.2
    world.bias_magnetic_field = (0, 0, 0.11)
    world.timesolver.set_method(method)
    tps = np.linspace(0, 5e-10, 20)
    out = world.timesolver.solve(tps, {"mx": lambda: magnet.magnetization.average()[0],
                                       "my": lambda: magnet.magnetization.average()[1],
                                       "mz": lambda: magnet.magnetization.average()[2]})
    exact = precession(tps, (1, 0, 0), 0.11, 0.2)
    for i, c in enumerate(("mx", "my", "mz")):
        assert np.max(np.abs(np.array(out[c]) - exact[i])) < 1e-4


@pytest.mark.parametrize("method", METHODS)
def test_two_uncoupled_spins(method):
    world = mp.World(cellsize=(1, 1, 1))
    mags = []
    for i in (0, 1):
        m = mp.Ferromagnet(world, grid=mp.Grid((1, 1, 1), origin=(i, 0, 0)))
        m.enable_demag = False
        # the reference couples the two through their (negligible, Ms = 1 A/m)
        # stray fields; coupling between magnets is not built here
        m.enable_as_stray_field_source = False
        m.magnetization = (1, 0, 0)
        m.alpha = 0.2
        mags.append(m)
    world.bias_magnetic_field = (0, 0, 0.11)
    world.timesolver.set_method(method)
    tps = np.linspace(0, 5e-11, 10)
    out = world.timesolver.solve(tps, {f"m{i}": (lambda mm=mm: mm.magnetization.average()) for i, mm in
                                       enumerate(mags)})
    exact = precession(tps, (1, 0, 0), 0.11, 0.2)
    for i in (0, 1):
        got = np.array(out[f"m{i}"]).T
        assert np.max(np.abs(got - exact)) < 1e-4


def test_fixed_timestep_time_bookkeeping():
    world = mp.World(cellsize=(1, 1, 1))
    world.timesolver.adaptive_timestep = False
    world.timesolver.timestep = 2.3e-9
    world.timesolver.steps(103)
    assert abs(world.timesolver.time - 103 * 2.3e-9) <= 1e-5 * 103 * 2.3e-9


@pytest.mark.parametrize("method", ["DormandPrince", "Fehlberg"])
def test_fixed_step_film_matches_oracle(method):
    """128x32x1 film with exchange + demag + Zeeman, fixed dt, vs FP64."""
    rng = np.random.default_rng(2)
    size = (64, 16, 1)
    cs = (500e-9 / 128, 125e-9 / 32, 3e-9)
    phi = rng.uniform(0, 2 * np.pi)
    m0 = np.array([1.0, 0.1, 0.0])[:, None, None, None] + 0.01 * rng.normal(size=(3, 1, 16, 64))
    world, mag, ow, om = make_pair(size, cs, {"msat": 8e5, "aex": 1.3e-11, "alpha": 0.02}, m0,
                                   world_bias=(0.025 * np.cos(phi), 0.025 * np.sin(phi), 0.0))
    for ts in (world.timesolver, ow.solver):
        ts.adaptive_timestep = False
    ow.solver.adaptive = False
    world.timesolver.set_method(method)
    ow.solver.set_method(method)
    world.timesolver.timestep = 1e-13
    ow.solver.dt = 1e-13
    world.timesolver.steps(200)
    ow.solver.steps(200)
    got = np.array(mag.magnetization.average())
    want = om.m.reshape(3, -1).mean(axis=1)
    assert np.abs(got - want).max() < 1e-4
    assert np.abs(mag.magnetization.get() - om.m).max() < 2e-3
    assert abs(world.timesolver.time - ow.solver.time) < 1e-20


def test_adaptive_run_matches_oracle():
    """Adaptive DormandPrince run: same controller, <m> within 1e-4 of FP64."""
    rng = np.random.default_rng(4)
    size = (32, 8, 2)
    cs = (3e-9, 3e-9, 3e-9)
    m0 = random_unit(rng, (2, 8, 32))
    world, mag, ow, om = make_pair(size, cs, {"msat": 8e5, "aex": 1.3e-11, "alpha": 0.3, "ku1": 3e5,
                                              "anisU": (0, 0, 1)}, m0, dmi={"xxz": 1e-3, "yyz": 1e-3})
    world.timesolver.set_method("DormandPrince")
    ow.solver.set_method("DormandPrince")
    world.timesolver.run(5e-11)
    ow.solver.run(5e-11)
    got = np.array(mag.magnetization.average())
    want = om.m.reshape(3, -1).mean(axis=1)
    assert np.abs(got - want).max() < 1e-4
    assert abs(world.timesolver.time - ow.solver.time) < 1e-3 * 5e-11


def test_rejection_path_and_errors():
    world = mp.World(cellsize=(2e-9, 2e-9, 2e-9))
    mag = mp.Ferromagnet(world, mp.Grid((8, 8, 1)))
    mag.aex = 1.3e-11
    mag.msat = 8e5
    ts = world.timesolver
    ts.max_error = 1e-12
    ts.timestep = 1e-12
    ts.adaptive_timestep = True
    ts._impl.step()
    assert ts.nrejected > 0
    ts.adaptive_timestep = False
    ts.timestep = 0.0
    with pytest.raises(RuntimeError):
        ts._impl.step()
    ts.adaptive_timestep = True
    ts.timestep = 1e-13
    ts.lower_bound = 3.0
    with pytest.raises(RuntimeError):
        ts._impl.step()
    with pytest.raises((TypeError, ValueError)):
        ts.set_method("asdf")
    with pytest.raises((TypeError, ValueError)):
        ts.set_method(None)


def test_film128x32_fixed_step_matches_fp64_fixture():
    """128x32x1 film, exchange + demag + 25 mT bias, Dormand-Prince at a fixed
    1e-13 s for 10000 steps (1 ns): <m> every 1000 steps within 1e-4 of the
    FP64 oracle trajectory stored in tests/golden (make_film128x32.py)."""
    import os
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "film128x32_dp_fixed.npz"))
    world = mp.World(cellsize=(500e-9 / 128, 125e-9 / 32, 3e-9))
    mag = mp.Ferromagnet(world, mp.Grid((128, 32, 1)))
    mag.msat = 8e5
    mag.aex = 1.3e-11
    mag.alpha = 0.02
    world.bias_magnetic_field = tuple(float(v) for v in g["bias"])
    mag.magnetization = g["m0"]  # FP32 unit vectors; renormalising moves them by at most an ulp
    ts = world.timesolver
    ts.set_method("DormandPrince")
    ts.adaptive_timestep = False
    ts.timestep = float(g["dt"])
    errs = []
    for want in g["averages"]:
        ts.steps(int(g["check_every"]))
        errs.append(float(np.abs(np.array(mag.magnetization.average()) - want).max()))
    assert max(errs) < 1e-4, errs
    assert abs(ts.time - float(g["time"])) <= 1e-9 * float(g["time"])
