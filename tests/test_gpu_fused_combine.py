"""The Runge-Kutta stage state formed inside the demag's first FFT pass.

When a world's only variable is one independent magnetisation with demag,
each stage's combination x = normalise(x0 + sum_j c_j k_j) runs inside the
x-forward FFT pass (world option "fuse_combine", off by default, with the
combination inside the field launch switched off), which reads
the step start and the derivatives once for both the new state and its
transform.  The arithmetic is the separate combination's; the compiler may
fuse its multiply-adds differently in the two kernels, so the runs agree to
FP32 rounding, not bit for bit.  Cases: every tableau, fixed and adaptive
steps with rejections, a geometry mask with per-cell Ms, a single layer,
and a relaxation.  The separate combination is compared with the FP64 oracle
by the time-stepping tests, which carries that parity over to the fused one.
"""

import numpy as np
import pytest

from helpers import random_unit

import mumaxplus as mp

pytestmark = pytest.mark.gpu


def _film(fuse, size=(24, 16, 3), seed=3, geometry=True, per_cell=True):
    world = mp.World((2e-9, 2.5e-9, 3e-9))
    nx, ny, nz = size
    geom = None
    if geometry:
        geom = np.random.default_rng(seed + 100).random((nz, ny, nx)) > 0.15
    mag = mp.Ferromagnet(world, mp.Grid(size), geometry=geom)
    mag.aex, mag.alpha, mag.ku1, mag.anisU = 1.3e-11, 0.05, 3e5, (0.1, 0.2, 1.0)
    if per_cell:
        mag.msat = np.random.default_rng(seed + 7).uniform(6e5, 9e5, size=(nz, ny, nx))
    else:
        mag.msat = 8e5
    mag.dmi_tensor.set_interfacial_dmi(2e-3)
    world.bias_magnetic_field = (0.01, 0.0, 0.02)
    mag.magnetization = random_unit(np.random.default_rng(seed), (nz, ny, nx)).astype(np.float32)
    # the combination inside the field launch (on by default) takes
    # precedence; this test compares the demag-pass placement with the
    # separate launch
    world._impl.set_option("fuse_stages", 0)
    world._impl.set_option("fuse_combine", int(fuse))
    assert world._impl.get_option("fuse_combine") == int(fuse)
    return world, mag


@pytest.mark.parametrize("method", ["Heun", "BogackiShampine", "CashKarp", "Fehlberg", "DormandPrince"])
def test_fixed_steps_match_separate_combination(method):
    out = []
    for fuse in (True, False):
        world, mag = _film(fuse)
        ts = world.timesolver
        ts.set_method(method)
        ts.adaptive_timestep = False
        ts.timestep = 4e-14
        ts.steps(12)
        ts.run(1.3e-13)
        out.append((mag.magnetization.get(), ts.time))
    (a, ta), (b, tb) = out
    assert np.abs(a - b).max() < 2e-6
    assert ta == tb


@pytest.mark.parametrize("size,geometry,per_cell", [((24, 16, 3), True, True), ((40, 20, 1), False, False)])
def test_adaptive_run_matches_separate_combination(size, geometry, per_cell):
    out = []
    for fuse in (True, False):
        world, mag = _film(fuse, size=size, seed=5, geometry=geometry, per_cell=per_cell)
        ts = world.timesolver
        ts.set_method("DormandPrince")
        ts.max_error = 1e-6
        ts.timestep = 2e-12  # far too large: the first attempts are rejected
        ts._impl.run(5e-12)
        out.append((mag.magnetization.get(), ts.time, ts._impl.nsteps, ts._impl.nrejected))
    a, b = out
    # rounding differences feed the step controller; both runs stay within
    # the controller's error per step (1e-6) times the number of steps
    assert a[3] > 0 and abs(a[2] - b[2]) <= 1
    assert np.abs(a[0] - b[0]).max() < 1e-6 * a[2]
    assert a[1] == b[1]


def test_relax_matches_separate_combination():
    out = []
    for fuse in (True, False):
        world, mag = _film(fuse, seed=8)
        mag.alpha = 0.5
        mag.relax()
        out.append(mag.magnetization.get())
    assert np.abs(out[0] - out[1]).max() < 1e-4


def test_fused_pass_replaces_the_combination_launch():
    world, mag = _film(True, geometry=False, per_cell=False)
    ts = world.timesolver
    ts.set_method("DormandPrince")
    ts.adaptive_timestep = False
    ts.timestep = 4e-14
    ts.steps(2)
    impl = world._impl
    impl.set_profiling(True)
    impl.profile_reset()
    ts.steps(3)
    rep = impl.profile_report()
    impl.set_profiling(False)
    assert rep["k_xfwd_combine"][1] == 18  # six stage states per step
    assert "k_combine" not in rep and "k_xfwd" not in rep
