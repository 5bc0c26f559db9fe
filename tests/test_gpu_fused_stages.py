"""Stage combinations fused into the field launch, on the GPU.

A world whose only variable is one magnetisation on the uniform-parameter
LLG path, stepped with a first-same-as-last tableau, computes each stage
state inside the field launch that produces the last derivative it needs
(and, with a fixed step, the next step's first stage state inside the step's
last evaluation).  The arithmetic is that of the separate combination, so
the fused run must equal the unfused one ("fuse_stages" off) bit for bit:
fixed and adaptive steps, rejections, a step change and a state change
between calls.  The fused path is the default, so the time-stepping tests
against the FP64 oracle cover it directly; equality with the separate
combination carries the same parity over to the other path.
"""

import numpy as np
import pytest

from helpers import random_unit

import mumaxplus as mp

pytestmark = pytest.mark.gpu


def _film(fuse, seed=3, geometry=True):
    world = mp.World((2e-9, 2.5e-9, 3e-9))
    nx, ny, nz = 24, 16, 3
    geom = None
    if geometry:
        geom = np.random.default_rng(seed + 100).random((nz, ny, nx)) > 0.15
    mag = mp.Ferromagnet(world, mp.Grid((nx, ny, nz)), geometry=geom)
    mag.msat, mag.aex, mag.alpha, mag.ku1, mag.anisU = 8e5, 1.3e-11, 0.05, 3e5, (0.1, 0.2, 1.0)
    mag.dmi_tensor.set_interfacial_dmi(2e-3)
    world.bias_magnetic_field = (0.01, 0.0, 0.02)
    mag.magnetization = random_unit(np.random.default_rng(seed), (nz, ny, nx)).astype(np.float32)
    world._impl.set_option("fuse_stages", int(fuse))
    return world, mag


def _state(world, mag):
    ts = world.timesolver
    return mag.magnetization.get(), ts.time, ts.timestep


@pytest.mark.parametrize("method", ["DormandPrince", "BogackiShampine"])
def test_fixed_steps_equal_unfused(method):
    out = []
    for fuse in (True, False):
        world, mag = _film(fuse)
        assert world._impl.get_option("fuse_stages") == int(fuse)
        ts = world.timesolver
        ts.set_method(method)
        ts.adaptive_timestep = False
        ts.timestep = 4e-14
        ts.steps(15)
        ts.timestep = 3e-14  # a new step: the precomputed first stage is stale
        ts.steps(5)
        mag.magnetization = mag.magnetization.get() * 1.0  # a state change between calls
        ts.steps(5)
        ts.run(2.3e-13)  # ends with a shortened step
        ts.steps(3)
        out.append(_state(world, mag))
    (m1, t1, dt1), (m2, t2, dt2) = out
    assert np.array_equal(m1, m2)
    assert t1 == t2 and dt1 == dt2


def test_adaptive_run_equals_unfused():
    out = []
    for fuse in (True, False):
        world, mag = _film(fuse, seed=5)
        ts = world.timesolver
        ts.set_method("DormandPrince")
        ts.max_error = 1e-6
        ts.timestep = 2e-12  # far too large: the first attempts are rejected
        ts._impl.run(5e-12)  # the binding's run keeps the step (no sensible-step reset)
        out.append(_state(world, mag) + (ts._impl.nsteps, ts._impl.nrejected))
    a, b = out
    assert np.array_equal(a[0], b[0])
    assert a[1:] == b[1:]
    assert a[4] > 0  # rejections happened


def test_fused_launches_replace_the_combination():
    """With fusion a fixed-step Dormand-Prince step is six field launches and
    no combination launch once the first step has run."""
    world, mag = _film(True, geometry=False)
    ts = world.timesolver
    ts.set_method("DormandPrince")
    ts.adaptive_timestep = False
    ts.timestep = 4e-14
    ts.steps(2)
    impl = world._impl
    impl.set_profiling(True)
    impl.profile_reset()
    ts.steps(4)
    rep = impl.profile_report()
    impl.set_profiling(False)
    assert rep["k_field_next"][1] == 24
    assert "k_combine" not in rep and "k_field" not in rep
