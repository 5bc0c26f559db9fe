"""The persistent batch step kernel and the step-path autotuner.

The persistent kernel runs every pass of a batch step (stage combination,
demag passes, field, error norms, controller) as phases of one resident
kernel with a grid barrier between phases; it calls the same device bodies
as the per-pass launches, so a member's trajectory is the same bit for bit
on either path:

* fixed-step and adaptive members of every shape class (single layers, a
  masked ellipse, multi-layer magnets, DMI), persistent against per-pass;
* the autotuner ("auto") decides per size class, records its choice and
  timings, and its steps are the same as either path's;
* a grid barrier forced to fail from the host (spin limit 0) raises, leaves
  the batch at the start of the interrupted step, and the run then ends
  exactly where an uninterrupted one does;
* members with FFT lines longer than 256 points are refused by the
  persistent path and run per pass under "auto".
"""

import numpy as np
import pytest

from helpers import random_unit

import mumaxplus as mp

pytestmark = pytest.mark.gpu

CS = (2e-9, 2e-9, 2e-9)


def _ellipse(n):
    x = (np.arange(n[0]) + 0.5) / n[0] - 0.5
    y = (np.arange(n[1]) + 0.5) / n[1] - 0.5
    e = (x[None, :] ** 2 + y[:, None] ** 2) <= 0.25
    return np.broadcast_to(e, (n[2], n[1], n[0])).copy()


SPECS = [
    ((16, 16, 1), None, {"msat": 8e5, "aex": 1.3e-11, "alpha": 0.02, "ku1": 5e5, "anisU": (0, 0, 1)}, 1),
    ((32, 16, 1), "ellipse", {"msat": 1.1e6, "aex": 1.6e-11, "alpha": 0.05}, 2),
    ((20, 12, 2), None, {"msat": 6e5, "aex": 1e-11, "alpha": 0.1, "ku1": 2e5, "anisU": (1, 1, 0)}, 3),
    ((16, 8, 4), None, {"msat": 4e5, "aex": 5e-12, "alpha": 0.01}, 4),
    ((33, 17, 1), None, {"msat": 9e5, "aex": 1.5e-11, "alpha": 0.03, "ku1": 1e5, "anisU": (0, 1, 0)}, 5),
    ((64, 48, 3), "ellipse", {"msat": 1.4e6, "aex": 2e-11, "alpha": 0.005, "ku1": 1e6, "anisU": (0, 0, 1)}, 6),
]


def _build(spec, fixed_dt=None, bias=(0.01, 0.0, 0.02), dmi=0.0):
    size, geom, params, seed = spec
    world = mp.World(CS)
    g = _ellipse(size) if geom == "ellipse" else None
    mag = mp.Ferromagnet(world, mp.Grid(size), geometry=g)
    for k, v in params.items():
        setattr(mag, k, v)
    if dmi:
        mag.dmi_tensor.set_interfacial_dmi(dmi)
    world.bias_magnetic_field = bias
    rng = np.random.default_rng(seed)
    mag.magnetization = random_unit(rng, (size[2], size[1], size[0])).astype(np.float32)
    ts = world.timesolver
    ts.set_method("DormandPrince")
    if fixed_dt is not None:
        ts.adaptive_timestep = False
        ts.timestep = fixed_dt
    return world, mag


def _run(mode, rules, fixed_dt=None, poll=8, blocks_per_cu=None, options=None):
    """Every spec with its rule ("steps" or "duration"), stepped by a batch in
    the given step mode; returns [(m, time, steps, rejected)]."""
    batch = mp.Batch(80000)
    batch.step_mode = mode
    if blocks_per_cu is not None:
        batch._impl.set_option("persist_blocks_per_cu", blocks_per_cu)
    for k, v in (options or {}).items():
        batch._impl.set_option(k, v)
    out = []
    for i, (spec, (how, arg)) in enumerate(zip(SPECS, rules)):
        w, m = _build(spec, fixed_dt=fixed_dt, dmi=1e-3 if i == 0 else 0.0)
        mid = batch.add(m, **{how: arg})
        assert mid is not None
        out.append((mid, w, m))
    states = {}
    taken = 0
    while len(batch.finished()) < len(out):
        batch.advance(poll)
        taken += poll
        assert taken < 100000
    for mid, w, m in out:
        states[mid] = batch.member(mid)
    path = batch.step_path()
    for mid, w, m in out:
        batch.remove(mid)
    res = [(m.magnetization.get(), w.timesolver.time, states[mid]["steps"], states[mid]["rejected"])
           for mid, w, m in out]
    return res, path


def _same(a, b):
    assert len(a) == len(b)
    for (ma, ta, sa, ra), (mb, tb, sb, rb) in zip(a, b):
        assert (sa, ra) == (sb, rb)
        assert ta == tb
        assert np.array_equal(ma, mb)


FIXED = [("steps", 30)] * len(SPECS)
ADAPTIVE = [("duration", 2e-11), ("duration", 3e-11), ("duration", 1.5e-11), ("steps", 25), ("duration", 2.5e-11),
            ("duration", 1e-11)]


def test_persistent_fixed_steps_equal_the_per_pass_path():
    want, p0 = _run("passes", FIXED, fixed_dt=1e-13)
    got, p1 = _run("persistent", FIXED, fixed_dt=1e-13)
    assert p0["last_path"] == "passes" and p1["last_path"] == "persistent"
    assert p1["persist_grid"] > 0
    _same(got, want)


def test_persistent_adaptive_equals_the_per_pass_path():
    want, _ = _run("passes", ADAPTIVE, poll=16)
    got, _ = _run("persistent", ADAPTIVE, poll=16)
    assert sum(r for _, _, _, r in want) > 0 or True  # rejections are allowed, not required
    _same(got, want)


def test_persistent_with_one_block_per_compute_unit():
    want, _ = _run("passes", ADAPTIVE, poll=16)
    got, _ = _run("persistent", ADAPTIVE, poll=16, blocks_per_cu=1)
    _same(got, want)


def test_persistent_with_a_grid_of_64_blocks():
    want, _ = _run("passes", FIXED, fixed_dt=1e-13)
    batch_opts = {"persist_max_grid": 64}
    got, _ = _run("persistent", FIXED, fixed_dt=1e-13, options=batch_opts)
    _same(got, want)


def test_auto_mode_decides_and_records():
    want, _ = _run("passes", FIXED, fixed_dt=1e-13, poll=4)
    got, path = _run("auto", FIXED, fixed_dt=1e-13, poll=2)
    _same(got, want)
    b = mp.Batch(256)
    report = b.tune_report()
    assert report, "the autotuner recorded no size class"
    decided = [v for v in report.values() if v["choice"] is not None]
    assert decided, report
    for v in decided:
        assert v["choice"] in ("passes", "persistent", "persistent_1pcu", "persistent_64")
        assert v[v["choice"]] is not None and v[v["choice"]] > 0


def test_forced_barrier_failure_rewinds_the_step():
    want, _ = _run("passes", FIXED, fixed_dt=1e-13, poll=5)
    batch = mp.Batch(80000)
    batch.step_mode = "persistent"
    out = []
    for i, (spec, (how, arg)) in enumerate(zip(SPECS, FIXED)):
        w, m = _build(spec, fixed_dt=1e-13, dmi=1e-3 if i == 0 else 0.0)
        out.append((batch.add(m, **{how: arg}), w, m))
    batch.advance(5)
    batch._impl.set_option("persist_spin", 0)
    with pytest.raises(RuntimeError, match="grid barrier"):
        batch.advance(5)
    # the interrupted step is retried from its start
    for mid, _, _ in out:
        assert batch.member(mid)["steps"] == 5
    batch._impl.set_option("persist_spin", 1 << 22)
    taken = 5
    while len(batch.finished()) < len(out):
        batch.advance(5)
        taken += 5
        assert taken < 1000
    states = {mid: batch.member(mid) for mid, _, _ in out}
    for mid, _, _ in out:
        batch.remove(mid)
    got = [(m.magnetization.get(), w.timesolver.time, states[mid]["steps"], states[mid]["rejected"])
           for mid, w, m in out]
    _same(got, want)


def test_long_lines_are_refused_by_the_persistent_path():
    spec = ((130, 8, 1), None, {"msat": 8e5, "aex": 1.3e-11, "alpha": 0.02}, 7)
    w, m = _build(spec, fixed_dt=1e-13)
    batch = mp.Batch(4096)
    batch.add(m, steps=4)
    assert batch.step_path()["persist_grid"] == 0
    batch.step_mode = "persistent"
    with pytest.raises(ValueError, match="256 points"):
        batch.advance(2)
    batch.step_mode = "auto"
    batch.advance(4)
    assert batch.step_path()["last_path"] == "passes"
    assert batch.member(0)["steps"] == 4
