"""Batched stepping of independent ferromagnets (mumaxplus.Batch) on the GPU.

* every member ends where the same magnet stepped alone through its World's
  TimeSolver ends: fixed-step DormandPrince (max |dm| <= 1e-5), adaptive
  run(duration) with the TimeSolver.run rules (<m> within 1e-5, same time);
* a member's trajectory is the same bit for bit alone, in a mix, admitted
  while others are mid-run, with and without the captured graph;
* damping-only relaxation stops below the torque threshold;
* accuracy against the FP64 oracle: <m> within 1e-4 (the path's bar);
* admission by free capacity, and refusals.
"""

import numpy as np
import pytest

from helpers import make_pair, random_unit

import mumaxplus as mp

pytestmark = pytest.mark.gpu

CS = (2e-9, 2e-9, 2e-9)


def _ellipse(n):
    x = (np.arange(n[0]) + 0.5) / n[0] - 0.5
    y = (np.arange(n[1]) + 0.5) / n[1] - 0.5
    e = (x[None, :] ** 2 + y[:, None] ** 2) <= 0.25
    return np.broadcast_to(e, (n[2], n[1], n[0])).copy()


# size, geometry, material, seed
SPECS = [
    ((16, 16, 1), None, {"msat": 8e5, "aex": 1.3e-11, "alpha": 0.02, "ku1": 5e5, "anisU": (0, 0, 1)}, 1),
    ((32, 16, 1), "ellipse", {"msat": 1.1e6, "aex": 1.6e-11, "alpha": 0.05}, 2),
    ((20, 12, 2), None, {"msat": 6e5, "aex": 1e-11, "alpha": 0.1, "ku1": 2e5, "anisU": (1, 1, 0)}, 3),
    ((16, 8, 4), None, {"msat": 4e5, "aex": 5e-12, "alpha": 0.01}, 4),
    ((33, 17, 1), None, {"msat": 9e5, "aex": 1.5e-11, "alpha": 0.03, "ku1": 1e5, "anisU": (0, 1, 0)}, 5),
]


def _build(spec, fixed_dt=None, bias=(0.0, 0.0, 0.0), dmi=0.0):
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


def test_fixed_steps_match_the_plain_path():
    batch = mp.Batch(20000)
    pairs = []
    for i, spec in enumerate(SPECS):
        bw, bm = _build(spec, fixed_dt=1e-13, dmi=1e-3 if i == 0 else 0.0)
        pw, pm = _build(spec, fixed_dt=1e-13, dmi=1e-3 if i == 0 else 0.0)
        assert batch.add(bm, steps=40) is not None
        pairs.append((bw, bm, pw, pm))
    batch.run_until_done()
    for bw, bm, pw, pm in pairs:
        pw.timesolver.steps(40)
        assert np.abs(bm.magnetization.get() - pm.magnetization.get()).max() <= 1e-5
        assert bw.timesolver.time == pytest.approx(pw.timesolver.time, rel=1e-12)
        assert bw.timesolver.timestep == 1e-13


def test_adaptive_run_matches_the_plain_path():
    batch = mp.Batch(20000)
    pairs = []
    for spec in SPECS[:3]:
        bw, bm = _build(spec, bias=(0.0, 0.0, 0.05))
        pw, pm = _build(spec, bias=(0.0, 0.0, 0.05))
        batch.add(bm, duration=3e-11)
        pairs.append((bw, bm, pw, pm))
    batch.run_until_done(poll=16)
    for bw, bm, pw, pm in pairs:
        pw.timesolver.run(3e-11)
        assert bw.timesolver.time == pytest.approx(pw.timesolver.time, rel=1e-9)
        assert np.abs(np.array(bm.magnetization.average()) - np.array(pm.magnetization.average())).max() <= 1e-5
        assert np.abs(bm.magnetization.get() - pm.magnetization.get()).max() <= 1e-4


def _alone(spec, how, arg, graph=True):
    w, m = _build(spec, bias=(0.01, 0.0, 0.02))
    b = mp.Batch(10000)
    b.use_graph = graph
    b.add(m, **{how: arg})
    b.run_until_done(poll=8)
    return m.magnetization.get(), w.timesolver.time


def test_member_is_bitwise_the_same_alone_and_in_a_mix():
    spec = SPECS[1]
    want, want_t = _alone(spec, "duration", 2e-11)
    # in a mix, admitted while the others are mid-run
    batch = mp.Batch(30000)
    others = []
    for s in (SPECS[0], SPECS[2], SPECS[3]):
        w, m = _build(s, bias=(0.0, 0.01, 0.0))
        batch.add(m, duration=5e-11)
        others.append((w, m))
    batch.advance(7)
    w, m = _build(spec, bias=(0.01, 0.0, 0.02))
    mid = batch.add(m, duration=2e-11)
    while mid not in batch.finished():
        batch.advance(5)
    batch.remove(mid)
    assert np.array_equal(m.magnetization.get(), want)
    assert w.timesolver.time == want_t
    # the others are unaffected by the admission: same as alone
    batch.run_until_done()
    for s, (w2, m2) in zip((SPECS[0], SPECS[2], SPECS[3]), others):
        w3, m3 = _build(s, bias=(0.0, 0.01, 0.0))
        b3 = mp.Batch(10000)
        b3.add(m3, duration=5e-11)
        b3.run_until_done()
        assert np.array_equal(m2.magnetization.get(), m3.magnetization.get())


def test_graph_and_eager_steps_agree():
    a, ta = _alone(SPECS[2], "steps", 25, graph=True)
    b, tb = _alone(SPECS[2], "steps", 25, graph=False)
    assert np.array_equal(a, b) and ta == tb


def test_relax_stops_below_threshold():
    spec = ((24, 16, 1), "ellipse", {"msat": 8e5, "aex": 1.3e-11, "alpha": 0.3, "ku1": 1e5, "anisU": (1, 0, 0)}, 7)
    w, m = _build(spec)
    e0 = m.total_energy.eval()
    b = mp.Batch(5000)
    mid = b.add(m, relax_torque=1e9)
    b.run_until_done(poll=16)
    tq = np.linalg.norm(m.damping_torque.eval(), axis=0).max()
    assert tq <= 1e9 * (1 + 1e-5)
    assert m.total_energy.eval() < e0
    assert w.timesolver.time > 0 and mid is not None


def test_fixed_steps_against_the_fp64_oracle():
    rng = np.random.default_rng(9)
    size = (32, 16, 1)
    m0 = np.array([1.0, 0.2, 0.1])[:, None, None, None] + 0.05 * rng.normal(size=(3, 1, 16, 32))
    world, mag, ow, om = make_pair(size, CS, {"msat": 8e5, "aex": 1.3e-11, "alpha": 0.02, "ku1": 3e5,
                                              "anisU": (0.0, 0.0, 1.0)}, m0, world_bias=(0.0, 0.02, 0.0))
    world.timesolver.set_method("DormandPrince")
    world.timesolver.adaptive_timestep = False
    world.timesolver.timestep = 1e-13
    ow.solver.adaptive = False
    ow.solver.set_method("DormandPrince")
    ow.solver.dt = 1e-13
    b = mp.Batch(1024)
    b.add(mag, steps=150)
    b.run_until_done()
    ow.solver.steps(150)
    got = np.array(mag.magnetization.average())
    want = om.m.reshape(3, -1).mean(axis=1)
    assert np.abs(got - want).max() < 1e-4
    assert abs(world.timesolver.time - ow.solver.time) < 1e-20


def test_admission_by_capacity_and_refusals():
    b = mp.Batch(600)  # 768 cells after rounding up to 256
    w1, m1 = _build(SPECS[0], fixed_dt=1e-13)  # 256 cells
    w2, m2 = _build(SPECS[1], fixed_dt=1e-13)  # 512 cells
    w3, m3 = _build(SPECS[0], fixed_dt=1e-13)
    i1 = b.add(m1, steps=3)
    i2 = b.add(m2, steps=3)
    assert i1 is not None and i2 is not None
    assert b.add(m3, steps=3) is None
    assert b.info()["free_cells"] == 0
    b.advance(3)
    assert sorted(b.finished()) == sorted([i1, i2])
    b.remove(i1)
    assert b.add(m3, steps=3) is not None
    b.run_until_done()
    assert w1.timesolver.time == pytest.approx(3e-13) and w3.timesolver.time == pytest.approx(3e-13)
    with pytest.raises(ValueError):
        mp.Batch(100, method="Fehlberg")
    w4, m4 = _build(SPECS[0], fixed_dt=1e-13)
    m4.alpha = np.full((1, 16, 16), 0.1)
    with pytest.raises(ValueError):
        b.add(m4, steps=1)
    w5 = mp.World(CS)
    a = mp.Ferromagnet(w5, mp.Grid((8, 8, 1)))
    mp.Ferromagnet(w5, mp.Grid((8, 8, 1), (20, 0, 0)))
    with pytest.raises(ValueError):
        b.add(a, steps=1)
    with pytest.raises(ValueError):
        b.add(m4, steps=1, duration=1e-12)
