"""Snapshots, checkpoint files and batch preemption on the GPU.

* restoring a snapshot and stepping again gives the state the first run
  reached (max |dm| <= 1e-5: the restored magnetisation is normalised again
  and the first step computes its derivative afresh);
* a checkpoint file holds every magnetisation, sublattice, elastic state and
  the solver settings, read back with allow_pickle=False;
* a batch member preempted mid-run and resumed ends where an uninterrupted
  run ends (same time; <m> within 1e-6);
* with thermal noise, a fresh world restored from a checkpoint continues the
  first world's run bit for bit (the noise stream's counters are part of it);
* a run continued from a checkpoint file in a fresh world ends within 1e-4 in
  <m> of the FP64 oracle's uninterrupted run.
"""

import numpy as np
import pytest

from helpers import random_unit

import mumaxplus as mp

pytestmark = pytest.mark.gpu


def _film(seed=1):
    world = mp.World((2e-9, 2e-9, 2e-9))
    mag = mp.Ferromagnet(world, mp.Grid((24, 12, 2)), name="film")
    mag.msat, mag.aex, mag.alpha, mag.ku1, mag.anisU = 8e5, 1.3e-11, 0.05, 2e5, (0, 0, 1)
    mag.magnetization = random_unit(np.random.default_rng(seed), (2, 12, 24)).astype(np.float32)
    world.timesolver.set_method("DormandPrince")
    return world, mag


def test_snapshot_restore_and_continue():
    world, mag = _film()
    world.timesolver.run(5e-12)
    snap = world.snapshot()
    world.timesolver.run(5e-12)
    m1, t1 = mag.magnetization.get(), world.timesolver.time
    world.restore(snap)
    assert world.timesolver.time == snap["solver"]["time"]
    world.timesolver.run(5e-12)
    assert world.timesolver.time == pytest.approx(t1, rel=1e-12)
    assert np.abs(mag.magnetization.get() - m1).max() <= 1e-5


def test_checkpoint_file_round_trip(tmp_path):
    world = mp.World((2e-9, 2e-9, 2e-9))
    fm = mp.Ferromagnet(world, mp.Grid((8, 6, 1)), name="fm")
    afm = mp.Antiferromagnet(world, mp.Grid((6, 6, 1), (20, 0, 0)), name="afm")
    for m in (fm, afm):
        m.enable_as_stray_field_source = False
        m.enable_as_stray_field_destination = False
    fm.enable_elastodynamics = True
    rng = np.random.default_rng(3)
    fm.magnetization = random_unit(rng, (1, 6, 8)).astype(np.float32)
    afm.sub1.magnetization = random_unit(rng, (1, 6, 6)).astype(np.float32)
    afm.sub2.magnetization = random_unit(rng, (1, 6, 6)).astype(np.float32)
    fm.elastic_displacement = rng.normal(size=(3, 1, 6, 8)).astype(np.float32) * 1e-12
    ts = world.timesolver
    ts.set_method("BogackiShampine")
    ts.adaptive_timestep = False
    ts.timestep = 3e-14
    ts.time = 2e-12
    path = tmp_path / "state.npz"
    world.save_checkpoint(str(path))
    want = {"m": fm.magnetization.get(), "u": fm.elastic_displacement.get(), "s1": afm.sub1.magnetization.get(),
            "s2": afm.sub2.magnetization.get()}
    fm.magnetization = (1, 0, 0)
    afm.sub1.magnetization = (0, 1, 0)
    fm.elastic_displacement = np.zeros((3, 1, 6, 8), dtype=np.float32)
    ts.set_method("Heun")
    ts.adaptive_timestep = True
    ts.time = 0.0
    world.load_checkpoint(str(path))
    assert np.array_equal(fm.elastic_displacement.get(), want["u"])
    for key, got in (("m", fm.magnetization.get()), ("s1", afm.sub1.magnetization.get()),
                     ("s2", afm.sub2.magnetization.get())):
        assert np.abs(got - want[key]).max() <= 1e-6
    assert ts._impl.method == "BogackiShampine" and not ts.adaptive_timestep
    assert ts.timestep == pytest.approx(3e-14) and ts.time == pytest.approx(2e-12)
    with np.load(str(path), allow_pickle=False) as f:
        assert "ferromagnet/fm/magnetization" in f.files and "antiferromagnet/afm/sub2" in f.files


def test_preempted_member_resumes_where_it_would_have_been():
    w1, m1 = _film(5)
    b = mp.Batch(4096)
    b.add(m1, duration=4e-11)
    b.run_until_done()
    w2, m2 = _film(5)
    b2 = mp.Batch(4096)
    mid = b2.add(m2, duration=4e-11)
    b2.advance(40)
    assert mid not in b2.finished()
    token = b2.preempt(mid)
    assert b2.info()["members"] == 0 and 0 < w2.timesolver.time < 4e-11
    b3 = mp.Batch(4096)
    b3.resume(token)
    b3.run_until_done()
    assert w2.timesolver.time == pytest.approx(w1.timesolver.time, rel=1e-12)
    assert np.abs(np.array(m2.magnetization.average()) - np.array(m1.magnetization.average())).max() <= 1e-6


def _noisy_film():
    world = mp.World((3e-9, 3e-9, 3e-9))
    mag = mp.Ferromagnet(world, mp.Grid((16, 8, 2)), name="film")
    mag.msat, mag.aex, mag.alpha, mag.ku1, mag.anisU = 8e5, 1.3e-11, 0.1, 2e5, (0, 0, 1)
    mag.temperature = 300.0
    mag.magnetization = random_unit(np.random.default_rng(9), (2, 8, 16)).astype(np.float32)
    ts = world.timesolver
    ts.set_method("Heun")
    ts.adaptive_timestep = False
    ts.timestep = 1e-14
    return world, mag


def test_noisy_restore_into_fresh_world_continues_the_run(tmp_path):
    """With temperature > 0 the snapshot carries the noise stream's draw
    counters: a fresh world restored from it draws the noise the first world
    draws next, so both end bit for bit alike.  (The first world is restored
    from the same snapshot too, so that both start from the same
    re-normalised magnetisation.)"""
    wa, ma = _noisy_film()
    wa.timesolver.steps(10)
    assert ma._impl.noise_state[1] == 10
    path = tmp_path / "noisy.npz"
    wa.save_checkpoint(str(path))
    at_snapshot = ma._impl.noise_state
    wa.load_checkpoint(str(path))
    wa.timesolver.steps(10)
    want = ma.magnetization.get()

    wb, mb = _noisy_film()
    assert mb._impl.noise_state[1] == 0
    wb.load_checkpoint(str(path))
    assert mb._impl.noise_state == at_snapshot
    wb.timesolver.steps(10)
    assert wb.timesolver.time == wa.timesolver.time
    assert np.array_equal(mb.magnetization.get(), want)

    # without the counters the fresh world would draw the first ten steps' noise again
    wc, mc = _noisy_film()
    wc.load_checkpoint(str(path))
    seed, _, q = mc._impl.noise_state
    mc._impl.noise_state = (seed, 0, q)
    wc.timesolver.steps(10)
    assert np.abs(mc.magnetization.get() - want).max() > 1e-4


def test_checkpoint_continue_matches_uninterrupted_oracle(tmp_path):
    """20 fixed Dormand-Prince steps, a checkpoint file, a fresh world loaded
    from it and 20 more steps, against 40 uninterrupted FP64 oracle steps:
    <m> within 1e-4 and every cell within 2e-4."""
    from oracle import model as omodel

    size, cs = (16, 8, 2), (3e-9, 3e-9, 3e-9)
    params = {"msat": 8e5, "aex": 1.3e-11, "alpha": 0.05, "ku1": 3e5, "anisU": (0.0, 0.0, 1.0)}

    def product():
        world = mp.World(cs)
        mag = mp.Ferromagnet(world, mp.Grid(size), name="film")
        for k, v in params.items():
            setattr(mag, k, v)
        mag.dmi_tensor.set_interfacial_dmi(1e-3)
        mag.magnetization = random_unit(np.random.default_rng(11), (2, 8, 16)).astype(np.float32)
        ts = world.timesolver
        ts.set_method("DormandPrince")
        ts.adaptive_timestep = False
        ts.timestep = 5e-14
        return world, mag

    w1, m1 = product()
    ow = omodel.World(cs)
    om = ow.add_ferromagnet(size)
    for k, v in params.items():
        om.set(k, v)
    om.set("xxz", 1e-3)
    om.set("yyz", 1e-3)
    om.set_m(m1.magnetization.get().astype(np.float64))
    ow.solver.set_method("DormandPrince")
    ow.solver.adaptive = False
    ow.solver.dt = 5e-14
    ow.solver.steps(40)

    w1.timesolver.steps(20)
    path = tmp_path / "film.npz"
    w1.save_checkpoint(str(path))
    del w1, m1
    w2, m2 = product()
    m2.magnetization = (1, 0, 0)
    w2.load_checkpoint(str(path))
    assert w2.timesolver.time == pytest.approx(20 * 5e-14, rel=1e-12)
    w2.timesolver.steps(20)
    got = m2.magnetization.get()
    assert np.abs(got.reshape(3, -1).mean(axis=1) - om.m.reshape(3, -1).mean(axis=1)).max() < 1e-4
    assert np.abs(got - om.m).max() < 2e-4
    assert w2.timesolver.time == pytest.approx(ow.solver.time, rel=1e-12)
