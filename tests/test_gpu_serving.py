"""The request server (mumaxplus.serving) on the GPU.

* a served switching run ends where the plain path (World.timesolver.run)
  ends: <m> within 1e-5, same time;
* a request's result is the same bit for bit served alone and served in a
  mix where it is admitted while others are running;
* a served switching run against the FP64 oracle: <m> within 1e-4;
* a served relaxation stops below its torque threshold, at the state the
  plain relax() reaches (<m> within 1e-2: different step rules), and within
  1e-4 in <m> of the FP64 oracle stepping the same damping-only dynamics to
  the same threshold;
* admission by footprint: the running requests never exceed the capacity.
"""

import numpy as np
import pytest

from helpers import make_pair

import mumaxplus as mp
from mumaxplus import serving

pytestmark = pytest.mark.gpu


def _switch(rid, size=(20, 10, 1), duration=2e-11, b=0.1, alpha=0.05, arrival=0.0):
    nx, ny, nz = size
    x = (np.arange(nx) + 0.5) / nx * 2 - 1
    y = (np.arange(ny) + 0.5) / ny * 2 - 1
    geom = np.broadcast_to((x[None, :] ** 2 + y[:, None] ** 2) <= 1.0, (nz, ny, nx)).copy()
    v = np.array([1.0, 0.05, 0.02])
    m0 = np.empty((3, nz, ny, nx), dtype=np.float32)
    m0[:] = (v / np.linalg.norm(v))[:, None, None, None]
    params = {"msat": 8e5, "aex": 1.3e-11, "ku1": 2e5, "anisU": (1.0, 0.0, 0.0), "alpha": alpha}
    return serving.Request("switch", size, (2e-9,) * 3, params, m0, geometry=geom, bias=(-b, 0.0, 0.0),
                           duration=duration, arrival=arrival, rid=rid)


def test_served_switch_equals_plain_run():
    reqs = [_switch(i, size=s, duration=d, b=b) for i, (s, d, b) in
            enumerate([((20, 10, 1), 2e-11, 0.1), ((32, 12, 2), 1.5e-11, 0.05), ((17, 17, 3), 2.5e-11, 0.15)])]
    res = {r.rid: r for r in serving.Server(20000).serve(reqs, realtime=False)}
    for req in reqs:
        w, m = req.solve_plain()
        got = res[req.rid]
        assert got.world.timesolver.time == pytest.approx(w.timesolver.time, rel=1e-9)
        assert np.abs(np.array(got.magnet.magnetization.average()) - np.array(m.magnetization.average())).max() <= 1e-5


def test_result_is_the_same_alone_and_in_a_mix():
    target = _switch(99, size=(24, 12, 2), duration=3e-11, b=0.12)
    alone = serving.Server(4096).serve([target], realtime=False)[0].magnet.magnetization.get()
    others = [_switch(i, size=(30, 20, 1), duration=1e-11 * (1 + i), b=0.02 * (i + 1)) for i in range(4)]
    # room for two at a time: the target waits and joins a running batch
    mix = serving.Server(1536, poll_steps=5).serve(others[:2] + [target] + others[2:], realtime=False)
    got = [r for r in mix if r.rid == 99][0]
    first_done = min(r.finished for r in mix)
    assert got.admitted >= first_done  # admitted once a running request had finished
    assert np.array_equal(got.magnet.magnetization.get(), alone)


def test_served_switch_against_fp64_oracle():
    req = _switch(0, size=(16, 8, 1), duration=1.5e-11, b=0.08, alpha=0.1)
    res = serving.Server(1024).serve([req], realtime=False)[0]
    world, mag, ow, om = make_pair(req.size, req.cellsize, req.params, req.m0, geometry=req.geometry,
                                   world_bias=req.bias)
    ow.solver.set_method("DormandPrince")
    ow.solver.run(req.duration)
    got = np.array(res.magnet.magnetization.average())
    want = om.m[:, om.geom_mask()].mean(axis=1)
    assert np.abs(got - want).max() < 1e-4
    assert abs(res.world.timesolver.time - ow.solver.time) < 1e-3 * req.duration


def test_served_relaxation_reaches_the_plain_relaxed_state():
    nx, ny, nz = 24, 12, 1
    x = (np.arange(nx) + 0.5) / nx * 2 - 1
    y = (np.arange(ny) + 0.5) / ny * 2 - 1
    geom = np.broadcast_to((x[None, :] ** 2 + y[:, None] ** 2) <= 1.0, (nz, ny, nx)).copy()
    v = np.array([np.cos(0.7), np.sin(0.7), 0.1])
    m0 = np.empty((3, nz, ny, nx), dtype=np.float32)
    m0[:] = (v / np.linalg.norm(v))[:, None, None, None]
    params = {"msat": 8e5, "aex": 1.3e-11, "ku1": 3e5, "anisU": (1.0, 0.0, 0.0), "alpha": 0.02}
    req = serving.Request("relax", (nx, ny, nz), (2e-9,) * 3, params, m0, geometry=geom, threshold=1e9, rid=0)
    res = serving.Server(1024).serve([req], realtime=False)[0]
    assert np.linalg.norm(res.magnet.damping_torque.eval(), axis=0).max() <= 1e9 * (1 + 1e-5)
    assert res.max_torque <= 1e9
    w, m = req.solve_plain()
    assert np.abs(np.array(res.magnet.magnetization.average()) - np.array(m.magnetization.average())).max() < 1e-2


def test_served_relaxation_against_fp64_oracle():
    """A served relaxation (damping-only Dormand-Prince with the adaptive step,
    until the max damping torque is below the threshold) against the FP64
    oracle stepping the same damping-only dynamics from the same state with
    the same stopping rule: <m> within 1e-4."""
    from oracle import relax as orelax

    nx, ny, nz = 20, 12, 1
    x = (np.arange(nx) + 0.5) / nx * 2 - 1
    y = (np.arange(ny) + 0.5) / ny * 2 - 1
    geom = np.broadcast_to((x[None, :] ** 2 + y[:, None] ** 2) <= 1.0, (nz, ny, nx)).copy()
    v = np.array([np.cos(0.6), np.sin(0.6), 0.15])
    m0 = np.empty((3, nz, ny, nx), dtype=np.float32)
    m0[:] = (v / np.linalg.norm(v))[:, None, None, None]
    params = {"msat": 8e5, "aex": 1.3e-11, "ku1": 3e5, "anisU": (1.0, 0.0, 0.0), "alpha": 0.02}
    thr = 2e7
    # a field across the easy axis: the relaxed state is canted
    req = serving.Request("relax", (nx, ny, nz), (2e-9,) * 3, params, m0, geometry=geom, bias=(0.0, 0.12, 0.03),
                          threshold=thr, rid=0)
    res = serving.Server(1024).serve([req], realtime=False)[0]
    assert res.max_torque <= thr
    world, mag, ow, om = make_pair(req.size, req.cellsize, req.params, req.m0, geometry=req.geometry,
                                   world_bias=req.bias)
    ts = ow.solver
    ts.set_method("DormandPrince")
    ts.adaptive = True
    ow.relax_fms = [om]
    ts.equations_changed()
    g = om.geom_mask()
    steps = 0
    while orelax._max_norm(orelax.damping_torque(ow, om), g) > thr:
        ts.step()
        steps += 1
        assert steps < 200000
    ow.relax_fms = None
    got = np.array(res.magnet.magnetization.average())
    want = om.m[:, g].mean(axis=1)
    assert np.abs(got - want).max() < 1e-4


def test_admission_respects_capacity_and_every_request_finishes():
    reqs = serving.disk_requests(3, 12, 1e6)
    for r in reqs:  # short runs for the test
        if r.kind == "switch":
            r.duration = 5e-12
        else:
            r.threshold = 3e9
    cap = 16384
    server = serving.Server(cap, poll_steps=8)
    used = []
    orig = server.batch.advance

    def advance(n):
        used.append(server.batch.info()["capacity"] - server.batch.info()["free_cells"])
        orig(n)

    server.batch.advance = advance
    res = server.serve(reqs, realtime=False)
    assert sorted(r.rid for r in res) == sorted(r.rid for r in reqs)
    assert max(used) <= server.batch.info()["capacity"]
    assert server.batch.info()["members"] == 0
    for r in res:
        if r.request.kind == "switch":
            assert r.world.timesolver.time == pytest.approx(5e-12, rel=1e-9)
