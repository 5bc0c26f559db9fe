"""A request server for small micromagnetic solves on one GPU.

Requests (relaxations and field-pulse switching runs of small magnets)
arrive over time.  The server keeps one :class:`Batch` whose packed device
state holds every running request; between batch steps it retires the
requests that have finished (copying their state back) and admits waiting
ones as soon as their cells fit in the free capacity (first come, first
served).  Requests that arrive while others are on the GPU join the running
batch at the next admission point; nothing waits for the batch to drain.

Each request is an ordinary World with one Ferromagnet, exactly what a
caller of the plain API would build, so the same request can be solved one
at a time through ``World.timesolver.run`` / ``Ferromagnet.relax`` (the
plain path) and its result compared with the served one.

Workload generator: :func:`disk_requests` draws the elliptical-disk requests
(major axis 20-100 nm, aspect 1-3, 2 nm cells, 1-3 layers; Ms, A, Ku and
alpha uniform in [4e5, 1.4e6] A/m, [5e-12, 2e-11] J/m, [0, 1e6] J/m^3,
[0.005, 0.1]; easy axis along the major axis; switching pulses of 10-200 mT
against the easy axis lasting 0.1-2 ns) with Poisson arrivals.
"""

import math
import time

import numpy as np

from .batch import Batch
from .ferromagnet import Ferromagnet
from .grid import Grid
from .world import World


class Request:
    """A relaxation (`kind` "relax": damping-only dynamics until the max
    torque is at most `threshold`) or a switching run (`kind` "switch": LLG
    dynamics under the field `bias` for `duration` seconds)."""

    def __init__(self, kind, size, cellsize, params, m0, geometry=None, bias=(0.0, 0.0, 0.0), duration=None,
                 threshold=None, arrival=0.0, rid=None):
        if kind not in ("relax", "switch"):
            raise ValueError("kind is 'relax' or 'switch'")
        if kind == "switch" and not (duration and duration > 0):
            raise ValueError("a switching request needs a positive duration")
        if kind == "relax" and not (threshold and threshold > 0):
            raise ValueError("a relaxation needs a positive torque threshold")
        self.kind = kind
        self.size = tuple(int(v) for v in size)
        self.cellsize = tuple(float(v) for v in cellsize)
        self.params = dict(params)
        self.m0 = np.asarray(m0, dtype=np.float32)
        self.geometry = geometry
        self.bias = tuple(float(v) for v in bias)
        self.duration = duration
        self.threshold = threshold
        self.arrival = float(arrival)
        self.rid = rid

    @property
    def cells(self):
        return self.size[0] * self.size[1] * self.size[2]

    def build(self):
        """The request as a World holding one Ferromagnet."""
        world = World(self.cellsize)
        mag = Ferromagnet(world, Grid(self.size), geometry=self.geometry)
        for k, v in self.params.items():
            setattr(mag, k, v)
        world.bias_magnetic_field = self.bias
        mag.magnetization = self.m0
        world.timesolver.set_method("DormandPrince")
        return world, mag

    def solve_plain(self):
        """Solve through the plain API, one request at a time; returns the magnet."""
        world, mag = self.build()
        if self.kind == "switch":
            world.timesolver.run(self.duration)
        else:
            mag.RelaxTorqueThreshold = self.threshold
            mag.relax()
        return world, mag


class Result:
    def __init__(self, request, admitted, finished, world, magnet, state):
        self.request = request
        self.rid = request.rid
        self.admitted = admitted
        self.finished = finished
        self.latency = finished - request.arrival
        self.world = world
        self.magnet = magnet
        self.steps = state["steps"]
        self.time = state["time"]
        self.max_torque = state["max_torque"]


class Server:
    """Server(capacity, poll_steps=32, step_mode="passes"): continuous
    batching of requests in a Batch of `capacity` packed cells; finished
    requests are retired and waiting ones admitted every `poll_steps` batch
    steps.  `step_mode` is the batch's step path (Batch.step_mode)."""

    def __init__(self, capacity, poll_steps=32, method="DormandPrince", step_mode="passes"):
        self.batch = Batch(capacity, method)
        self.batch.step_mode = step_mode
        self.poll_steps = int(poll_steps)
        # wall seconds spent admitting (building and packing requests),
        # stepping and retiring, and the batch steps taken
        self.stats = {"admit_s": 0.0, "advance_s": 0.0, "retire_s": 0.0, "batch_steps": 0, "polls": 0}

    def serve(self, requests, realtime=True):
        """Serve `requests` (sorted by arrival, seconds from the start).  With
        `realtime`, a request is admitted only once the wall clock has passed
        its arrival; otherwise all are available at once.  Returns one Result
        per request, in completion order; latencies are wall seconds from
        arrival to the result being back in the request's magnet."""
        queue = sorted(requests, key=lambda r: r.arrival)
        cap = self.batch.info()["capacity"]
        for r in queue:
            if -(-r.cells // 256) * 256 > cap:
                raise ValueError(f"request {r.rid} needs {r.cells} cells, more than the capacity {cap}")
        t0 = time.perf_counter()
        clock = (lambda: time.perf_counter() - t0) if realtime else (lambda: math.inf)
        head = 0
        running = {}  # member id -> (request, world, magnet, admitted)
        results = []
        batch = self.batch
        while head < len(queue) or running:
            now = clock()
            # admit in arrival order while the head fits
            ta = time.perf_counter()
            while head < len(queue) and queue[head].arrival <= now:
                req = queue[head]
                world, mag = req.build()
                if req.kind == "switch":
                    mid = batch.add(mag, duration=req.duration)
                else:
                    mid = batch.add(mag, relax_torque=req.threshold)
                if mid is None:
                    break
                running[mid] = (req, world, mag, time.perf_counter() - t0)
                head += 1
            self.stats["admit_s"] += time.perf_counter() - ta
            if not running:
                wait = queue[head].arrival - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(wait)
                continue
            ta = time.perf_counter()
            batch.advance(self.poll_steps)
            done = batch.finished()
            tb = time.perf_counter()
            self.stats["advance_s"] += tb - ta
            self.stats["batch_steps"] += self.poll_steps
            self.stats["polls"] += 1
            for mid in done:
                req, world, mag, admitted = running.pop(mid)
                state = batch.member(mid)
                batch.remove(mid)
                results.append(Result(req, admitted, time.perf_counter() - t0, world, mag, state))
            self.stats["retire_s"] += time.perf_counter() - tb
        return results


def disk_requests(seed, n, rate, cellsize=2e-9, relax_fraction=0.5, threshold=1e9):
    """`n` elliptical-disk requests with Poisson arrivals of `rate` per second."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        a = rng.uniform(20e-9, 100e-9)
        r = rng.uniform(1.0, 3.0)
        layers = int(rng.integers(1, 4))
        nx = int(math.ceil(a / cellsize))
        ny = max(1, int(math.ceil(a / r / cellsize)))
        x = (np.arange(nx) + 0.5) / nx * 2 - 1
        y = (np.arange(ny) + 0.5) / ny * 2 - 1
        mask = (x[None, :] ** 2 + y[:, None] ** 2) <= 1.0
        geom = np.broadcast_to(mask, (layers, ny, nx)).copy()
        params = {"msat": rng.uniform(4e5, 1.4e6), "aex": rng.uniform(5e-12, 2e-11), "ku1": rng.uniform(0, 1e6),
                  "anisU": (1.0, 0.0, 0.0), "alpha": rng.uniform(0.005, 0.1)}
        if rng.uniform() < relax_fraction:
            th = rng.uniform(0, 2 * np.pi)
            m0 = np.empty((3, layers, ny, nx), dtype=np.float32)
            v = np.array([np.cos(th), np.sin(th), 0.1])
            v /= np.linalg.norm(v)
            m0[:] = v[:, None, None, None]
            out.append(Request("relax", (nx, ny, layers), (cellsize,) * 3, params, m0, geometry=geom,
                               threshold=threshold, arrival=t, rid=i))
        else:
            b = rng.uniform(0.010, 0.200)
            dur = rng.uniform(0.1e-9, 2e-9)
            v = np.array([1.0, 0.05, 0.02])
            v /= np.linalg.norm(v)
            m0 = np.empty((3, layers, ny, nx), dtype=np.float32)
            m0[:] = v[:, None, None, None]
            out.append(Request("switch", (nx, ny, layers), (cellsize,) * 3, params, m0, geometry=geom,
                               bias=(-b, 0.0, 0.0), duration=dur, arrival=t, rid=i))
    return out
