"""Throughput benchmark: cell-steps/s of Dormand-Prince RK45 with the full
effective field on a 256 x 256 x 64 ferromagnet (one MI355X).

Workload (fixed by the benchmark definition): 2 nm cubic cells, Msat 1.1e6 A/m,
Aex 1.6e-11 J/m, interfacial DMI D = 3e-3 J/m^2 (Dxxz = Dyyz = D), uniaxial
Ku1 = 8e5 J/m^3 along z, demag on, alpha = 0 (the reference default), m0 =
normalised N(0,1) vectors from numpy.random.default_rng(seed), FP32 state,
fixed dt = 5e-14 s.  cell-steps/s = ncells * steps / wall seconds, timed
around `steps` after warm-up (kernel construction excluded), as in the
reference's examples/bench.py.

Output: one JSON line with the metric, a per-kernel roofline (HIP events
around every launch of a second, profiled pass over the same steps) and the
FP64 CPU oracle timed on a bounded sample of the same workload.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline] [--no-sub]

The same line carries, as "batched" and "served", the measurements of
`--workload batch` and `--workload serve` below (their own step counts; --no-sub
leaves them out, e.g. under a profiler).

`--workload afm` measures instead the two-sublattice antiferromagnet with
magnetoelastic coupling and elastodynamics (512 x 512 x 1, 2 nm cells,
periodic in x and y with two repetitions, Ms 566e3 A/m, A 2.48e-12 J/m,
afmex_cell -4e-12 J/m, bulk DMI 1e-3 J/m^2 on the inter- and intra-sublattice
tensors, alpha 0.004, B1 = B2 = -8.8e6 J/m^3, C11/C12/C44 = 283/166/58 GPa,
rho 8900 kg/m^3, eta 1e10, u = v = 0, m1 random and m2 = -m1 + 1% noise,
Dormand-Prince at a fixed dt of 1e-14 s) and prints one JSON line for it.

`--workload batch` measures batched solves/s: 256 independent ferromagnets
(log2 of nx, ny uniform in [4, 7], log2 nz in [0, 2], 2 nm cells, Ms, A, Ku
along z and alpha drawn uniformly from [4e5, 1.4e6], [5e-12, 2e-11], [0, 1e6],
[0.005, 0.1]; exchange + demag + anisotropy), one solve being 100 fixed
Dormand-Prince steps at 1e-13 s, stepped together by mumaxplus.Batch, next to
the same instances stepped one World at a time.

`--workload serve` measures served solves/s and p50 / p99 latency of a
Poisson stream of 128 relaxation and switching requests on elliptical disks
(mumaxplus.serving, continuous batching), next to the first 12 requests
solved one at a time through the plain API with the same arrival times (their
mean service time gives the plain path's saturated solves/s).
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "mumax-plus_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK = 8.0e12  # B/s, MI355X
SIZE = (256, 256, 64)
CELL = 2e-9
MSAT, AEX, DMI, KU1, DT = 1.1e6, 1.6e-11, 3e-3, 8e5, 5e-14


def cpu_model():
    """The host CPU's model name (from /proc/cpuinfo), for the CPU baselines."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def initial_m(seed):
    nx, ny, nz = SIZE
    rng = np.random.default_rng(seed)
    m = rng.normal(size=(3, nz, ny, nx))
    m /= np.linalg.norm(m, axis=0)
    return m


def build_world(seed):
    import mumaxplus as mp
    world = mp.World((CELL, CELL, CELL))
    mag = mp.Ferromagnet(world, mp.Grid(SIZE))
    mag.msat = MSAT
    mag.aex = AEX
    mag.ku1 = KU1
    mag.anisU = (0, 0, 1)
    mag.dmi_tensor.set_interfacial_dmi(DMI)
    mag.magnetization = initial_m(seed).astype(np.float32)
    ts = world.timesolver
    ts.set_method("DormandPrince")
    ts.adaptive_timestep = False
    ts.timestep = DT
    return world, mag


AFM_SIZE = (512, 512, 1)
AFM_DT = 1e-14


def build_afm_world(seed):
    import mumaxplus as mp
    world = mp.World((CELL, CELL, CELL), mastergrid=mp.Grid((512, 512, 0)), pbc_repetitions=(2, 2, 0))
    afm = mp.Antiferromagnet(world, mp.Grid(AFM_SIZE))
    afm.msat = 566e3
    afm.aex = 2.48e-12
    afm.alpha = 0.004
    afm.afmex_cell = -4e-12
    afm.dmi_tensors.set_bulk_dmi(1e-3)
    afm.B1 = -8.8e6
    afm.B2 = -8.8e6
    afm.enable_elastodynamics = True
    afm.C11 = 283e9
    afm.C12 = 166e9
    afm.C44 = 58e9
    afm.rho = 8900.0
    afm.eta = 1e10
    nx, ny, nz = AFM_SIZE
    rng = np.random.default_rng(seed)
    m1 = rng.normal(size=(3, nz, ny, nx))
    m1 /= np.linalg.norm(m1, axis=0)
    afm.sub1.magnetization = m1.astype(np.float32)
    afm.sub2.magnetization = (-m1 + 0.01 * rng.normal(size=m1.shape)).astype(np.float32)
    ts = world.timesolver
    ts.set_method("DormandPrince")
    ts.adaptive_timestep = False
    ts.timestep = AFM_DT
    return world, afm


def afm_step_bytes(N, demag_bytes):
    """Algorithmic HBM bytes of one DP step (FSAL: 6 RHS evaluations) of the
    antiferromagnet with elastodynamics: per evaluation the two-sublattice
    demag (demag_bytes, both magnetisations read), strain and stress (read u,
    write 12 floats: 60 B), two field kernels (read m, sister m, h, strain;
    write the torque: 72 B each) and the acceleration (read stress, v, m1,
    m2; write a: 72 B); per step the stage combinations of the four 3-vector
    variables m1, m2, u, v (384 B each) and, per stage, the stage velocity
    written once more as du/dt (12 B)."""
    per_eval = demag_bytes + (60.0 + 2 * 72.0 + 72.0) * N
    return 6 * per_eval + 4 * 384.0 * N + 6 * 12.0 * N


def set_world_options(impl, args):
    """--world-option KEY=VALUE (repeatable): a world switch of the C ABI
    (mpa_world_set_option), e.g. split_demag=1."""
    for kv in args.world_option or []:
        k, v = kv.split("=", 1)
        impl.set_option(k, int(v))


def run_afm(args):
    import mumaxplus as mp
    if not mp._cpp.device_available():
        raise RuntimeError("no HIP device visible")
    world, afm = build_afm_world(args.seed)
    ts = world.timesolver
    impl = world._impl
    if args.serial_elastic:
        impl.set_option("split_elastic", 0)
        impl.set_option("split_demag", 0)
    if args.split_elastic:
        impl.set_option("split_elastic", 1)
    set_world_options(impl, args)
    nx, ny, nz = AFM_SIZE
    N = nx * ny * nz
    steps = args.steps if args.steps_given else 500
    ts.steps(args.warmup)
    t0 = time.perf_counter()
    ts.steps(steps)
    wall = time.perf_counter() - t0
    impl.set_profiling(True)
    impl.profile_reset()
    ts.steps(args.profile_steps)
    rep = impl.profile_report()
    impl.set_profiling(False)
    tot_ms = sum(v[0] for v in rep.values())
    dom = max(rep, key=lambda k: rep[k][0])
    d_ms, d_n, d_b = rep[dom]
    achieved = (d_b / d_n) / (d_ms / d_n * 1e-3)
    kernels = {k: {"avg_us": round(v[0] / v[1] * 1e3, 2), "launches": v[1],
                   "GBps": round((v[2] / v[1]) / (v[0] / v[1] * 1e-3) / 1e9, 1),
                   "share": round(v[0] / tot_ms, 4)} for k, v in sorted(rep.items(), key=lambda kv: -kv[1][0])}
    plan = mp._cpp.DemagPlan(AFM_SIZE, (CELL, CELL, CELL), (512, 512, 0), (2, 2, 0))
    demag_bytes = plan.info()[1] + 12.0 * N
    b_step = afm_step_bytes(N, demag_bytes)
    t_step = wall / steps
    out = {
        "metric": "cell-steps/s (RK45, full Heff) at 512x512x1 AFM + magnetoelastic + elastodynamics",
        "value": N * steps / wall,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "antiferromagnet 512x512x1 with magnetoelastic coupling and elastodynamics, "
                               "Dormand-Prince fixed dt 1e-14 s",
                   "grid": list(AFM_SIZE), "cellsize_m": CELL, "pbc": [2, 2, 0], "msat": 566e3, "aex": 2.48e-12,
                   "afmex_cell": -4e-12, "bulk_dmi": 1e-3, "alpha": 0.004, "B1": -8.8e6, "B2": -8.8e6,
                   "C11": 283e9, "C12": 166e9, "C44": 58e9, "rho": 8900.0, "eta": 1e10, "dt": AFM_DT,
                   "method": "DormandPrince", "seed": args.seed,
                   "split_elastic": impl.get_option("split_elastic"), "split_demag": impl.get_option("split_demag")},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": None,
                     "bytes_per_launch": d_b / d_n,
                     "step": {"algorithmic_bytes": b_step, "achieved_GBps": b_step / t_step / 1e9,
                              "frac": b_step / t_step / HBM_PEAK}},
        "kernels": kernels,
        "cpu_baseline": None,
        "neel_average": [round(float(v), 7) for v in afm.neel_vector.average()],
        "u_max": float(np.abs(afm.elastic_displacement.get()).max()),
    }
    print(json.dumps(out))


BATCH_N, BATCH_STEPS, BATCH_DT = 256, 100, 1e-13


def batch_specs(seed):
    """The batched workload: BATCH_N ferromagnets with log2(nx), log2(ny)
    uniform in [4, 7] and log2(nz) uniform in [0, 2] (rounded), 2 nm cells,
    Ms in [4e5, 1.4e6] A/m, A in [5e-12, 2e-11] J/m, Ku in [0, 1e6] J/m^3
    along z, alpha in [0.005, 0.1], random unit m0; exchange + demag +
    anisotropy."""
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(BATCH_N):
        nx, ny = (int(2 ** round(v)) for v in rng.uniform(4, 7, 2))
        nz = int(2 ** round(rng.uniform(0, 2)))
        p = {"msat": rng.uniform(4e5, 1.4e6), "aex": rng.uniform(5e-12, 2e-11), "ku1": rng.uniform(0, 1e6),
             "alpha": rng.uniform(0.005, 0.1)}
        m = rng.normal(size=(3, nz, ny, nx))
        m /= np.linalg.norm(m, axis=0)
        specs.append(((nx, ny, nz), p, m.astype(np.float32)))
    return specs


def build_batch_member(spec):
    import mumaxplus as mp
    size, p, m0 = spec
    world = mp.World((CELL, CELL, CELL))
    mag = mp.Ferromagnet(world, mp.Grid(size))
    mag.msat = p["msat"]
    mag.aex = p["aex"]
    mag.ku1 = p["ku1"]
    mag.anisU = (0, 0, 1)
    mag.alpha = p["alpha"]
    mag.magnetization = m0
    ts = world.timesolver
    ts.set_method("DormandPrince")
    ts.adaptive_timestep = False
    ts.timestep = BATCH_DT
    mag.demag_field.eval()  # builds the demag plan (kernel construction is not timed)
    return world, mag


def batch_cpu_baseline(specs, budget_s=20.0):
    """FP64 oracle: one Dormand-Prince step of every 32nd instance, time per
    cell-step extrapolated to BATCH_STEPS steps of all instances."""
    from oracle import demag as odemag
    from oracle import model as omodel

    cores = os.cpu_count() or 1
    odemag.set_fft_workers(cores)
    t_tot, cells = 0.0, 0
    picked = 0
    for size, p, m0 in specs[::32]:
        ow = omodel.World((CELL, CELL, CELL))
        om = ow.add_ferromagnet(size, defer_timestep=True)
        for k, v in p.items():
            om.set(k, v)
        om.set("anisU", (0.0, 0.0, 1.0))
        om.set_m(m0.astype(np.float64))
        om.demag_plan()
        ow.solver.set_method("DormandPrince", defer_timestep=True)
        ow.solver.adaptive = False
        ow.solver.dt = BATCH_DT
        t0 = time.perf_counter()
        ow.solver.step()
        t_tot += time.perf_counter() - t0
        cells += size[0] * size[1] * size[2]
        picked += 1
        if t_tot > budget_s:
            break
    all_cells = sum(sz[0] * sz[1] * sz[2] for sz, _, _ in specs)
    t_all = t_tot / cells * all_cells * BATCH_STEPS
    return {"value": len(specs) / t_all, "unit": "solves/s", "cores": cores, "cpu_model": cpu_model(), "kind": "port",
            "sample": f"one FP64 Dormand-Prince step of {picked} of the {len(specs)} instances (every 32nd), "
                      f"time per cell-step scaled to {BATCH_STEPS} steps of all instances",
            "seconds": round(t_tot, 3)}


def measure_batch(args, steps):
    """Batched solves: every instance advanced `steps` fixed Dormand-Prince
    steps; solves/s = instances / wall seconds from the first admission to
    the last result copied back, against the same instances stepped one
    World at a time."""
    import mumaxplus as mp
    if not mp._cpp.device_available():
        raise RuntimeError("no HIP device visible")
    specs = batch_specs(args.seed)
    cells = sum(sz[0] * sz[1] * sz[2] for sz, _, _ in specs)
    span = sum(-(-sz[0] * sz[1] * sz[2] // 256) * 256 for sz, _, _ in specs)
    # warm-up: every kernel once, on a few instances
    warm = [build_batch_member(sp) for sp in specs[:8]]
    wb = mp.Batch(span)
    for _, mg in warm:
        wb.add(mg, steps=2)
    wb.run_until_done(poll=2)
    for _, mg in warm:
        _, _ = mg, mg.magnetization.average()
    # the step path: under "auto" the autotuner first times the candidates
    # on the full membership (its size class), then the timed run uses its
    # choice
    mode = args.batch_step_mode
    prof = [build_batch_member(sp) for sp in specs]
    pb = mp.Batch(span)
    pb.step_mode = mode
    for _, mg in prof:
        pb.add(mg, steps=10 ** 9)
    for _ in range(12 if mode == "auto" else 1):
        pb.advance(5)
    # batched path
    members = [build_batch_member(sp) for sp in specs]
    batch = mp.Batch(span)
    batch.step_mode = mode
    t0 = time.perf_counter()
    for _, mg in members:
        if batch.add(mg, steps=steps) is None:
            raise RuntimeError("batch capacity too small")
    batch.run_until_done(poll=steps)
    wall = time.perf_counter() - t0
    path = batch.step_path()
    # plain sequential path: the same instances one World at a time
    wall_seq, dm, dm_default = None, None, None
    if not args.no_plain:
        plain = [build_batch_member(sp) for sp in specs]
        plain[0][0].timesolver.steps(1)  # warm-up of the single-magnet kernels
        plain[0] = build_batch_member(specs[0])
        t0 = time.perf_counter()
        for w, _ in plain:
            w.timesolver.steps(steps)
        wall_seq = time.perf_counter() - t0
        dm_default = max(float(np.abs(a.magnetization.get() - b.magnetization.get()).max())
                         for (_, a), (_, b) in zip(members, plain))
        # the same instances through the plain path with the batch's own
        # kernels (per-cell field kernel, separate stage combination): equal
        # bit for bit.  The default plain kernels (halo tiles, stage states
        # in the field launch) round differently in the last bit, and the
        # random initial states amplify that over 100 steps.
        matched = [build_batch_member(sp) for sp in specs]
        for w, _ in matched:
            w._impl.set_option("fuse_stages", 0)
            w._impl.set_option("field_zchunk", 0)
            w.timesolver.steps(steps)
        dm = max(float(np.abs(a.magnetization.get() - b.magnetization.get()).max())
                 for (_, a), (_, b) in zip(members, matched))
    # steady-state step time on the full membership (the chosen path), and
    # the per-kernel roofline of the per-pass launches
    t0 = time.perf_counter()
    pb.advance(args.profile_steps)
    t_step = (time.perf_counter() - t0) / args.profile_steps
    info = pb.info()
    pb.set_profiling(True)
    pb.advance(min(args.profile_steps, 20))
    rep = pb.profile_report()
    pb.set_profiling(False)
    tot_ms = sum(v[0] for v in rep.values())
    with_bytes = {k: v for k, v in rep.items() if v[2] > 0}
    dom = max(with_bytes, key=lambda k: with_bytes[k][0])
    d_ms, d_n, d_b = rep[dom]
    achieved = (d_b / d_n) / (d_ms / d_n * 1e-3)
    kernels = {k: {"avg_us": round(v[0] / v[1] * 1e3, 2), "launches": v[1],
                   "GBps": round((v[2] / v[1]) / (v[0] / v[1] * 1e-3) / 1e9, 1) if v[2] > 0 else None,
                   "share": round(v[0] / tot_ms, 4)} for k, v in sorted(rep.items(), key=lambda kv: -kv[1][0])}
    b_step = info["step_bytes"]
    cpu = None if args.no_cpu_baseline else batch_cpu_baseline(specs)
    out = {
        "metric": "batched solves/s (256 independent ferromagnets, RK45)",
        "value": len(specs) / wall,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 2,
        "ms_per_step": 1e3 * t_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"{len(specs)} ferromagnets, log2(nx), log2(ny) in [4, 7], log2(nz) in [0, 2], "
                               "2 nm cells, exchange + demag + anisotropy, one solve = "
                               f"{steps} Dormand-Prince steps at a fixed dt of 1e-13 s",
                   "instances": len(specs), "cells": cells, "packed_cells": span, "dt": BATCH_DT,
                   "method": "DormandPrince", "seed": args.seed},
        "sequential": None if wall_seq is None else {
            "value": len(specs) / wall_seq, "unit": "solves/s", "wall_s": wall_seq, "speedup": wall_seq / wall,
            "max_abs_dm_vs_batched": dm, "kernels_for_dm": "plain path with the batch's kernels "
            "(world options fuse_stages 0, field_zchunk 0)",
            "max_abs_dm_vs_batched_default_kernels": dm_default},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": None,
                     "bytes_per_launch": d_b / d_n,
                     "step": {"algorithmic_bytes": b_step, "achieved_GBps": b_step / t_step / 1e9,
                              "frac": b_step / t_step / HBM_PEAK,
                              "launches": info["launches_per_step"]}},
        "kernels": kernels,
        "step_path": {"mode": mode, "taken": path["last_path"], "persist_grid": path["persist_grid"],
                      "tuning": pb.tune_report() if mode == "auto" else None},
        "cpu_baseline": cpu,
    }
    return out


def run_batch(args):
    print(json.dumps(measure_batch(args, args.steps if args.steps_given else BATCH_STEPS)))


SERVE_N, SERVE_RATE, SERVE_CAP, SERVE_PLAIN = 128, 100.0, 262144, 24


def _pct(v, q):
    return float(np.percentile(np.asarray(v), q)) if len(v) else None


def serve_cpu_baseline(req, steps, budget_s=20.0):
    """FP64 oracle: a few Dormand-Prince steps of one switching request at
    the step size the GPU ended with, scaled to the steps the GPU needed."""
    from oracle import demag as odemag
    from oracle import model as omodel

    cores = os.cpu_count() or 1
    odemag.set_fft_workers(cores)
    ow = omodel.World(req.cellsize)
    om = ow.add_ferromagnet(req.size, geometry=req.geometry, defer_timestep=True)
    for k, v in req.params.items():
        om.set(k, v)
    ow.bias = req.bias
    om.set_m(req.m0.astype(np.float64))
    om.demag_plan()
    ow.solver.set_method("DormandPrince", defer_timestep=True)
    ow.solver.adaptive = False
    ow.solver.dt = req.duration / max(steps, 1)
    n, t = 0, 0.0
    while n < 20 and t < budget_s:
        t0 = time.perf_counter()
        ow.solver.step()
        t += time.perf_counter() - t0
        n += 1
    per_solve = t / n * steps
    return {"value": 1.0 / per_solve, "unit": "solves/s", "cores": cores, "cpu_model": cpu_model(), "kind": "port",
            "sample": f"{n} FP64 Dormand-Prince steps of request {req.rid} ({req.size[0]}x{req.size[1]}x{req.size[2]}, "
                      f"{req.duration * 1e9:.2f} ns switching run), scaled to the {steps} steps the GPU took",
            "seconds": round(t, 3)}


def measure_serve(args):
    """Served solves: Poisson arrivals of relaxation and switching requests on
    elliptical disks, continuous batching; solves/s over the whole stream and
    p50 / p99 latency from arrival to result, next to the same requests
    solved one at a time through the plain API (the first SERVE_PLAIN
    requests, with their arrival times)."""
    import mumaxplus as mp
    from mumaxplus import serving
    if not mp._cpp.device_available():
        raise RuntimeError("no HIP device visible")
    reqs = serving.disk_requests(args.seed, SERVE_N, SERVE_RATE)
    # warm-up of every kernel
    warm = serving.disk_requests(args.seed + 1, 4, 1e6)
    for r in warm:
        r.duration = 2e-12 if r.kind == "switch" else None
        r.threshold = 1e11 if r.kind == "relax" else None
    serving.Server(SERVE_CAP, step_mode=args.batch_step_mode).serve(warm, realtime=False)
    for r in warm[:1]:
        r.solve_plain()
    # served stream
    server = serving.Server(SERVE_CAP, poll_steps=args.serve_poll, step_mode=args.batch_step_mode)
    t0 = time.perf_counter()
    res = server.serve(reqs, realtime=True)
    wall = time.perf_counter() - t0
    print(f"served {len(res)} requests in {wall:.2f} s", file=sys.stderr, flush=True)
    lat = [r.latency for r in res]
    byid = {r.rid: r for r in res}
    # the first requests one at a time through the plain API, same arrivals
    sub = reqs[:SERVE_PLAIN] if not args.no_plain else []
    plain_lat, dms, service = [], [], []
    t0 = time.perf_counter()
    for r in sub:
        now = time.perf_counter() - t0
        if r.arrival > now:
            time.sleep(r.arrival - now)
        ts = time.perf_counter()
        w, m = r.solve_plain()
        service.append(time.perf_counter() - ts)
        plain_lat.append(time.perf_counter() - t0 - r.arrival)
        print(f"plain request {r.rid} ({r.kind}) done at {time.perf_counter() - t0:.2f} s", file=sys.stderr,
              flush=True)
        if r.kind == "switch":
            dms.append(float(np.abs(np.array(m.magnetization.average())
                                    - np.array(byid[r.rid].magnet.magnetization.average())).max()))
    wall_plain = time.perf_counter() - t0
    # the same first requests served, for a like-for-like latency
    sub_server = serving.Server(SERVE_CAP, poll_steps=args.serve_poll, step_mode=args.batch_step_mode)
    res_sub = sub_server.serve(sub, realtime=True) if sub else []
    sub_prof = reqs[:SERVE_PLAIN]
    # per-kernel roofline of a profiled pass over the first requests
    pserver = serving.Server(SERVE_CAP, poll_steps=args.serve_poll)
    pserver.batch.set_profiling(True)
    pserver.serve(sub_prof, realtime=False)
    rep = pserver.batch.profile_report()
    tot_ms = sum(v[0] for v in rep.values())
    with_bytes = {k: v for k, v in rep.items() if v[2] > 0}
    dom = max(with_bytes, key=lambda k: with_bytes[k][0])
    d_ms, d_n, d_b = rep[dom]
    achieved = (d_b / d_n) / (d_ms / d_n * 1e-3)
    kernels = {k: {"avg_us": round(v[0] / v[1] * 1e3, 2), "launches": v[1],
                   "GBps": round((v[2] / v[1]) / (v[0] / v[1] * 1e-3) / 1e9, 1) if v[2] > 0 else None,
                   "share": round(v[0] / tot_ms, 4)} for k, v in sorted(rep.items(), key=lambda kv: -kv[1][0])}
    steps_total = sum(r.steps for r in res)
    first_switch = next(r for r in reqs if r.kind == "switch")
    cpu = None if args.no_cpu_baseline else serve_cpu_baseline(first_switch, byid[first_switch.rid].steps)
    out = {
        "metric": "served solves/s + p50/p99 latency (Poisson stream of disk relaxation / switching requests)",
        "value": len(res) / wall,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": steps_total,
        "warmup": len(warm),
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"{SERVE_N} requests, Poisson arrivals at {SERVE_RATE}/s, elliptical disks "
                               "(major axis 20-100 nm, aspect 1-3, 1-3 layers, 2 nm cells), half relaxations "
                               "(damping-only until max torque <= 1e9 /s), half switching runs (10-200 mT "
                               "pulses against the easy axis for 0.1-2 ns), adaptive Dormand-Prince",
                   "requests": SERVE_N, "rate_per_s": SERVE_RATE, "capacity_cells": SERVE_CAP,
                   "poll_steps": args.serve_poll, "seed": args.seed},
        "latency_s": {"p50": _pct(lat, 50), "p99": _pct(lat, 99), "max": max(lat)},
        "wall_s": wall,
        "server": dict(server.stats, ms_per_batch_step=1e3 * server.stats["advance_s"] / max(server.stats["batch_steps"], 1)),
        "step_path": {"mode": args.batch_step_mode,
                      "tuning": server.batch.tune_report() if args.batch_step_mode == "auto" else None},
        "plain": None if not sub else {"requests": len(sub), "value": len(sub) / wall_plain, "unit": "solves/s",
                  "mean_service_s": float(np.mean(service)),
                  "saturated_value": 1.0 / float(np.mean(service)),
                  "served_over_saturated_plain": (len(res) / wall) * float(np.mean(service)),
                  "latency_s": {"p50": _pct(plain_lat, 50), "p99": _pct(plain_lat, 99)},
                  "served_same_requests": {"value": len(sub) / max(r.finished for r in res_sub),
                                           "latency_s": {"p50": _pct([r.latency for r in res_sub], 50),
                                                         "p99": _pct([r.latency for r in res_sub], 99)}},
                  "max_abs_dm_switch_vs_served": max(dms) if dms else None},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": None, "bytes_per_launch": d_b / d_n},
        "kernels": kernels,
        "cpu_baseline": cpu,
    }
    return out


def run_dft8(args):
    """One radix-8 FFT stage, 32 rounds in LDS over 2^20 lines, by vector-ALU
    butterflies and as a matrix-core product: the measurement behind the FFT
    passes' choice of butterflies."""
    import mumaxplus as mp
    from mumaxplus import _mumaxpluscpp as cpp
    if not mp._cpp.device_available():
        raise RuntimeError("no HIP device visible")
    nlines, rounds = 1 << 20, 32
    rng = np.random.default_rng(args.seed)
    lines = ((rng.random((nlines, 8)) - 0.5) + 1j * (rng.random((nlines, 8)) - 0.5)).astype(np.complex64) * 0.1
    res = {}
    for v, name in ((0, "valu_butterflies"), (1, "mfma_16x16x4_f32")):
        best = None
        for _ in range(3):
            _, ms = cpp._bench_dft8(v, lines, rounds)
            best = ms if best is None else min(best, ms)
        # per 8-point DFT: butterflies 64 real adds and 8 real multiplies by
        # sqrt(1/2) (plus the twiddles a pass adds); the matrix form 16 x 16
        # multiply-adds
        res[name] = {"ms": best, "dfts_per_s": nlines * rounds / (best * 1e-3)}
    out = {
        "metric": "8-point DFTs/s of one FFT stage in LDS (vector-ALU butterflies vs matrix cores)",
        "value": res["valu_butterflies"]["dfts_per_s"],
        "unit": "DFT/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"{nlines} lines of 8 complex points, {rounds} rounds each in LDS, block barrier "
                               "between rounds", "lines": nlines, "rounds": rounds},
        "variants": res,
        "mfma_over_valu": res["valu_butterflies"]["ms"] / res["mfma_16x16x4_f32"]["ms"],
    }
    print(json.dumps(out))


def run_serve(args):
    print(json.dumps(measure_serve(args)))


def step_bytes(N, Hx, Ky, Hy, Hz, ny, nz):
    """Algorithmic HBM bytes of one DP step with FSAL (6 RHS evaluations)."""
    b_dm = 12 * N + 48.0 * Hx * ny * nz + 48.0 * Hx * Ky * nz + 24.0 * Hx * Hy * Hz + 48.0 * Hx * Ky * nz \
        + 48.0 * Hx * ny * nz + 12 * N
    b_loc = 36.0 * N
    b_rk = sum(12.0 * (2 + s) * N for s in range(1, 7))  # stage combinations, s = 1..6 (384 N)
    return 6 * (b_dm + b_loc) + b_rk


def cpu_baseline(seed):
    """FP64 oracle on the host cores: one timed Dormand-Prince step of the
    same workload (7 RHS evaluations, as the reference evaluates every
    stage), after the kernel construction."""
    import scipy.fft

    from oracle import model as omodel
    from oracle import demag as odemag

    cores = os.cpu_count() or 1
    odemag.set_fft_workers(cores)
    ow = omodel.World((CELL, CELL, CELL))
    om = ow.add_ferromagnet(SIZE, defer_timestep=True)
    om.set("msat", MSAT)
    om.set("aex", AEX)
    om.set("ku1", KU1)
    om.set("anisU", (0.0, 0.0, 1.0))
    om.set("xxz", DMI)
    om.set("yyz", DMI)
    om.set_m(initial_m(seed))
    om.demag_plan()  # kernel construction is outside the timed sample
    ow.solver.set_method("DormandPrince", defer_timestep=True)
    ow.solver.adaptive = False
    ow.solver.dt = DT
    N = SIZE[0] * SIZE[1] * SIZE[2]
    t0 = time.perf_counter()
    ow.solver.step()
    t = time.perf_counter() - t0
    sample = ("one timed Dormand-Prince step (7 RHS evaluations) of the 256x256x64 workload in FP64, "
              "numpy + scipy.fft on all host cores")
    del scipy
    return {"value": N / t, "unit": "cell-steps/s", "cores": cores, "cpu_model": cpu_model(), "kind": "port",
            "sample": sample, "seconds": round(t, 3)}


def pmc_traffic(kernel):
    """HBM bytes per launch of `kernel` measured with rocprofv3 FETCH_SIZE and
    WRITE_SIZE counters (scaled by a copy of known size), from the committed
    summary profiles/pmc_traffic_256x256x64.txt; None if it has no entry."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_256x256x64.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            fields = line.split()
            if len(fields) >= 4 and not fields[0].startswith("#") and fields[0] == kernel:
                return float(fields[3])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--profile-steps", type=int, default=50)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true",
                    help="batch / serve: skip the one-at-a-time comparison (for profiler runs)")
    ap.add_argument("--workload", choices=("ferromagnet", "afm", "batch", "serve", "dft8"), default="ferromagnet")
    ap.add_argument("--no-sub", action="store_true",
                    help="default workload only: leave out the batched and served sub-measurements")
    ap.add_argument("--batch-step-mode", choices=("passes", "persistent", "auto"), default="passes",
                    help="step path of the batched and served workloads (Batch.step_mode)")
    ap.add_argument("--serial-elastic", action="store_true",
                    help="afm workload: strains, accelerations and sublattice fields on the world's stream, "
                         "not beside the demag passes (the default; also switches split_demag off)")
    ap.add_argument("--serve-poll", type=int, default=32,
                    help="served workload: batch steps per server poll (admissions and retirements between polls)")
    ap.add_argument("--world-option", action="append", metavar="KEY=VALUE",
                    help="default and afm workloads: set a world option (mpa_world_set_option), repeatable")
    ap.add_argument("--split-elastic", action="store_true",
                    help="afm workload: strains and accelerations on a second stream beside the demag passes "
                         "(world option split_elastic 1)")
    ap.add_argument("--field-zchunk", type=int, default=None,
                    help="default workload: z planes per workgroup of the field kernel's halo-tile variant "
                         "(world option field_zchunk; -1: the tile-schedule autotuner picks per grid)")
    ap.add_argument("--unfused", action="store_true",
                    help="run the stage combination as its own launch (world option fuse_stages 0), e.g. to "
                         "calibrate counters on a kernel of known traffic")
    args = ap.parse_args()
    args.steps_given = any(a == "--steps" or a.startswith("--steps=") for a in sys.argv[1:])
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported by this build", file=sys.stderr)
        sys.exit(2)
    if args.workload == "afm":
        run_afm(args)
        return
    if args.workload == "batch":
        run_batch(args)
        return
    if args.workload == "serve":
        run_serve(args)
        return
    if args.workload == "dft8":
        run_dft8(args)
        return

    import mumaxplus as mp
    if not mp._cpp.device_available():
        raise RuntimeError("no HIP device visible")
    world, mag = build_world(args.seed)
    ts = world.timesolver
    impl = world._impl
    if args.unfused:
        impl.set_option("fuse_stages", 0)
    if args.field_zchunk is not None:
        impl.set_option("field_zchunk", args.field_zchunk)
    set_world_options(impl, args)
    N = SIZE[0] * SIZE[1] * SIZE[2]

    ts.steps(args.warmup)
    t0 = time.perf_counter()
    ts.steps(args.steps)  # returns after the stream has drained
    wall = time.perf_counter() - t0

    # second pass with HIP events around every launch: per-kernel roofline
    impl.set_profiling(True)
    impl.profile_reset()
    ts.steps(args.profile_steps)
    rep = impl.profile_report()
    impl.set_profiling(False)
    tot_ms = sum(v[0] for v in rep.values())
    dom = max(rep, key=lambda k: rep[k][0])
    d_ms, d_n, d_b = rep[dom]
    achieved = (d_b / d_n) / (d_ms / d_n * 1e-3)
    kernels = {k: {"avg_us": round(v[0] / v[1] * 1e3, 2), "launches": v[1],
                   "GBps": round((v[2] / v[1]) / (v[0] / v[1] * 1e-3) / 1e9, 1),
                   "share": round(v[0] / tot_ms, 4)} for k, v in sorted(rep.items(), key=lambda kv: -kv[1][0])}

    kx, ky, kz = (512, 512, 128)
    Hx, Hy, Hz = kx // 2 + 1, ky // 2 + 1, kz // 2 + 1
    b_step = step_bytes(N, Hx, ky, Hy, Hz, SIZE[1], SIZE[2])
    t_step = wall / args.steps
    avg = mag.magnetization.average()

    traffic = pmc_traffic(dom)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args.seed)

    out = {
        "metric": "cell-steps/s (RK45, full Heff) at 256²×64",
        "value": N * args.steps / wall,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "ferromagnet 256x256x64, 2 nm cells, exchange + interfacial DMI + uniaxial "
                               "anisotropy + demag, Dormand-Prince fixed dt 5e-14 s",
                   "grid": list(SIZE), "cellsize_m": CELL, "msat": MSAT, "aex": AEX, "dmi": DMI, "ku1": KU1,
                   "alpha": 0.0, "dt": DT, "method": "DormandPrince", "seed": args.seed, "fft": [kx, ky, kz],
                   "field_zchunk": impl.get_option("field_zchunk"),
                   "field_tile": {f"{k[0]}x{k[1]}x{k[2]}/{k[3]}": v["choice"]
                                  for k, v in mp._cpp._field_tile_report().items()}},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "bytes_per_launch": d_b / d_n,
                     "step": {"algorithmic_bytes": b_step, "achieved_GBps": b_step / t_step / 1e9,
                              "frac": b_step / t_step / HBM_PEAK}},
        "kernels": kernels,
        "cpu_baseline": cpu,
        "m_average": [round(float(v), 7) for v in avg],
    }
    if not args.no_sub:
        # the batched (256 small ferromagnets) and served (request stream)
        # workloads, each with its own steps, roofline, plain-path comparison
        # and CPU baseline
        drop = ("n_gpus", "scaling", "vs_baseline", "data", "higher_is_better", "warmup")
        del world, mag
        out["batched"] = {k: v for k, v in measure_batch(args, BATCH_STEPS).items() if k not in drop}
        out["served"] = {k: v for k, v in measure_serve(args).items() if k not in drop}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
