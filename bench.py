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

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]

`--workload afm` measures instead the two-sublattice antiferromagnet with
magnetoelastic coupling and elastodynamics (512 x 512 x 1, 2 nm cells,
periodic in x and y with two repetitions, Ms 566e3 A/m, A 2.48e-12 J/m,
afmex_cell -4e-12 J/m, bulk DMI 1e-3 J/m^2 on the inter- and intra-sublattice
tensors, alpha 0.004, B1 = B2 = -8.8e6 J/m^3, C11/C12/C44 = 283/166/58 GPa,
rho 8900 kg/m^3, eta 1e10, u = v = 0, m1 random and m2 = -m1 + 1% noise,
Dormand-Prince at a fixed dt of 1e-14 s) and prints one JSON line for it.
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


def run_afm(args):
    import mumaxplus as mp
    if not mp._cpp.device_available():
        raise RuntimeError("no HIP device visible")
    world, afm = build_afm_world(args.seed)
    ts = world.timesolver
    impl = world._impl
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
                   "method": "DormandPrince", "seed": args.seed},
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


def step_bytes(N, Hx, Ky, Hy, Hz, ny, nz):
    """Algorithmic HBM bytes of one DP step with FSAL (6 RHS evaluations)."""
    b_dm = 12 * N + 48.0 * Hx * ny * nz + 48.0 * Hx * Ky * nz + 24.0 * Hx * Hy * Hz + 48.0 * Hx * Ky * nz \
        + 48.0 * Hx * ny * nz + 12 * N
    b_loc = 36.0 * N
    b_rk = sum(12.0 * (2 + s) * N for s in range(1, 7))  # stage combinations, s = 1..6 (384 N)
    return 6 * (b_dm + b_loc) + b_rk


def cpu_baseline(seed, budget_s=30.0):
    """FP64 oracle on the host cores: one Dormand-Prince step of the same
    workload (7 RHS evaluations as in the reference), or, if one RHS
    evaluation already exceeds the budget, that evaluation scaled to a step."""
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
    ow.rhs(ow.get_states())
    t_rhs = time.perf_counter() - t0
    if 7 * t_rhs <= budget_s:
        t0 = time.perf_counter()
        ow.solver.step()
        t = time.perf_counter() - t0
        sample = "one Dormand-Prince step (7 RHS evaluations) of the 256x256x64 workload in FP64"
    else:
        t = 7 * t_rhs
        sample = "one RHS evaluation of the 256x256x64 workload in FP64, times 7 (one step)"
    del scipy
    return {"value": N / t, "unit": "cell-steps/s", "cores": cores, "kind": "port", "sample": sample,
            "seconds": round(t, 3)}


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
    ap.add_argument("--workload", choices=("ferromagnet", "afm"), default="ferromagnet")
    args = ap.parse_args()
    args.steps_given = any(a == "--steps" or a.startswith("--steps=") for a in sys.argv[1:])
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported by this build", file=sys.stderr)
        sys.exit(2)
    if args.workload == "afm":
        run_afm(args)
        return

    import mumaxplus as mp
    if not mp._cpp.device_available():
        raise RuntimeError("no HIP device visible")
    world, mag = build_world(args.seed)
    ts = world.timesolver
    impl = world._impl
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
                   "alpha": 0.0, "dt": DT, "method": "DormandPrince", "seed": args.seed, "fft": [kx, ky, kz]},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "bytes_per_launch": d_b / d_n,
                     "step": {"algorithmic_bytes": b_step, "achieved_GBps": b_step / t_step / 1e9,
                              "frac": b_step / t_step / HBM_PEAK}},
        "kernels": kernels,
        "cpu_baseline": cpu,
        "m_average": [round(float(v), 7) for v in avg],
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
