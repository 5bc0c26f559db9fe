"""Writes film128x32_dp_fixed.npz: the FP64 oracle's trajectory of the
128x32x1 permalloy film (cells 3.90625 x 3.90625 x 3 nm, Ms=8e5 A/m,
A=1.3e-11 J/m, alpha=0.02, exchange + demag + a 25 mT in-plane bias at a
seeded angle, m0 = normalise((1, 0.1, 0) + 0.01 N(0,1)) per cell) under
Dormand-Prince at a fixed dt of 1e-13 s.  Stored: the inputs (m0 as the
FP32 unit vectors both sides start from, the bias), <m> every 1000 steps up
to 10000, and the final m.

Run from the repository root:  python tests/golden/make_film128x32.py
"""

import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import demag as odemag  # noqa: E402
from oracle import model as omodel  # noqa: E402

SEED = 0
SIZE = (128, 32, 1)
CELL = (500e-9 / 128, 125e-9 / 32, 3e-9)
PARAMS = {"msat": 8e5, "aex": 1.3e-11, "alpha": 0.02}
DT = 1e-13
CHECK_EVERY = 1000
NSTEPS = 10000


def inputs(seed=SEED):
    rng = np.random.default_rng(seed)
    phi = rng.uniform(0, 2 * np.pi)
    bias = 0.025 * np.array([np.cos(phi), np.sin(phi), 0.0])
    nx, ny, nz = SIZE
    m = np.array([1.0, 0.1, 0.0])[:, None, None, None] + 0.01 * rng.normal(size=(3, nz, ny, nx))
    m = (m / np.linalg.norm(m, axis=0)).astype(np.float32)
    # the same FP32 unit vectors start both the product and the oracle
    m = (m / np.linalg.norm(m.astype(np.float64), axis=0)).astype(np.float32)
    return m, bias


def main():
    odemag.set_fft_workers(os.cpu_count() or 1)
    m0, bias = inputs()
    w = omodel.World(CELL)
    mag = w.add_ferromagnet(SIZE)
    for k, v in PARAMS.items():
        mag.set(k, v)
    w.bias = tuple(bias)
    mag.set_m(m0.astype(np.float64))
    w.solver.set_method("DormandPrince")
    w.solver.adaptive = False
    w.solver.dt = DT
    avgs = []
    t0 = time.time()
    for i in range(NSTEPS // CHECK_EVERY):
        w.solver.steps(CHECK_EVERY)
        avgs.append(mag.m.reshape(3, -1).mean(axis=1))
        print(f"{(i + 1) * CHECK_EVERY} steps  <m>={avgs[-1]}  {time.time() - t0:.0f} s", flush=True)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "film128x32_dp_fixed.npz")
    np.savez_compressed(out, m0=m0, bias=bias, dt=DT, check_every=CHECK_EVERY,
                        averages=np.array(avgs), m_final=mag.m.astype(np.float32), time=w.solver.time)
    print("wrote", out)


if __name__ == "__main__":
    main()
