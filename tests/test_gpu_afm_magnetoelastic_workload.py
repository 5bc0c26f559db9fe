"""The antiferromagnet benchmark workload as configured, on the GPU.

The workload: a single-layer two-sublattice antiferromagnet, 2 nm cells,
periodic in x and y (mastergrid (Nx, Ny, 0), 2 repetitions each way), Ms 566e3
A/m, A 2.48e-12 J/m, afmex_cell -4e-12 J/m, bulk DMI 1e-3 J/m^2 on the
inter-sublattice tensor and on both sublattices' own tensors, alpha 0.004,
magnetoelastic B1 = B2 = -8.8e6 J/m^3, elastodynamics with C11/C12/C44 =
283/166/58 GPa, rho 8900 kg/m^3, eta 1e10, random m1, m2 = -m1 + 1 % noise,
u = v = 0, fixed-step Dormand-Prince at dt = 1e-14 s (examples/magnetoelastic.py:19-59
for the elastic constants; the rest as the benchmark's afm workload).

* 32x32x1: 40 fixed steps against the FP64 oracle on the same seeded state.
  Both magnetisations within 2e-4 per component (an FP32 rounding of the
  initial state grows about 30-fold over these 40 steps in FP64 itself),
  the sublattice averages within 1e-5, u and v within 2e-3 of their maxima.
* 512x512x1 (the full size): one right-hand side (both sublattices' torques and
  the elastic acceleration) against the oracle at the survey's per-term
  tolerance of 1e-5; then 20 steps keep |m| = 1 to 1e-6 and repeat bit for bit.
"""

import warnings

import numpy as np
import pytest

from conftest import rel_max_err
from oracle import model as omodel

import mumaxplus as mp

pytestmark = pytest.mark.gpu

CELL = 2e-9
DT = 1e-14
SUB = {"msat": 566e3, "aex": 2.48e-12, "alpha": 0.004, "B1": -8.8e6, "B2": -8.8e6}
BULK = {"xyz": 1e-3, "yxz": -1e-3, "zxy": 1e-3}
ELASTIC = {"C11": 283e9, "C12": 166e9, "C44": 58e9, "rho": 8900.0, "eta": 1e10}


def initial_state(n, seed):
    rng = np.random.default_rng(seed)
    m1 = rng.normal(size=(3, 1, n, n))
    m1 /= np.linalg.norm(m1, axis=0)
    m2 = -m1 + 0.01 * rng.normal(size=m1.shape)
    return m1, m2


def build_product(n, seed):
    world = mp.World((CELL, CELL, CELL), mastergrid=mp.Grid((n, n, 0)), pbc_repetitions=(2, 2, 0))
    afm = mp.Antiferromagnet(world, mp.Grid((n, n, 1)))
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", UserWarning)  # sign warnings of afmex_cell, B1, B2
        for k, v in SUB.items():
            setattr(afm, k, v)
        afm.afmex_cell = -4e-12
    afm.dmi_tensors.set_bulk_dmi(1e-3)
    afm.enable_elastodynamics = True
    for k, v in ELASTIC.items():
        setattr(afm, k, v)
    m1, m2 = initial_state(n, seed)
    afm.sub1.magnetization = m1.astype(np.float32)
    afm.sub2.magnetization = m2.astype(np.float32)
    ts = world.timesolver
    ts.set_method("DormandPrince")
    ts.adaptive_timestep = False
    ts.timestep = DT
    return world, afm


def build_oracle(n, afm):
    """The oracle twin, started from the product's (normalised FP32) state."""
    ow = omodel.World((CELL, CELL, CELL), mastergrid=(n, n, 0), pbc=(2, 2, 0))
    oa = ow.add_antiferromagnet((n, n, 1), defer_timestep=True)
    for s in oa.sublattices:
        for k, v in SUB.items():
            s.set(k, v)
        for k, v in BULK.items():
            s.set(k, v)
    for k, v in BULK.items():
        oa.set(k, v)
    oa.set("afmex_cell", -4e-12)
    oa.enable_elastodynamics = True
    for k, v in ELASTIC.items():
        oa.set(k, v)
    oa.sub1.set_m(afm.sub1.magnetization.get().astype(np.float64))
    oa.sub2.set_m(afm.sub2.magnetization.get().astype(np.float64))
    ow.solver.equations_changed()
    ow.solver.set_method("DormandPrince")
    ow.solver.adaptive = False
    ow.solver.dt = DT
    return ow, oa


def test_bulk_dmi_is_set_on_all_three_tensors():
    world, afm = build_product(8, 0)
    for t in (afm.dmi_tensor, afm.sub1.dmi_tensor, afm.sub2.dmi_tensor):
        for k, v in BULK.items():
            assert getattr(t, k).uniform_value == pytest.approx(v)
        assert t.xxz.uniform_value == 0 and t.yyz.uniform_value == 0


def test_fixed_steps_match_oracle_32x32x1():
    n = 32
    world, afm = build_product(n, 7)
    ow, oa = build_oracle(n, afm)
    world.timesolver.steps(40)
    ow.solver.steps(40)
    for sub, osub in ((afm.sub1, oa.sub1), (afm.sub2, oa.sub2)):
        got = sub.magnetization.get()
        assert np.abs(got - osub.m).max() < 2e-4
        assert np.abs(got.reshape(3, -1).mean(axis=1) - osub.m.reshape(3, -1).mean(axis=1)).max() < 1e-5
    assert rel_max_err(afm.elastic_displacement.get(), oa.u) < 2e-3
    assert rel_max_err(afm.elastic_velocity.get(), oa.v) < 2e-3
    assert np.abs(oa.u).max() > 0  # the magnetoelastic force has set the lattice moving
    assert abs(world.timesolver.time - ow.solver.time) < 1e-25


def test_full_size_rhs_matches_oracle():
    n = 512
    world, afm = build_product(n, 3)
    # a displacement and velocity so that every elastic term is non-zero
    rng = np.random.default_rng(4)
    x = np.arange(n) * CELL
    L = n * CELL
    u = np.zeros((3, 1, n, n))
    u[0] = 1e-11 * np.sin(2 * np.pi * x / L)[None, None, :] * np.cos(4 * np.pi * x / L)[None, :, None]
    u[1] = 5e-12 * np.cos(6 * np.pi * x / L)[None, :, None]
    u[2] = 2e-12 * rng.normal(size=(1, n, n))
    afm.elastic_displacement = u.astype(np.float32)
    afm.elastic_velocity = (10 * rng.normal(size=(3, 1, n, n))).astype(np.float32)
    ow, oa = build_oracle(n, afm)
    oa.u = afm.elastic_displacement.get().astype(np.float64)
    oa.v = afm.elastic_velocity.get().astype(np.float64)
    want = ow.rhs(ow.get_states())  # [tau1, tau2, du/dt, dv/dt]
    assert rel_max_err(afm.sub1.torque.eval(), want[0]) < 1e-5
    assert rel_max_err(afm.sub2.torque.eval(), want[1]) < 1e-5
    assert rel_max_err(afm.elastic_acceleration.eval(), want[3]) < 1e-5


def test_full_size_steps_keep_norm_and_repeat():
    n = 512
    runs = []
    for _ in range(2):
        world, afm = build_product(n, 5)
        world.timesolver.steps(20)
        runs.append([afm.sub1.magnetization.get(), afm.sub2.magnetization.get(),
                     afm.elastic_displacement.get(), afm.elastic_velocity.get()])
        del world, afm
    for a, b in zip(*runs):
        assert np.array_equal(a, b)
    for m in runs[0][:2]:
        assert np.abs(np.linalg.norm(m.astype(np.float64), axis=0) - 1).max() < 1e-6
    assert np.abs(runs[0][2]).max() > 0
