"""Electrical potential (PoissonSystem, poissonsystem.cu; LinSolver,
linsolver.cpp; conductivity_tensor, conductivitytensor.cu) on the GPU
against the FP64 oracle (oracle/poisson.py):

* conductivity_tensor with AMR and per-cell conductivity: 1e-6 relative;
* electrical_potential with applied potentials from a function (NaN where
  none), a hole in the geometry and a strip without conductivity, solved to
  a residual of 1e-12 with the default Jacobi stepper: within 2e-7 V (FP32) of the
  direct solution;
* every stepper, step by step, on an anisotropic (AMR) 3-D system: x after
  25 steps within 1e-7 relative of the oracle's same steps (minimal
  residual: 8 steps, 1e-6);
* the solver API: set_method, max_iter, tol, max_norm_residual, errors.
"""

import numpy as np
import pytest

from oracle import model as omodel
from oracle import poisson as op

import mumaxplus as mp

pytestmark = pytest.mark.gpu

CS = (2e-9, 3e-9, 4e-9)


def _pair(size, amr=0.0, geometry=None, m=None):
    nx, ny, nz = size
    world = mp.World(CS)
    mag = mp.Ferromagnet(world, mp.Grid(size), geometry=geometry)
    ow = omodel.World(CS)
    om = ow.add_ferromagnet(size, geometry=geometry, defer_timestep=True)
    if m is not None:
        mag.magnetization = m.astype(np.float32)
    om.set_m(mag.magnetization.get().astype(np.float64))
    mag.amr_ratio = amr
    om.set("amr_ratio", float(np.float32(amr)))  # the parameter is FP32
    return world, mag, om


def _both(mag, om, key, value):
    setattr(mag, key, value)
    om.set(key, getattr(mag, key).eval()[0].astype(np.float64))


def test_conductivity_tensor():
    rng = np.random.default_rng(1)
    size = (7, 5, 3)
    m = rng.normal(size=(3, 3, 5, 7))
    world, mag, om = _pair(size, amr=0.25, m=m / np.linalg.norm(m, axis=0))
    _both(mag, om, "conductivity", (1 + rng.random((3, 5, 7))).astype(np.float32))
    want = op.conductivity_tensor(om)
    got = mag.conductivity_tensor.eval()
    assert np.abs(got - want).max() <= 1e-6 * np.abs(want).max()


def test_potential_matches_the_direct_solution():
    nx, ny = 24, 16
    geom = np.ones((1, ny, nx), dtype=bool)
    geom[0, 6:9, 10:13] = False
    world, mag, om = _pair((nx, ny, 1), geometry=geom)
    sigma = np.full((1, ny, nx), 2.3, dtype=np.float32)
    sigma[0, 12, 2:18] = 0.0
    _both(mag, om, "conductivity", sigma)

    def applied(x, y, z):
        if y < 0.1 * ny * CS[1]:
            return 1.0
        if y > 0.9 * ny * CS[1] and abs(x - nx * CS[0] / 2) < 0.3 * nx * CS[0]:
            return -1.0
        return np.nan

    mag.applied_potential = applied
    om.set("applied_potential", mag.applied_potential.eval()[0].astype(np.float64))
    mag.poisson_system.tol = 1e-12
    mag.poisson_system.max_iter = 100000  # a bound on the loop; about 3000 Jacobi steps are needed
    got = mag.electrical_potential.eval()[0]
    want = op.potential(om)
    assert np.abs(got - want).max() < 2e-7  # the quantity is FP32
    assert mag.poisson_system.max_norm_residual() <= 1e-12
    assert np.array_equal(got[~geom], np.zeros(np.count_nonzero(~geom)))


@pytest.mark.parametrize("method", ["jacobi", "conjugategradient", "minimalresidual", "steepestdescent"])
def test_steppers_step_by_step(method):
    size = (6, 5, 4)
    m0 = np.zeros((3, 4, 5, 6))
    m0[0], m0[1], m0[2] = 0.6, 0.48, 0.64
    world, mag, om = _pair(size, amr=0.35, m=m0)
    _both(mag, om, "conductivity", 4.0)
    pot = np.full((4, 5, 6), np.nan, dtype=np.float32)
    pot[:, :, 0] = 0.5
    pot[:, 4, 5] = -0.25
    _both(mag, om, "applied_potential", pot)
    ps = mag.poisson_system
    ps._init()
    solver = ps._solver
    solver.set_method(method)
    # the minimal-residual iteration grows small differences (about x1.3 a
    # step on this system), so it is compared over fewer steps
    nstep, rel = (8, 1e-6) if method == "minimalresidual" else (25, 1e-7)
    for _ in range(nstep):
        solver.step()
    solver.tol = 1e300  # solve() then returns x without stepping
    got = solver.solve()
    A, b = op.rows(om)
    st = op.Stepper(A, b, method)
    for _ in range(nstep):
        st.step()
    assert np.abs(got - st.x).max() <= rel * np.abs(st.x).max()


def test_solver_api():
    world, mag, om = _pair((8, 4, 1))
    mag.conductivity = 1.0
    pot = np.full((1, 4, 8), np.nan)
    pot[0, :, 0], pot[0, :, -1] = 1.0, 0.0
    mag.applied_potential = pot
    ps = mag.poisson_system
    assert ps.tol == 1e-5 and ps.max_iter == -1
    with pytest.raises(ValueError):
        ps.set_method("gmres")
    ps._init()
    r0 = ps.max_norm_residual()
    for _ in range(10):
        ps._solver.step()
    assert ps.max_norm_residual() < r0
    ps.max_iter = 3
    ps.tol = 1e-14
    v = ps.solve()
    assert v.shape == (1, 1, 4, 8) and ps.max_norm_residual() > 1e-14
    ps.max_iter = 20000  # a bound on the loop, far above what Jacobi needs here
    ps.tol = 1e-10
    v = ps.solve()[0, 0]
    assert np.abs(v - np.linspace(1.0, 0.0, 8)[None, :]).max() < 1e-6
