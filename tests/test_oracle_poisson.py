"""Pins of the oracle's electrical potential (oracle/poisson.py).

* the AMR tensor (conductivitytensor.cu:11-32): for m along x it is
  diag(sigma0 (1 - 2f/3), sigma0 (1 + f/3), sigma0 (1 + f/3)), f = 6r/(6+r);
  its trace is 3 sigma0; its components are those of the reference's own
  test_conductivitytensor.py:6-21;
* the un-normalised row is -V_cell div(sigma grad V) (the flux out of the
  cell): on V = x^2 and V = x y with a uniform anisotropic tensor it is
  -2 sigma_xx V_cell and -2 sigma_xy V_cell at interior cells, exactly;
* a bar with fixed potentials on its two end columns has the linear profile;
* the steppers: Jacobi's fixed point and the direct solution agree; on a
  symmetric positive definite system conjugate gradient ends within n
  steps, steepest descent and minimal residual (eigenvalues in (0.5, 1.5))
  converge.
"""

import numpy as np
import scipy.sparse as sp

from oracle import model as om
from oracle import poisson as op

CS = (2e-9, 3e-9, 4e-9)


def _mag(size, amr=0.0, m=(1.0, 0.0, 0.0)):
    mag = om.Ferromagnet(size, CS)
    mag.set("conductivity", 5.0)
    mag.set("amr_ratio", amr)
    mag.set("applied_potential", np.nan)
    mm = np.asarray(m, dtype=float) / np.linalg.norm(m)
    mag.set_m(np.broadcast_to(mm[:, None, None, None], (3,) + mag.shape).copy())
    return mag


def test_amr_tensor():
    f = 6 * 0.3 / 6.3
    # m along x: diagonal, sigma0 (1 - 2f/3) along m and sigma0 (1 + f/3) across it
    t = op.conductivity_tensor(_mag((2, 2, 1), amr=0.3, m=(1.0, 0.0, 0.0)))[:, 0, 0, 0]
    assert np.allclose(t, 5.0 * np.array([1 - 2 * f / 3, 1 + f / 3, 1 + f / 3, 0, 0, 0]), rtol=1e-13, atol=1e-13)
    # a general m: the trace is 3 sigma0, and the components are those of
    # test_conductivitytensor.py:6-21 (off-diagonal + sigma0 f m_i m_j)
    mag = _mag((3, 2, 2), amr=0.3, m=(0.3, -0.5, 0.8))
    t = op.conductivity_tensor(mag)[:, 0, 0, 0]
    mx, my, mz = mag.m[:, 0, 0, 0]
    assert np.isclose(t[:3].sum(), 15.0, rtol=1e-13)
    want = 5.0 * np.array([1 - f * (mx * mx - 1 / 3), 1 - f * (my * my - 1 / 3), 1 - f * (mz * mz - 1 / 3),
                           f * mx * my, f * mx * mz, f * my * mz])
    assert np.allclose(t, want, rtol=1e-13)


def test_rows_are_the_flux_balance_of_quadratics():
    mag = _mag((5, 5, 5), amr=0.4, m=(0.6, 0.3, 0.74))
    A, b = op.rows(mag, normalize=False)
    t = op.conductivity_tensor(mag)[:, 0, 0, 0]
    z, y, x = np.indices(mag.shape)
    X, Y, Z = (x * CS[0], y * CS[1], z * CS[2])
    V = CS[0] * CS[1] * CS[2]
    i = (2 * 5 + 2) * 5 + 2  # the centre cell
    for pot, want in ((X * X, 2 * t[0]), (Y * Y, 2 * t[1]), (Z * Z, 2 * t[2]), (X * Y, 2 * t[3]),
                      (X * Z, 2 * t[4]), (Y * Z, 2 * t[5])):
        got = (A @ pot.reshape(-1))[i]
        assert np.isclose(got, -want * V, rtol=1e-10, atol=1e-12 * abs(want * V))


def test_bar_with_fixed_ends_is_linear():
    nx = 9
    mag = _mag((nx, 3, 2))
    pot = np.full(mag.shape, np.nan)
    pot[:, :, 0] = 1.0
    pot[:, :, -1] = -0.5
    mag.set("applied_potential", pot)
    v = op.potential(mag)
    want = 1.0 - 1.5 * np.arange(nx) / (nx - 1)
    assert np.abs(v - want[None, None, :]).max() < 1e-12


def test_steppers():
    mag = _mag((6, 4, 1))
    pot = np.full(mag.shape, np.nan)
    pot[0, 0, 0], pot[0, 3, 5] = 1.0, -1.0
    mag.set("applied_potential", pot)
    A, b = op.rows(mag)
    exact = op.potential(mag).reshape(-1)
    s = op.Stepper(A, b, "jacobi")
    for _ in range(3000):
        s.step()
    assert np.abs(s.x - exact).max() < 1e-10
    rng = np.random.default_rng(0)
    n = 12
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    M = sp.csr_matrix(Q @ np.diag(np.linspace(0.6, 1.4, n)) @ Q.T)
    rhs = rng.normal(size=n)
    sol = np.linalg.solve(M.toarray(), rhs)
    cg = op.Stepper(M, rhs, "conjugategradient")
    for _ in range(n):
        cg.step()
    assert np.abs(cg.x - sol).max() < 1e-10
    for method in ("steepestdescent", "minimalresidual"):
        st = op.Stepper(M, rhs, method)
        for _ in range(200):
            st.step()
        assert np.abs(st.x - sol).max() < 1e-10, method
