"""Pins of the oracle's elastic and magnetoelastic terms (oracle/elastic.py)
against the closed forms of the reference's own tests: sine displacements
and magnetisations on periodic 1-D and 2-D grids (test/test_strain.py,
test_elasticforce_uniform_c.py, test_magnetoelasticfield.py,
test_magnetoelasticforce.py), plus the traction-free boundary and the
fourth-order accuracy of the stencils."""

import math

import numpy as np
import pytest

from oracle import elastic as oe
from oracle import model as om

CS = (1.5e-9, 2e-9, 2.5e-9)
N, P, AMP = 128, 2, 4.0
SRTOL = 1e-4


def semirel(result, wanted):
    """max |result - wanted| (vector norm) / max |wanted| (test_strain.py:14-23)."""
    return np.max(np.linalg.norm(result - wanted, axis=0)) / np.max(np.abs(wanted))


def line_magnet(axis, n=N, second=None):
    """Magnet periodic along `axis` (and `second`, if given), one cell across."""
    size, mg, reps = [1, 1, 1], [0, 0, 0], [0, 0, 0]
    size[axis], mg[axis], reps[axis] = n, n, 1
    if second is not None:
        a2, n2 = second
        size[a2], mg[a2], reps[a2] = n2, n2, 1
    w = om.World(CS, mastergrid=tuple(mg), pbc=tuple(reps))
    mag = w.add_ferromagnet(tuple(size), defer_timestep=True)
    mag.enable_elastodynamics = True
    return mag


def coords(mag):
    z, y, x = np.indices(mag.shape)
    return [x * CS[0], y * CS[1], z * CS[2]]


@pytest.mark.parametrize("i,j", [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2), (1, 0), (2, 0), (2, 1)])
def test_strain_of_sine_displacement(i, j):
    """test/test_strain.py:26-100."""
    mag = line_magnet(i)
    k = P * 2 * math.pi / (N * CS[i])
    u = np.zeros((3,) + mag.shape)
    u[j] = AMP * np.sin(k * coords(mag)[i])
    got = oe.strain_tensor(mag, u)
    want = np.zeros_like(got)
    cos = AMP * np.cos(k * coords(mag)[i])
    if i == j:
        want[j] = k * cos
    else:
        want[2 + i + j] = 0.5 * k * cos
    assert semirel(got, want) < SRTOL


def test_stress_from_strain():
    """test/test_stress_tensor.py:27-56 with a random displacement, plus the
    isotropic check: a uniform expansion e gives sigma_ii = (C11 + 2 C12) e."""
    rng = np.random.default_rng(0)
    w = om.World(CS)
    mag = w.add_ferromagnet((16, 8, 1), defer_timestep=True)
    for k, v in (("C11", 283e9), ("C12", 58e9), ("C44", 166e9)):
        mag.set(k, v)
    u = rng.random((3,) + mag.shape)
    e = oe.strain_tensor(mag, u)
    s = oe.stress_tensor(mag, e)
    for i in range(3):
        want = 283e9 * e[i] + 58e9 * (e[(i + 1) % 3] + e[(i + 2) % 3])
        assert np.allclose(s[i], want, rtol=1e-12, atol=0)
        assert np.allclose(s[i + 3], 2 * 166e9 * e[i + 3], rtol=1e-12, atol=0)
    c = coords(mag)
    u = np.array([1e-3 * c[0], 1e-3 * c[1], 1e-3 * c[2]])
    s = oe.stress_tensor(mag, oe.strain_tensor(mag, u))
    # one cell thick: no z derivative, so e_zz = 0 and the in-plane expansion gives
    # sigma_xx = sigma_yy = (C11 + C12) e, sigma_zz = 2 C12 e
    assert np.allclose(s[:2], (283e9 + 58e9) * 1e-3, rtol=1e-9)
    assert np.allclose(s[2], 2 * 58e9 * 1e-3, rtol=1e-9)
    assert np.allclose(s[3:], 0.0, atol=1e-3)


C11, C12, C44 = 283e9, 166e9, 58e9


@pytest.mark.parametrize("d,uc,ckey", [(0, 0, "C11"), (1, 1, "C11"), (2, 2, "C11"), (1, 0, "C44"), (2, 0, "C44"),
                                       (0, 1, "C44"), (2, 1, "C44"), (0, 2, "C44"), (1, 2, "C44")])
def test_elastic_force_of_sine(d, uc, ckey):
    """test/test_elasticforce_uniform_c.py:45-137: f = -k^2 C u."""
    mag = line_magnet(d)
    if ckey == "C11":
        mag.set("C11", C11)
        C = C11
    else:
        mag.set("C44", C44)
        mag.set("C12", -C44)  # removes the mixed derivative
        C = C44
    k = P * 2 * math.pi / (N * CS[d])
    u = np.zeros((3,) + mag.shape)
    u[uc] = 1e-15 * np.sin(k * coords(mag)[d])
    f = oe.elastic_force(mag, oe.stress_tensor(mag, oe.strain_tensor(mag, u)))
    assert semirel(f, -k * k * C * u) < SRTOL


@pytest.mark.parametrize("outer,inner", [(0, 1), (1, 0), (0, 2), (2, 0), (1, 2), (2, 1)])
def test_elastic_force_mixed_derivative(outer, inner):
    """test/test_elasticforce_uniform_c.py:143-205: f_inner = C12 d_inner d_outer u_outer."""
    n1, n2, p1, p2 = 128, 256, 2, 3
    mag = line_magnet(outer, n1, second=(inner, n2))
    mag.set("C12", C12)
    ko = p1 * 2 * math.pi / (n1 * CS[outer])
    ki = p2 * 2 * math.pi / (n2 * CS[inner])
    c = coords(mag)
    u = np.zeros((3,) + mag.shape)
    u[outer] = 1e-15 * np.sin(ko * c[outer]) * np.sin(ki * c[inner])
    f = oe.elastic_force(mag, oe.stress_tensor(mag, oe.strain_tensor(mag, u)))
    want = np.zeros_like(f)
    want[inner] = 1e-15 * C12 * ko * ki * np.cos(ki * c[inner]) * np.cos(ko * c[outer])
    assert semirel(f, want) < SRTOL


def test_fourth_order_interior_and_traction_free_edges():
    """Open 1-D bar: the interior stencil error of d^2/dx^2 sin falls by 16x
    when the grid is refined twice (4th order); at the free ends the force is
    the one-sided traction-free difference of elasticforce.cu:69-82."""
    errs = []
    for n in (32, 64):
        size, mg, reps = (n, 1, 1), (n, 0, 0), (1, 0, 0)
        w = om.World(CS, mastergrid=mg, pbc=reps)
        mag = w.add_ferromagnet(size, defer_timestep=True)
        mag.set("C11", 1.0)
        k = 2 * math.pi / (n * CS[0])
        u = np.zeros((3,) + mag.shape)
        u[0] = np.sin(k * coords(mag)[0])
        f = oe.elastic_force(mag, oe.stress_tensor(mag, oe.strain_tensor(mag, u)))
        errs.append(np.abs(f[0] + k * k * u[0]).max() / (k * k))
    assert math.log2(errs[0] / errs[1]) == pytest.approx(4.0, abs=0.3)
    w = om.World(CS)
    mag = w.add_ferromagnet((8, 1, 1), defer_timestep=True)
    mag.set("C11", 2.0)
    u = np.zeros((3,) + mag.shape)
    u[0] = 0.3 * coords(mag)[0]  # uniform strain 0.3: stress 0.6 inside, zero traction outside
    f = oe.elastic_force(mag, oe.stress_tensor(mag, oe.strain_tensor(mag, u)))
    w0 = 1 / CS[0]
    assert f[0, 0, 0, 0] == pytest.approx(w0 * (0.6 + 0.6 / 3))
    assert f[0, 0, 0, -1] == pytest.approx(w0 * (-0.6 / 3 - 0.6))
    assert np.allclose(f[0, 0, 0, 1:-1], 0.0, atol=1e-6 * w0)


@pytest.mark.parametrize("mc,i,j,b1,b2", [(0, 0, 0, -8.8e6, 0), (1, 1, 1, -8.8e6, 0), (2, 2, 2, -8.8e6, 0),
                                          (1, 0, 1, 0, -8.8e6), (2, 0, 2, 0, -8.8e6), (2, 1, 2, 0, -8.8e6),
                                          (0, 1, 0, 0, -8.8e6)])
def test_magnetoelastic_field_of_sine_strain(mc, i, j, b1, b2):
    """test/test_magnetoelasticfield.py:30-95."""
    mag = line_magnet(i)
    ms = 800e3
    mag.set("msat", ms)
    mag.set("B1", b1)
    mag.set("B2", b2)
    m = np.zeros((3,) + mag.shape)
    m[mc] = 1.0
    k = P * 2 * math.pi / (N * CS[i])
    u = np.zeros((3,) + mag.shape)
    u[j] = AMP * np.sin(k * coords(mag)[i])
    e = np.zeros((6,) + mag.shape)
    cos = AMP * np.cos(k * coords(mag)[i])
    if i == j:
        e[j] = k * cos
    else:
        e[2 + i + j] = 0.5 * k * cos
    want = np.zeros_like(m)
    for a in range(3):
        a1, a2 = (a + 1) % 3, (a + 2) % 3
        want[a] = -2 / ms * (b1 * e[a] * m[a] + b2 * (e[a + a1 + 2] * m[a1] + e[a + a2 + 2] * m[a2]))
    got = oe.magnetoelastic_field_dynamic(mag, m, oe.strain_tensor(mag, u))
    assert semirel(got, want) < SRTOL


@pytest.mark.parametrize("d,b1,b2", [(0, -8.8e6, 0), (1, -8.8e6, 0), (2, -8.8e6, 0), (0, 0, -8.8e6),
                                     (1, 0, -8.8e6), (2, 0, -8.8e6)])
def test_magnetoelastic_force_of_rotating_m(d, b1, b2):
    """test/test_magnetoelasticforce.py:33-60: m_d = sin(kx), m_c = cos(kx)."""
    mag = line_magnet(d)
    mag.set("B1", b1)
    mag.set("B2", b2)
    c = (d + 1) % 3
    k = P * 2 * math.pi / (N * CS[d])
    m = np.zeros((3,) + mag.shape)
    m[d] = np.sin(k * coords(mag)[d])
    m[c] = np.cos(k * coords(mag)[d])
    got = oe.magnetoelastic_force(mag, m)
    want = np.zeros_like(m)
    want[d] = 2 * b1 * k * m[d] * m[c]
    want[c] = b2 * k * (m[c] ** 2 - m[d] ** 2)
    assert semirel(got, want) < SRTOL


def test_rigid_strain_field_matches_dynamic_formula():
    """A rigid strain equal to a uniform dynamic strain gives the same field
    (magnetoelasticfield.cu:72-103 vs :44-70)."""
    rng = np.random.default_rng(1)
    w = om.World(CS)
    mag = w.add_ferromagnet((6, 5, 2), defer_timestep=True)
    mag.set("msat", 5e5)
    mag.set("B1", -3e6)
    mag.set("B2", 7e6)
    m = rng.normal(size=(3,) + mag.shape)
    m /= np.linalg.norm(m, axis=0)
    n, s = np.array([1e-3, -2e-3, 5e-4]), np.array([3e-4, -1e-4, 2e-4])
    e = np.zeros((6,) + mag.shape)
    e[:3] = n[:, None, None, None]
    e[3:] = s[:, None, None, None]
    dyn = oe.magnetoelastic_field_dynamic(mag, m, e)
    rig = oe.magnetoelastic_field_rigid(mag, m, np.broadcast_to(n[:, None, None, None], m.shape),
                                        np.broadcast_to(s[:, None, None, None], m.shape))
    assert np.allclose(dyn, rig, rtol=1e-13, atol=0)
