"""Pins of the oracle's local effective-field terms (oracle/fields.py)."""

import numpy as np
import pytest
import scipy.optimize

from oracle import fields as of
from oracle import model as om


def _magnet(size, cs, **params):
    w = om.World(cs)
    m = w.add_ferromagnet(size, defer_timestep=True)
    m.enable_demag = False
    for k, v in params.items():
        m.set(k, v)
    return w, m


def _rand(rng, shape):
    v = rng.normal(size=(3,) + shape)
    return v / np.linalg.norm(v, axis=0)


def test_exchange_reference_numpy_formulation():
    """test/test_exchange.py:6-46 (Neumann boundary, no DMI)."""
    rng = np.random.default_rng(0)
    w, mag = _magnet((16, 16, 4), (1e3, 2e3, 3e3), aex=3.2e7, msat=5.4)
    m = _rand(rng, (4, 16, 16))
    got = of.exchange_field(mag, m)
    cs = mag.cellsize
    h = np.zeros(m.shape)
    for axis, c in ((1, cs[2]), (2, cs[1]), (3, cs[0])):
        for shift, keep in ((1, slice(1, None)), (-1, slice(0, -1))):
            m_ = np.roll(m, shift, axis=axis)
            sl = [slice(None)] * 4
            sl[axis] = keep
            h[tuple(sl)] += (m_ - m)[tuple(sl)] / c ** 2
    want = 2 * 3.2e7 * h / 5.4
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


def test_interfacial_dmi_reference_numpy_formulation():
    """test/test_interfacialdmi.py:8-56 (open boundaries)."""
    rng = np.random.default_rng(1)
    w, mag = _magnet((4, 5, 3), (1.1, 2.0, 3.2), msat=3.1, xxz=7.1, yyz=-7.1)
    mag.openbc = True
    m = _rand(rng, (3, 5, 4))
    got = of.dmi_field(mag, m)
    cs = mag.cellsize
    h = np.zeros(m.shape)
    X, Y, Z = 0, 1, 2
    Dx, Dy = 7.1, -7.1
    m_ = np.roll(m, 1, axis=3)
    h[X][:, :, 1:] -= Dx * m_[Z][:, :, 1:] / cs[0]
    h[Z][:, :, 1:] += Dx * m_[X][:, :, 1:] / cs[0]
    m_ = np.roll(m, -1, axis=3)
    h[X][:, :, :-1] += Dx * m_[Z][:, :, :-1] / cs[0]
    h[Z][:, :, :-1] -= Dx * m_[X][:, :, :-1] / cs[0]
    m_ = np.roll(m, 1, axis=2)
    h[Y][:, 1:, :] -= Dy * m_[Z][:, 1:, :] / cs[1]
    h[Z][:, 1:, :] += Dy * m_[Y][:, 1:, :] / cs[1]
    m_ = np.roll(m, -1, axis=2)
    h[Y][:, :-1, :] += Dy * m_[Z][:, :-1, :] / cs[1]
    h[Z][:, :-1, :] -= Dy * m_[Y][:, :-1, :] / cs[1]
    want = h / 3.1
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


def _energy(mag, m, term):
    if term == "exchange":
        e = of.energy_density(mag, m, of.exchange_field(mag, m), 0.5)
    elif term == "dmi":
        e = of.energy_density(mag, m, of.dmi_field(mag, m), 0.5)
    else:
        e = of.anisotropy_energy_density(mag, m)
    return e.sum() * np.prod(mag.cellsize)


@pytest.mark.parametrize("term", ["exchange", "dmi", "anisotropy"])
def test_field_is_energy_gradient(term):
    """h = -(1/(Ms V)) dE/dm for the bulk (periodic) discretisation: pins
    signs, factors of two and the harmonic mean of non-uniform A."""
    rng = np.random.default_rng(2)
    size = (6, 5, 4)
    cs = (2e-9, 3e-9, 2.5e-9)
    w = om.World(cs, mastergrid=size, pbc=(1, 1, 1))
    mag = w.add_ferromagnet(size, defer_timestep=True)
    mag.enable_demag = False
    ms = 8e5
    mag.set("msat", ms)
    if term == "exchange":
        mag.set("aex", 1e-11 * (1 + rng.random((4, 5, 6))))
    elif term == "dmi":
        for k, v in (("xyz", 2e-3), ("yxz", -2e-3), ("zxy", 2e-3), ("xxz", 1e-3), ("yyz", 1e-3)):
            mag.set(k, v)
        mag.set("aex", 1e-11)
    else:
        mag.set("ku1", 5e5)
        mag.set("ku2", -1e5)
        mag.set("anisU", (0.3, 0.1, 1.0))
    m = _rand(rng, (4, 5, 6))
    h = {"exchange": of.exchange_field, "dmi": of.dmi_field, "anisotropy": of.anisotropy_field}[term](mag, m)
    V = np.prod(cs)
    eps = 1e-6
    for (c, k, j, i) in [(0, 1, 2, 3), (1, 0, 4, 0), (2, 3, 0, 5), (0, 2, 2, 2)]:
        mp, mm = m.copy(), m.copy()
        mp[c, k, j, i] += eps
        mm[c, k, j, i] -= eps
        dE = (_energy(mag, mp, term) - _energy(mag, mm, term)) / (2 * eps)
        want = -dE / (ms * V)
        if term == "anisotropy":
            assert h[c, k, j, i] == pytest.approx(want, rel=1e-6)
        else:
            # quadratic energy: the energy density -1/2 Ms m.h gives dE/dm = -Ms V h
            assert h[c, k, j, i] == pytest.approx(want, rel=1e-6, abs=1e-9 * np.abs(h).max())


def test_exchange_spiral_energy():
    """test/test_exchange.py:48-85: discrete spiral energy vs A k^2 (< 1e-2)."""
    msat, aex = 800e3, 5e-12
    n = 1000
    c = 1e-9
    w, mag = _magnet((n, 1, 1), (c, c, c), msat=msat, aex=aex)
    x = np.arange(n) * c
    for k in (1e7, 5e7, 1e8):
        m = np.zeros((3, 1, 1, n))
        m[0, 0, 0] = np.cos(k * x)
        m[1, 0, 0] = np.sin(k * x)
        e = of.energy_density(mag, m, of.exchange_field(mag, m), 0.5).mean()
        assert e == pytest.approx(aex * k * k, rel=1e-2)


def test_anisotropy_closed_form_and_llg_orthogonal():
    rng = np.random.default_rng(3)
    w, mag = _magnet((3, 2, 2), (1e-9, 1e-9, 1e-9), msat=1e6, ku1=4e5, ku2=2e5, anisU=(0.0, 0.0, 2.0), alpha=0.3)
    m = _rand(rng, (2, 2, 3))
    h = of.anisotropy_field(mag, m)
    mz = m[2]
    assert np.allclose(h[2], (2 * 4e5 * mz + 4 * 2e5 * mz ** 3) / 1e6)
    assert np.allclose(h[:2], 0.0)
    t = of.llg_torque(mag, m, h)
    assert np.abs((t * m).sum(axis=0)).max() < 1e-6 * np.abs(t).max()


# ---------------------------------------------------------------- canting
# test/test_boundaries.py:17-20: mumax3 reference values of the edge canting
# of a 1-D wire with interfacial DMI at 90 % of the critical strength.
DMI = 0.9 * 4 / np.pi
CANT = {"neumann": 0.44716486657643906, "open": 0.5958252040658198, "periodic": 0.5958256832017506}


def _canting(kind, axis):
    ncell, cs = 1024, 0.05
    size = (ncell, 1, 1) if axis == 0 else (1, ncell, 1)
    if kind == "periodic":
        mg = (0, 1, 0) if axis == 0 else (1, 0, 0)
        w = om.World((cs, cs, cs), mastergrid=mg, pbc=tuple(int(v > 0) for v in mg))
    else:
        w = om.World((cs, cs, cs))
    mag = w.add_ferromagnet(size, defer_timestep=True)
    mag.enable_demag = False
    mag.openbc = kind == "open"
    for k, v in (("xxz", DMI), ("yyz", DMI), ("aex", 1.0), ("ku1", 1.0), ("msat", 1.0), ("anisU", (0, 0, 1))):
        mag.set(k, v)
    shape = mag.shape

    def m_of(th):
        m = np.zeros((3,) + shape)
        s = np.sin(th).reshape(shape)
        m[axis] = s
        m[2] = np.cos(th).reshape(shape)
        return m

    def torque(th):
        m = m_of(th)
        h = of.exchange_field(mag, m) + of.dmi_field(mag, m) + of.anisotropy_field(mag, m)
        t = of.cross(m, h)
        # component perpendicular to the rotation plane
        return t[1 if axis == 0 else 0].ravel()

    x = np.arange(ncell) * cs
    th0 = 0.5 * (np.exp(-x) + np.exp(-(x[-1] - x)))
    if axis == 1:
        th0 = -th0
    # the torque couples nearest neighbours only: tridiagonal Jacobian
    sol = scipy.optimize.root(torque, th0, method="hybr", options={"xtol": 1e-13, "band": (1, 1)})
    assert sol.success, sol.message
    m = m_of(sol.x)
    return abs(np.arctan2(m[axis].ravel()[0], m[2].ravel()[0]))


@pytest.mark.parametrize("kind", ["neumann", "open", "periodic"])
@pytest.mark.parametrize("axis", [0, 1])
def test_edge_canting_golden(kind, axis):
    got = _canting(kind, axis)
    assert abs(got - CANT[kind]) / CANT[kind] < 2e-4
    if kind != "neumann":
        # test/test_boundaries.py:52-55,97-140: open and periodic are compared
        # with the analytic 1-D canting shifted to the cell centre
        theta0 = np.arcsin(DMI / 2)
        want = 2 * np.arctan(np.exp(-0.05 / 2) * np.tan(theta0 / 2))
        assert abs(got - want) / want < 2e-4
