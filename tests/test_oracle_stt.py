"""Pins of the oracle's spin-transfer torques (oracle/fields.py)."""

import numpy as np
import pytest

from oracle import fields as of
from oracle import model as omodel


def _magnet(size, cs=(2e-9, 3e-9, 1e-9), geometry=None):
    ow = omodel.World(cs)
    return ow, ow.add_ferromagnet(size, geometry=geometry)


@pytest.mark.parametrize("k", [0.3, 1.1])
def test_zhang_li_on_a_spiral_is_the_closed_form(k):
    """m = (cos kx, sin kx, 0) with a current along x: the central difference
    gives h = (u_x sin(k c) / c) z x m, so m x h = (u_x sin(kc)/c) z and
    m x (m x h) = -h (periodic wrap, no ends)."""
    nx, cs = 40, 2e-9
    ow = omodel.World((cs, cs, cs), mastergrid=(nx, 0, 0))
    mag = ow.add_ferromagnet((nx, 1, 1))
    kk = 2 * np.pi * 3 / (nx * cs)  # an integer number of turns for the wrap
    if k > 1:
        kk *= 2
    x = np.arange(nx) * cs
    m = np.stack([np.cos(kk * x), np.sin(kk * x), 0 * x])[:, None, None, :]
    ms, pol, xi, al, jx = 8e5, 0.6, 0.05, 0.02, 3e11
    for key, v in {"msat": ms, "pol": pol, "xi": xi, "alpha": al, "jcur": (jx, 0.0, 0.0)}.items():
        mag.set(key, v)
    t = of.zhang_li_torque(mag, m)
    u = of.MUB * pol / (of.QE * ms * (1 + xi * xi)) * jx
    a = u * np.sin(kk * cs) / cs
    h = a * np.stack([-np.sin(kk * x), np.cos(kk * x), 0 * x])
    zhat = np.array([0.0, 0.0, 1.0])[:, None]
    want = (-1 / (1 + al * al)) * ((1 + xi * al) * (-h) + (xi - al) * a * zhat)
    assert np.allclose(t[:, 0, 0, :], want, rtol=1e-10, atol=1e-10 * abs(a))


def test_zhang_li_is_zero_for_uniform_m_and_perpendicular_to_m():
    rng = np.random.default_rng(0)
    ow, mag = _magnet((6, 5, 3))
    for key, v in {"msat": 6e5, "pol": 0.5, "xi": 0.1, "alpha": 0.05, "jcur": (1e12, -2e11, 5e11)}.items():
        mag.set(key, v)
    mag.set_m((1.0, 2.0, 0.5))
    # uniform m: the difference is zero inside and parallel to m at the ends
    scale = of.MUB * 0.5 / (of.QE * 6e5) * 1e12 / 1e-9
    assert np.abs(of.zhang_li_torque(mag, mag.m)).max() <= 1e-9 * scale
    m = rng.normal(size=(3, 3, 5, 6))
    m /= np.linalg.norm(m, axis=0)
    t = of.zhang_li_torque(mag, m)
    assert np.abs(np.sum(t * m, axis=0)).max() < 1e-9 * np.abs(t).max()


def test_slonczewski_free_layer_thickness_and_empty_layers():
    """test/test_free_layer_thickness.py: empty layers added to the grid do not
    change the torque once the free layer thickness is set; the torque is
    perpendicular to m and zero for m along the fixed layer."""
    m0 = np.array([0.3, 0.5, 0.81])
    m0 /= np.linalg.norm(m0)
    vals = {"msat": 4.3, "jcur": (0, 0, 5.4), "pol": 0.6, "fixed_layer": (1, 0, 1),
            "free_layer_thickness": 1e-9, "Lambda": 1.0}
    ow, a = _magnet((1, 1, 2), cs=(1e-9, 1e-9, 1e-9))
    geom = np.array([[[True]], [[True]], [[False]], [[False]]])
    ow2, b = _magnet((1, 1, 4), cs=(1e-9, 1e-9, 1e-9), geometry=geom)
    for mag in (a, b):
        for key, v in vals.items():
            mag.set(key, v)
        mag.set_m(m0)
    ta = of.slonczewski_torque(a, a.m)
    tb = of.slonczewski_torque(b, b.m)
    assert np.allclose(ta[:, 0, 0, 0], tb[:, 0, 0, 0], rtol=1e-12)
    assert abs(np.dot(ta[:, 0, 0, 0], m0)) < 1e-12 * np.linalg.norm(ta[:, 0, 0, 0])
    p = np.array([1.0, 0.0, 1.0]) / np.sqrt(2)
    a.set_m(p)
    assert np.abs(of.slonczewski_torque(a, a.m)).max() < 1e-12 * np.linalg.norm(ta[:, 0, 0, 0])


def test_slonczewski_fixed_layer_position_flips_the_sign():
    ow, mag = _magnet((2, 2, 1))
    for key, v in {"msat": 8e5, "jcur": (0, 0, -1e11), "pol": 0.56, "Lambda": 2.0, "epsilon_prime": 1.0,
                   "fixed_layer": (0.94, 0.34, 0.0), "alpha": 0.01}.items():
        mag.set(key, v)
    mag.set_m((1.0, 0.0, 0.2))
    top = of.slonczewski_torque(mag, mag.m)
    mag.fixed_on_top = False
    bottom = of.slonczewski_torque(mag, mag.m)
    assert np.allclose(top, -bottom, rtol=1e-12)


def test_torque_choice_and_switches():
    ow, mag = _magnet((4, 4, 1))
    for key, v in {"msat": 8e5, "jcur": (1e11, 0, 1e11), "pol": 0.5, "Lambda": 1.0, "fixed_layer": (0, 0, 1),
                   "xi": 0.1}.items():
        mag.set(key, v)
    rng = np.random.default_rng(1)
    m = rng.normal(size=(3, 1, 4, 4))
    m /= np.linalg.norm(m, axis=0)
    assert np.allclose(of.spin_transfer_torque(mag, m), of.slonczewski_torque(mag, m))
    mag.enable_slonc = False
    assert np.allclose(of.spin_transfer_torque(mag, m), of.zhang_li_torque(mag, m))
    mag.enable_zhangli = False
    assert not np.any(of.spin_transfer_torque(mag, m))
