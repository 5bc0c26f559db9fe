"""Pins of the oracle's time-dependent parameter terms (oracle/model.py,
DynamicParameter of src/core/dynamic_parameter.cpp:8-29).

* a single spin without damping in a field b(t) z precesses by the phase
  gamma * integral(b): b(t) = B1 sin(w t) from a term, on top of a static
  B0, gives phi(T) = gamma (B0 T + B1 (1 - cos w T) / w).  A term read at the
  step start instead of at each stage time misses this by
  gamma dt B1 |sin(w T)| / 2, about 2e-3 rad here;
* masked terms: two uncoupled cells with mask (1, 0.5) precess with their
  own phases; a term's mask of shape (nz, ny, nx) for a vector parameter
  scales every component;
* a time term makes the parameter not assured zero (zeeman.cu:38-51), so
  the Zeeman term enters with a zero static bias;
* terms add up, and removing them restores the static value.
"""

import numpy as np
import pytest

from oracle import model as om

GAMMALL = 1.7595e11
B0, B1, W = 0.1, 0.05, 2 * np.pi * 10e9


def _spin(n=1):
    world = om.World((2e-9, 2e-9, 2e-9))
    mag = world.add_ferromagnet((n, 1, 1), defer_timestep=True)
    mag.enable_demag = False
    mag.set("msat", 8e5)
    mag.set("alpha", 0.0)
    m = np.zeros((3, 1, 1, n))
    m[0] = 1.0
    mag.set_m(m)
    s = world.solver
    s.set_method("DormandPrince", defer_timestep=True)
    s.adaptive = False
    s.dt = 5e-13
    return world, mag


def _phase(mag, i=0):
    return np.arctan2(mag.m[1, 0, 0, i], mag.m[0, 0, 0, i])


def _wrap(a):
    return (a + np.pi) % (2 * np.pi) - np.pi


def test_single_spin_phase_follows_the_integral_of_the_term():
    world, mag = _spin()
    mag.set("bias", (0.0, 0.0, B0))
    mag.add_time_term("bias", lambda t: (0.0, 0.0, B1 * np.sin(W * t)))
    world.solver.steps(350)
    T = world.solver.time
    want = GAMMALL * (B0 * T + B1 * (1 - np.cos(W * T)) / W)
    assert abs(_wrap(_phase(mag) - want)) < 1e-7
    assert abs(mag.m[2, 0, 0, 0]) < 1e-12


def test_masked_terms_scale_per_cell():
    world, mag = _spin(2)
    mask = np.array([[[1.0, 0.5]]])  # (nz, ny, nx), broadcast over the components
    mag.add_time_term("bias", lambda t: (0.0, 0.0, B0 + B1 * np.sin(W * t)), mask=np.broadcast_to(mask, (3, 1, 1, 2)))
    world.solver.steps(350)
    T = world.solver.time
    integral = B0 * T + B1 * (1 - np.cos(W * T)) / W
    for i, k in enumerate((1.0, 0.5)):
        assert abs(_wrap(_phase(mag, i) - GAMMALL * k * integral)) < 1e-7


def test_terms_add_up_enter_zeeman_and_can_be_removed():
    world, mag = _spin()
    assert mag.assured_zero("bias")
    mag.add_time_term("bias", lambda t: (0.0, 0.0, 1.0))
    mag.add_time_term("bias", lambda t: (0.0, 2.0 * t, 0.0), mask=np.full((3, 1, 1, 1), 3.0))
    assert not mag.assured_zero("bias")
    world.solver.time = 0.25
    np.testing.assert_allclose(mag.vec("bias")[:, 0, 0, 0], [0.0, 1.5, 1.0])
    mag.add_time_term("alpha", lambda t: 0.5 * t)
    assert mag.val("alpha")[0, 0, 0] == pytest.approx(0.125)
    mag.remove_time_terms("bias")
    assert mag.assured_zero("bias")
    assert mag.vec("bias")[:, 0, 0, 0].tolist() == [0.0, 0.0, 0.0]
