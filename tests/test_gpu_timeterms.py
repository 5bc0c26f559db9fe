"""Time-dependent parameter terms (Parameter.add_time_term, parameter.py:61-113;
DynamicParameter, dynamic_parameter.cpp:8-29) on the GPU.

* a single undamped spin in b(t) z precesses by gamma * integral(b) (the
  closed form the oracle is pinned to), which needs the term at every stage
  time;
* fixed-step Dormand-Prince with a masked bias term, an unmasked term on a
  per-cell alpha and a vector term, against the FP64 oracle: max |dm| 1e-4,
  <m> 1e-5 (the path's bar); an adaptive run against the oracle;
* Parameter.eval / external_field at the solver time: static + sum of the
  terms, also with more masked terms than one fill launch takes;
* is_dynamic / is_uniform / uniform_value, set() dropping earlier terms,
  set(callable) and set((term, mask)), remove_time_terms;
* a term that raises stops the step with its own exception; a batch refuses
  a magnet with time terms.
"""

import numpy as np
import pytest

from helpers import make_pair

import mumaxplus as mp

pytestmark = pytest.mark.gpu

GAMMALL = 1.7595e11
B0, B1, W = 0.1, 0.05, 2 * np.pi * 10e9


def _wrap(a):
    return (a + np.pi) % (2 * np.pi) - np.pi


def test_single_spin_phase_follows_the_integral_of_the_term():
    world = mp.World((2e-9, 2e-9, 2e-9))
    mag = mp.Ferromagnet(world, mp.Grid((1, 1, 1)))
    mag.enable_demag = False
    mag.msat, mag.alpha = 8e5, 0.0
    mag.magnetization = (1.0, 0.0, 0.0)
    mag.bias_magnetic_field = (0.0, 0.0, B0)
    mag.bias_magnetic_field.add_time_term(lambda t: (0.0, 0.0, B1 * np.sin(W * t)))
    ts = world.timesolver
    ts.set_method("DormandPrince")
    ts.adaptive_timestep = False
    ts.timestep = 5e-13
    ts.steps(350)
    T = ts.time
    m = mag.magnetization.get()[:, 0, 0, 0]
    want = GAMMALL * (B0 * T + B1 * (1 - np.cos(W * T)) / W)
    # a term read only at step starts is off by ~2e-3 rad here
    assert abs(_wrap(np.arctan2(m[1], m[0]) - want)) < 2e-4


def _pair(rng):
    size, cs = (16, 8, 2), (2e-9, 2e-9, 2e-9)
    alpha = 0.02 + 0.03 * rng.random((2, 8, 16))
    params = {"msat": 8e5, "aex": 1.3e-11, "alpha": alpha, "ku1": 2e5, "anisU": (0.0, 0.0, 1.0)}
    m0 = np.array([1.0, 0.2, 0.1])[:, None, None, None] + 0.05 * rng.normal(size=(3, 2, 8, 16))
    world, mag, ow, om = make_pair(size, cs, params, m0 / np.linalg.norm(m0, axis=0))
    x = np.linspace(0.0, 1.0, 16)
    mask = np.broadcast_to(x[None, None, :] ** 2, (2, 8, 16)).astype(np.float32)

    def bias(t):
        return (0.02 * np.cos(W * t), 0.0, 0.05 * np.sin(W * t))

    def alpha_term(t):
        return 0.01 * (1 + np.sin(0.5 * W * t))

    def push(t):
        return (0.0, 0.01, 0.0)

    mag.bias_magnetic_field.add_time_term(bias, mask)
    om.add_time_term("bias", bias, mask=np.broadcast_to(mask, (3, 2, 8, 16)).astype(np.float64))
    mag.bias_magnetic_field.add_time_term(push)
    om.add_time_term("bias", push)
    mag.alpha.add_time_term(alpha_term)
    om.add_time_term("alpha", alpha_term)
    return world, mag, ow, om


def test_fixed_steps_match_the_oracle():
    world, mag, ow, om = _pair(np.random.default_rng(11))
    ts = world.timesolver
    ts.set_method("DormandPrince")
    ts.adaptive_timestep = False
    ts.timestep = 1e-13
    ow.solver.set_method("DormandPrince", defer_timestep=True)
    ow.solver.adaptive = False
    ow.solver.dt = 1e-13
    ts.steps(150)
    ow.solver.steps(150)
    assert np.abs(mag.magnetization.get() - om.m).max() < 1e-4
    assert np.abs(np.array(mag.magnetization.average()) - om.m.reshape(3, -1).mean(axis=1)).max() < 1e-5


def test_adaptive_run_matches_the_oracle():
    world, mag, ow, om = _pair(np.random.default_rng(12))
    world.timesolver.set_method("DormandPrince")
    ow.solver.set_method("DormandPrince")
    world.timesolver.run(2e-11)
    ow.solver.run(2e-11)
    assert world.timesolver.time == pytest.approx(ow.solver.time, rel=1e-12)
    assert np.abs(np.array(mag.magnetization.average()) - om.m.reshape(3, -1).mean(axis=1)).max() < 1e-4


def test_values_at_the_solver_time():
    rng = np.random.default_rng(5)
    world = mp.World((1e-9, 1e-9, 1e-9))
    world.bias_magnetic_field = (0.0, 0.0, 0.01)
    mag = mp.Ferromagnet(world, mp.Grid((9, 5, 3)))
    mag.enable_demag = False
    shape = (3, 5, 9)
    masks = [rng.random((3,) + shape) for _ in range(11)]  # more than one fill launch takes
    base = (0.1, -0.2, 0.3)
    mag.bias_magnetic_field = base
    for k, mk in enumerate(masks):
        mag.bias_magnetic_field.add_time_term(lambda t, k=k: (k * t, 1.0, -t), mk)
    mag.bias_magnetic_field.add_time_term(lambda t: (2.0, 0.0, t * t))
    assert mag.bias_magnetic_field.is_dynamic and not mag.bias_magnetic_field.is_uniform
    for t in (0.0, 0.5, 3.0):
        world.timesolver.time = t
        want = np.array(base)[:, None, None, None] + np.array([2.0, 0.0, t * t])[:, None, None, None]
        for k, mk in enumerate(masks):
            want = want + np.array([k * t, 1.0, -t])[:, None, None, None] * mk
        got = mag.bias_magnetic_field.eval()
        assert np.abs(got - want).max() <= 1e-5 * np.abs(want).max()
        ext = mag.external_field.eval()
        assert np.abs(ext - want - np.array([0, 0, 0.01])[:, None, None, None]).max() <= 1e-5 * np.abs(want).max()
    # a per-cell static value plus an unmasked term
    a0 = rng.random((3, 5, 9))
    mag.alpha = a0
    mag.alpha.add_time_term(lambda t: 0.25 * t)
    world.timesolver.time = 2.0
    assert np.abs(mag.alpha.eval()[0] - (a0 + 0.5)).max() < 1e-6
    mag.alpha.remove_time_terms()
    assert not mag.alpha.is_dynamic
    assert np.abs(mag.alpha.eval()[0] - a0).max() < 1e-7


def test_flags_set_and_remove():
    world = mp.World((1e-9, 1e-9, 1e-9))
    mag = mp.Ferromagnet(world, mp.Grid((4, 4, 1)))
    p = mag.ku1
    assert not p.is_dynamic and p.is_uniform
    mag.ku1 = 3.0
    p.add_time_term(lambda t: t)
    assert p.is_dynamic and p.is_uniform and p.uniform_value == 3.0
    world.timesolver.time = 2.0
    assert np.allclose(p.eval(), 5.0)
    mag.ku1 = 7.0  # set() drops the terms (parameter.py:170-214)
    assert not p.is_dynamic and np.allclose(p.eval(), 7.0)
    mag.ku1 = lambda t: 10 * t  # static 0 + term
    assert p.is_dynamic and np.allclose(p.eval(), 20.0)
    mask = np.zeros((1, 4, 4))
    mask[0, 1, :] = 1.0
    mag.ku1 = (lambda t: 4.0, mask)
    assert p.is_dynamic and not p.is_uniform
    assert np.allclose(p.eval()[0], 4.0 * mask[0])
    # a mask given as a function of the position
    mag.ku1 = 1.0
    p.add_time_term(lambda t: 1.0, lambda x, y, z: 1.0 if x > 2e-9 else 0.0)
    assert np.allclose(p.eval()[0, 0], [1, 1, 1, 2])
    p.remove_time_terms()
    assert np.allclose(p.eval(), 1.0)
    with pytest.raises(ValueError):
        p.add_time_term(lambda t: 1.0, np.ones((2, 4, 4, 2)))


def test_failing_term_raises_its_exception_and_batch_refuses():
    world = mp.World((2e-9, 2e-9, 2e-9))
    mag = mp.Ferromagnet(world, mp.Grid((8, 8, 1)))
    mag.msat, mag.aex, mag.alpha = 8e5, 1.3e-11, 0.1
    world.timesolver.adaptive_timestep = False
    world.timesolver.timestep = 1e-13

    def bad(t):
        if t > 2.5e-13:
            raise ZeroDivisionError("term failed")
        return (0.0, 0.0, 0.1)

    mag.bias_magnetic_field = bad
    with pytest.raises(ZeroDivisionError):
        world.timesolver.steps(5)
    mag.bias_magnetic_field.remove_time_terms()
    world.timesolver.steps(2)
    b = mp.Batch(1024)
    w2 = mp.World((2e-9, 2e-9, 2e-9))
    m2 = mp.Ferromagnet(w2, mp.Grid((8, 8, 1)))
    m2.msat, m2.aex, m2.alpha = 8e5, 1.3e-11, 0.1
    m2.bias_magnetic_field = lambda t: (0.0, 0.0, 0.1)
    w2.timesolver.adaptive_timestep = False
    w2.timesolver.timestep = 1e-13
    with pytest.raises(ValueError):
        b.add(m2, steps=1)
