"""Pins of the oracle's Runge-Kutta stepper and time solver (oracle/rk.py)."""

import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import model as om
from oracle import rk

GAMMALL = 1.7595e11


class LinearSystem:
    """y' = B y for a 3-vector (a rotation plus decay), closed-form solution."""

    def __init__(self):
        w, g = 2.0, 0.3
        self.B = np.array([[-g, -w, 0.0], [w, -g, 0.0], [0.0, 0.0, -2 * g]])
        self.y = np.array([1.0, 0.2, 0.5]).reshape(3, 1, 1, 1)
        self.geom = np.ones((1, 1, 1), dtype=bool)

    def get_states(self):
        return [self.y]

    def set_states(self, s):
        self.y = s[0]

    def rhs(self, states):
        return [np.tensordot(self.B, states[0], axes=(1, 0))]

    def geom_masks(self):
        return [self.geom]

    def exact(self, t):
        from scipy.linalg import expm
        return expm(self.B * t) @ np.array([1.0, 0.2, 0.5])


def test_tableau_rationals_consistent():
    for name, tab in rk.TABLEAUS.items():
        for i, row in enumerate(tab["A"]):
            assert sum(row) == pytest.approx(tab["nodes"][i], abs=1e-14), name
        assert sum(tab["w1"]) == pytest.approx(1.0, abs=1e-14)
        assert sum(tab["w2"]) == pytest.approx(1.0, abs=1e-14)


def test_dormand_prince_published_coefficients():
    """A few Dormand & Prince (1980) rationals, checked as exact fractions."""
    tab = rk.TABLEAUS["DormandPrince"]
    assert Fraction(tab["A"][4][1]).limit_denominator(10000) == Fraction(-25360, 2187)
    assert Fraction(tab["w2"][4]).limit_denominator(1000000) == Fraction(-92097, 339200)
    # first same as last: the last stage is the 5th-order solution
    assert tab["A"][6] == tab["w1"][:6]
    assert tab["w1"][6] == 0.0 and tab["nodes"][6] == 1.0


@pytest.mark.parametrize("method,order", [("Heun", 2), ("BogackiShampine", 3), ("Fehlberg", 5),
                                          ("CashKarp", 5), ("DormandPrince", 5)])
def test_convergence_order(method, order):
    errs = []
    T = 1.0
    for n in (20, 40):
        sys_ = LinearSystem()
        ts = rk.TimeSolver(sys_)
        ts.method = method
        ts.adaptive = False
        ts.dt = T / n
        for _ in range(n):
            ts.step()
        errs.append(np.abs(sys_.y.ravel() - sys_.exact(T)).max())
    p = math.log(errs[0] / errs[1], 2)
    assert p == pytest.approx(order, abs=0.35)


def test_embedded_error_estimates_lower_order():
    """dt * sum (w1 - w2) k scales like dt^(order2 + 1)."""
    for method in ("Fehlberg", "DormandPrince", "CashKarp"):
        tab = rk.TABLEAUS[method]
        est = []
        for dt in (0.1, 0.05):
            sys_ = LinearSystem()
            k = []
            y0 = sys_.y.copy()
            for s in range(len(tab["nodes"])):
                y = y0 + sum(dt * tab["A"][s][j] * k[j] for j in range(s))
                k.append(sys_.rhs([y])[0])
            e = sum(dt * (tab["w1"][j] - tab["w2"][j]) * k[j] for j in range(len(k)))
            est.append(np.abs(e).max())
        assert math.log(est[0] / est[1], 2) == pytest.approx(tab["order2"] + 1, abs=0.3)


def test_controller_rules():
    sys_ = LinearSystem()
    ts = rk.TimeSolver(sys_)
    ts.dt = 1.0
    ts.adapt(float("nan"))
    assert ts.dt == pytest.approx(0.8)  # NaN -> 1, times headroom
    ts.adapt(100.0)
    assert ts.dt == pytest.approx(1.6)  # clamped to the upper bound 2
    ts.adapt(1e-6)
    assert ts.dt == pytest.approx(0.8)  # clamped to the lower bound 0.5
    ts.lower_bound = 3.0
    with pytest.raises(RuntimeError):
        ts.adapt(1.0)


def test_rejections_shrink_the_step():
    sys_ = LinearSystem()
    ts = rk.TimeSolver(sys_)
    ts.method = "DormandPrince"
    ts.dt = 0.5
    ts.max_error = 1e-10
    ts.step()
    assert ts.nrejected >= 1
    assert ts.dt < 0.5


def precession(t, hz, damping):
    freq = GAMMALL * hz / (1 + damping ** 2)
    phi = freq * t
    theta = np.pi - 2 * np.arctan(np.exp(damping * freq * t) * np.tan(np.pi / 4))
    return np.array([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)])


@pytest.mark.parametrize("method", rk.METHOD_NAMES)
def test_single_spin_precession(method):
    """test/test_timesolver.py:84-124 (max error < 1e-4, adaptive steps)."""
    w = om.World((1, 1, 1))
    mag = w.add_ferromagnet((1, 1, 1))
    mag.enable_demag = False
    mag.set_m((1, 0, 0))
    mag.set("alpha", 0.2)
    w.bias = (0, 0, 0.11)
    w.solver.set_method(method)
    w.solver.assure_sensible_timestep()
    err = 0.0
    for tp in np.linspace(0, 5e-10, 20):
        w.solver.run(tp - w.solver.time, assure=False)
        err = max(err, np.abs(mag.m[:, 0, 0, 0] - precession(w.solver.time, 0.11, 0.2)).max())
    assert err < 1e-4


def test_run_lands_on_stop_time():
    w = om.World((1, 1, 1))
    mag = w.add_ferromagnet((1, 1, 1))
    mag.enable_demag = False
    mag.set_m((1, 0, 0))
    w.bias = (0, 0, 0.1)
    w.solver.set_method("DormandPrince")
    w.solver.run(3.3e-11)
    assert w.solver.time == pytest.approx(3.3e-11, rel=1e-12)
