"""Explicit embedded Runge-Kutta stepping in float64 (test infrastructure only).

* ``TABLEAUS`` restates ButcherTableau::{Heun, BogackiShampine, Fehlberg,
  CashKarp, DormandPrince} (src/core/butchertableau.cpp:89-224): nodes,
  rkMatrix, weights1, weights2, order1, order2.
* ``TimeSolver.step`` restates RungeKuttaStepper::step and its stage executor
  (src/core/rungekutta.cpp:25-151): stage states x0 + dt*sum a_sj k_j,
  normalised on assignment; final state with weights1; error =
  max_cells |dt * sum (w1_j - w2_j) k_j|; accept if error < max_error;
  correction (max_error/error)^(1/(order2+1)) on success, ^(1/(order1+1))
  on failure; rollback on failure.
* ``TimeSolver.adapt`` restates TimeSolver::adaptTimeStep
  (src/core/timesolver.cpp:60-75): NaN -> 1, x headroom, clamp to
  [lower_bound, upper_bound].
* ``TimeSolver.sensible_timestep`` restates timesolver.cpp:37-47.
* ``steps``/``run`` restate the Python wrapper (mumaxplus/timesolver.py:61-112)
  together with TimeSolver::steps/run (timesolver.cpp:85-111): a sensible dt
  is assured first, and run() ends with one truncated step.

Time and dt are float64 here (the reference keeps them in FP32,
src/core/timesolver.hpp:79-80).

Pinned by (tests/test_oracle_rk.py): the published rational coefficients
(row sums = nodes, weights sum to one), DP FSAL identity (last row of the
matrix == weights1), convergence order of every method on a linear ODE with
closed-form solution, and single-spin damped precession
(test/test_timesolver.py:19-29) for all five methods with adaptive steps.
"""

import math

import numpy as np

TABLEAUS = {}


def _tab(name, nodes, A, w1, w2, o1, o2):
    TABLEAUS[name] = dict(nodes=[float(v) for v in nodes], A=[[float(v) for v in r] for r in A],
                          w1=[float(v) for v in w1], w2=[float(v) for v in w2], order1=o1, order2=o2)


_tab("Heun", [0.0, 1.0], [[], [1.0]], [0.5, 0.5], [1.0, 0.0], 2, 1)
_tab("BogackiShampine", [0.0, 1 / 2, 3 / 4, 1.0],
     [[], [1 / 2], [0.0, 3 / 4], [2 / 9, 1 / 3, 4 / 9]],
     [2 / 9, 1 / 3, 4 / 9, 0.0], [7 / 24, 1 / 4, 1 / 3, 1 / 8], 3, 2)
_tab("Fehlberg", [0.0, 1 / 4, 3 / 8, 12 / 13, 1.0, 1 / 2],
     [[], [1 / 4], [3 / 32, 9 / 32], [1932 / 2197, -7200 / 2197, 7296 / 2197],
      [439 / 216, -8.0, 3680 / 513, -845 / 4104],
      [-8 / 27, 2.0, -3544 / 2565, 1859 / 4104, -11 / 40]],
     [16 / 135, 0.0, 6656 / 12825, 28561 / 56430, -9 / 50, 2 / 55],
     [25 / 216, 0.0, 1408 / 2565, 2197 / 4104, -1 / 5, 0.0], 5, 4)
_tab("CashKarp", [0.0, 1 / 5, 3 / 10, 3 / 5, 1.0, 7 / 8],
     [[], [1 / 5], [3 / 40, 9 / 40], [3 / 10, -9 / 10, 6 / 5],
      [-11 / 54, 5 / 2, -70 / 27, 35 / 27],
      [1631 / 55296, 175 / 512, 575 / 13824, 44275 / 110592, 253 / 4096]],
     [37 / 378, 0.0, 250 / 621, 125 / 594, 0.0, 512 / 1771],
     [2825 / 27648, 0.0, 18575 / 48384, 13525 / 55296, 277 / 14336, 1 / 4], 5, 4)
_tab("DormandPrince", [0.0, 1 / 5, 3 / 10, 4 / 5, 8 / 9, 1.0, 1.0],
     [[], [1 / 5], [3 / 40, 9 / 40], [44 / 45, -56 / 15, 32 / 9],
      [19372 / 6561, -25360 / 2187, 64448 / 6561, -212 / 729],
      [9017 / 3168, -355 / 33, 46732 / 5247, 49 / 176, -5103 / 18656],
      [35 / 384, 0.0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84]],
     [35 / 384, 0.0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84, 0.0],
     [5179 / 57600, 0.0, 7571 / 16695, 393 / 640, -92097 / 339200, 187 / 2100, 1 / 40], 5, 4)

METHOD_NAMES = tuple(TABLEAUS)


def max_vec_norm(v, geom):
    n = np.sqrt(v[0] ** 2 + v[1] ** 2 + v[2] ** 2)
    n = n[geom]
    return float(n.max()) if n.size else 0.0


class TimeSolver:
    """Controller state and defaults of src/core/timesolver.hpp:74-82."""

    def __init__(self, system):
        self.system = system
        self.time = 0.0
        self.dt = 0.0
        self.adaptive = True
        self.max_error = 1e-5
        self.headroom = 0.8
        self.lower_bound = 0.5
        self.upper_bound = 2.0
        self.sensible_factor = 0.01
        self.sensible_default = 1e-14
        self.method = "Fehlberg"
        self.nsteps = 0
        self.nrejected = 0

    def equations_changed(self):
        if self.adaptive:
            self.dt = self.sensible_timestep()

    def set_method(self, name, defer_timestep=False):
        if name not in TABLEAUS:
            raise ValueError(f"'{name}' is not a valid Runge Kutta method name")
        self.method = name
        if self.adaptive and not defer_timestep:
            self.dt = self.sensible_timestep()

    def sensible_timestep(self):
        states = self.system.get_states()
        if not states:
            return 0.0
        ks = self.system.rhs(states)
        mx = max(max_vec_norm(k, g) for k, g in zip(ks, self.system.geom_masks()))
        if mx == 0:
            return self.sensible_default
        return self.sensible_factor / mx

    def adapt(self, corr):
        if not self.adaptive:
            return
        if math.isnan(corr):
            corr = 1.0
        corr *= self.headroom
        if self.lower_bound >= self.upper_bound:
            raise RuntimeError("The lower bound should be lower than the upper bound.")
        corr = min(corr, self.upper_bound)
        corr = max(corr, self.lower_bound)
        self.dt *= corr

    def step(self):
        if self.dt <= 0:
            raise RuntimeError("Timesolver can not make a step because the timestep is "
                               "smaller than or equal to zero.")
        tab = TABLEAUS[self.method]
        nodes, A, w1, w2 = tab["nodes"], tab["A"], tab["w1"], tab["w2"]
        S = len(nodes)
        x0 = [x.copy() for x in self.system.get_states()]
        if not x0:
            self.time += self.dt
            return
        geoms = self.system.geom_masks()
        t0 = self.time
        # one thermal draw per step for every magnetisation with a temperature,
        # kept through the stages and rejected attempts (rungekutta.cpp:85-93)
        noisy = False
        for kind, obj in getattr(self.system, "_equations", lambda: [])():
            if kind == "m" and not obj.assured_zero("temperature"):
                obj.step_draw = obj.noise_draws[0]
                obj.noise_draws[0] += 1
                noisy = True
        while True:
            dt = self.dt
            if noisy:
                self.system.noise_dt = dt
            k = []
            for s in range(S):
                self.time = t0 + dt * nodes[s]
                if s > 0:
                    xs = []
                    for e in range(len(x0)):
                        x = x0[e].copy()
                        for j in range(s):
                            x = x + (dt * A[s][j]) * k[j][e]
                        xs.append(x)
                    self.system.set_states(xs)
                k.append(self.system.rhs(self.system.get_states()))
            self.time = t0 + dt
            xf = []
            for e in range(len(x0)):
                x = x0[e].copy()
                for j in range(S):
                    x = x + (dt * w1[j]) * k[j][e]
                xf.append(x)
            self.system.set_states(xf)
            self.nsteps += 1
            if not self.adaptive:
                break
            err = 0.0
            for e in range(len(x0)):
                ev = np.zeros_like(x0[e])
                for j in range(S):
                    ev = ev + (dt * (w1[j] - w2[j])) * k[j][e]
                err = max(err, max_vec_norm(ev, geoms[e]))
            success = err < self.max_error
            with np.errstate(divide="ignore"):
                ratio = self.max_error / err if err > 0 else math.inf
            if success:
                corr = ratio ** (1.0 / (tab["order2"] + 1))
            else:
                corr = ratio ** (1.0 / (tab["order1"] + 1))
            self.adapt(corr)
            if success:
                break
            self.nrejected += 1
            self.system.set_states(x0)
            self.time = t0
        if noisy:
            self.system.noise_dt = None

    # ------------------------------------------ Python-wrapper level calls
    def assure_sensible_timestep(self):
        if self.adaptive:
            s = self.sensible_timestep()
            if self.dt == 0.0 or self.dt > s:
                self.dt = s

    def steps(self, n):
        self.assure_sensible_timestep()
        for _ in range(n):
            self.step()

    def run(self, duration, assure=True):
        if assure:
            self.assure_sensible_timestep()
        if duration <= 0:
            return
        stop = self.time + duration
        while self.time < stop - self.dt:
            self.step()
        old = self.dt
        self.dt = stop - self.time
        self.step()
        if not self.adaptive:
            self.dt = old
