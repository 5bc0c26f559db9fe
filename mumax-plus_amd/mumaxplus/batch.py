"""Many small independent ferromagnets stepped together on one GPU.

The reference steps one World at a time; a sweep or an ensemble of small
magnets then costs a few kernel launches per term, per magnet, per stage.
``Batch`` packs the magnetisations of its members into one device state and
advances every member with one set of launches per Runge-Kutta stage (a
captured HIP graph per step).  Each member is a Ferromagnet alone in its
World, with uniform material parameters, and keeps its own time step, time,
step controller and stopping rule: a fixed number of steps, ``run`` for a
duration (the rules of TimeSolver.run), or damping-only relaxation until its
max torque is below a threshold.  A member's trajectory does not depend on
the other members.

    batch = Batch(capacity_cells)
    ids = [batch.add(magnet, steps=100) for magnet in magnets]
    batch.run_until_done()          # results are back in each magnet
"""

from . import _mumaxpluscpp as _cpp


class Batch:
    """Batch(capacity, method="DormandPrince"): room for `capacity` cells
    (every member takes its cell count rounded up to 256)."""

    def __init__(self, capacity, method="DormandPrince"):
        self._impl = _cpp.Batch(capacity, method)
        self._magnets = {}
        self._rules = {}

    def add(self, magnet, steps=None, duration=None, relax_torque=None, keep_step=False):
        """Admit `magnet` with exactly one stopping rule; returns its member
        id, or None when its cells do not fit in the free capacity now.
        `keep_step` starts from the world's current step instead of the
        sensible-step rule (as a resumed member does)."""
        rules = [(k, v) for k, v in (("steps", steps), ("run", duration), ("relax", relax_torque)) if v is not None]
        if len(rules) != 1:
            raise ValueError("give exactly one of steps, duration, relax_torque")
        mode, arg = rules[0]
        mid = self._impl.add(magnet._impl, mode, arg, keep_step)
        if mid is not None:
            self._magnets[mid] = magnet
            # a run ends at the world time of admission plus its duration
            stop = magnet._impl._world.timesolver.time + arg if mode == "run" else None
            self._rules[mid] = (mode, arg, stop)
        return mid

    def remove(self, member):
        """Copy the member's state back into its magnet and world, free its
        cells and return the magnet."""
        self._impl.remove(member)
        self._rules.pop(member, None)
        return self._magnets.pop(member)

    def preempt(self, member):
        """Take a running member off the GPU: its state goes back into its
        magnet and world (as with :meth:`remove`) and the returned token
        holds the work left, for :meth:`resume` here or in another batch."""
        st = self.member(member)
        mode, arg, stop = self._rules[member]
        magnet = self.remove(member)
        if mode == "steps":
            left = ("steps", int(arg) - int(st["steps"]))
        elif mode == "run":
            left = ("duration", stop - st["time"])
        else:
            left = ("relax_torque", arg)
        return magnet, left

    def resume(self, token):
        """Admit a preempted member again with the work it had left; returns
        its new member id (None when it does not fit now)."""
        magnet, (key, value) = token
        if key == "steps" and value <= 0 or key == "duration" and value <= 0:
            return None
        return self.add(magnet, keep_step=True, **{key: value})

    def advance(self, nsteps=1):
        """`nsteps` batch steps: one step attempt of every active member each."""
        self._impl.advance(nsteps)

    def finished(self):
        """Ids of the members whose stopping rule is met."""
        return self._impl.finished()

    def member(self, member):
        """time, timestep, steps, rejected, done, error and max_torque of a member."""
        return self._impl.member(member)

    @property
    def members(self):
        return dict(self._magnets)

    def run_until_done(self, poll=32, max_steps=10**9):
        """Advance until every member is finished (checking every `poll`
        steps), then return all members' states to their magnets."""
        taken = 0
        while self._magnets:
            done = set(self.finished())
            if len(done) == len(self._magnets):
                break
            if taken >= max_steps:
                raise RuntimeError("batch members did not finish within max_steps")
            self.advance(poll)
            taken += poll
        for mid in list(self._magnets):
            self.remove(mid)

    def info(self):
        return self._impl.info()

    @property
    def use_graph(self):
        return getattr(self, "_graph", True)

    @use_graph.setter
    def use_graph(self, value):
        self._impl.set_option("graph", bool(value))
        self._graph = bool(value)

    _STEP_MODES = ("passes", "persistent", "auto")

    @property
    def step_mode(self):
        """How a batch step runs: "passes" (one launch per pass, replayed
        from a captured graph), "persistent" (one resident kernel for all the
        steps of an advance, a grid barrier between passes; FFT lines of up
        to 256 points) or "auto" (whichever of these the autotuner measures
        faster for the batch's size class, see :meth:`tune_report`)."""
        return getattr(self, "_step_mode", "passes")

    @step_mode.setter
    def step_mode(self, value):
        if value not in self._STEP_MODES:
            raise ValueError("step_mode must be one of " + ", ".join(self._STEP_MODES))
        self._impl.set_option("step_mode", self._STEP_MODES.index(value))
        self._step_mode = value

    def step_path(self):
        """{"last_path": the path the last advance took, "persist_grid": the
        persistent kernel's grid for the current members (0: cannot run)}."""
        return self._impl.step_path()

    def tune_report(self):
        """The step-path autotuner's record in this process (every batch's
        size classes): {size class: {"choice": path or None, candidate: best
        ms per step, or None when not measured}}."""
        text = self._impl.tune_report()
        out = {}
        for line in text.splitlines():
            parts = [p.strip() for p in line.split("|")]
            entry = {"choice": None if parts[1] == "undecided" else parts[1]}
            for cand in parts[2:]:
                name, _, val = cand.partition(" ")
                entry[name] = float(val.split()[0]) if val.endswith("ms") else None
            out[parts[0]] = entry
        return out

    @staticmethod
    def save_tuning(path):
        """Write the autotuner's record (every size class of this process) to
        a text file."""
        _cpp._batch_tune_save(path)

    @staticmethod
    def load_tuning(path):
        """Merge a record written by save_tuning: its decided size classes
        are used without timing them again."""
        _cpp._batch_tune_load(path)

    @staticmethod
    def clear_tuning():
        _cpp._batch_tune_clear()

    def set_profiling(self, enable):
        self._impl.set_option("profiling", bool(enable))

    def profile_report(self):
        """{kernel name: (total ms, launches, algorithmic bytes)} since profiling was enabled."""
        return self._impl.profile_report()
