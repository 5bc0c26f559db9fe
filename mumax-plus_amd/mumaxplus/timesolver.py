"""TimeSolver (mumaxplus/timesolver.py of the reference)."""

import os as _os
from typing import Callable


class TimeSolverOutput:
    """Collect values of a dict of quantities at specified time points."""

    def __init__(self, quantity_dict, file_name=None):
        self._quantities = quantity_dict
        self._data = {"time": []}
        for key in self._quantities:
            self._data[key] = []
        self._keys = list(self._data.keys())
        self._file_name = file_name
        if self._file_name is not None:
            directory = _os.path.dirname(self._file_name)
            if directory:
                _os.makedirs(directory, exist_ok=True)
            with open(self._file_name, "w") as f:
                print("# " + "\t".join(self._keys), file=f)

    def write_line(self, time):
        self._data["time"].append(time)
        for key, func in self._quantities.items():
            self._data[key].append(func())
        if self._file_name is not None:
            with open(self._file_name, "a") as f:
                print(*[self._data[key][-1] for key in self._keys], sep="\t", file=f)

    def __getitem__(self, key):
        return self._data[key]


class TimeSolver:
    """Evolve the world in time (use ``world.timesolver``)."""

    def __init__(self, impl):
        object.__setattr__(self, "_impl", impl)

    def _assure_sensible_timestep(self):
        """timesolver.py:61-70 of the reference."""
        if self.adaptive_timestep:
            sensible = self._impl.sensible_timestep
            if self.timestep == 0.0 or self.timestep > sensible:
                self.timestep = sensible

    def set_method(self, method_name):
        """'Heun', 'BogackiShampine', 'CashKarp', 'Fehlberg' (default) or 'DormandPrince'."""
        self._impl.set_method(method_name)

    def steps(self, nsteps):
        self._assure_sensible_timestep()
        self._impl.steps(nsteps)

    def run_while(self, condition: Callable[[], bool]):
        self._assure_sensible_timestep()
        self._impl.run_while(condition)

    def run(self, duration):
        self._assure_sensible_timestep()
        self._impl.run(duration)

    def solve(self, timepoints, quantity_dict, file_name=None, tqdm=False):
        assert all(i1 <= i2 for i1, i2 in zip(timepoints, timepoints[1:])), \
            "The list of timepoints should be increasing."
        assert self.time <= timepoints[0], "The list of timepoints should lie in the future."
        self._assure_sensible_timestep()
        output = TimeSolverOutput(quantity_dict, file_name)
        for tp in timepoints:
            self._impl.run(tp - self.time)
            output.write_line(self.time)
        return output

    _PROPS = ("timestep", "time", "max_error", "headroom", "lower_bound", "upper_bound", "sensible_factor",
              "sensible_timestep_default")
    _ASSERT_POSITIVE = {"max_error": "The maximum error should be bigger than 0.",
                        "headroom": "The headroom should be bigger than 0.",
                        "lower_bound": "The lower bound should be bigger than 0.",
                        "upper_bound": "The upper bound should be bigger than 0.",
                        "sensible_factor": "The sensible factor should be bigger than 0."}

    def __getattr__(self, key):
        if key in TimeSolver._PROPS or key == "adaptive_timestep" or key in ("nsteps", "nrejected", "nrhs"):
            return getattr(object.__getattribute__(self, "_impl"), key)
        raise AttributeError(key)

    def __setattr__(self, key, value):
        if key in TimeSolver._ASSERT_POSITIVE:
            assert value > 0, TimeSolver._ASSERT_POSITIVE[key]
        if key in TimeSolver._PROPS or key == "adaptive_timestep":
            setattr(self._impl, key, value)
        else:
            object.__setattr__(self, key, value)

    @property
    def sensible_timestep(self):
        return self._impl.sensible_timestep
