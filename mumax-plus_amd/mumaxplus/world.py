"""World (mumaxplus/world.py of the reference)."""

import warnings

from . import _mumaxpluscpp as _cpp
from .grid import Grid
from .timesolver import TimeSolver


class World:
    """Construct a world with a given cell size, optionally periodic.

    pbc_repetitions : number of image repetitions of the mastergrid per axis
    mastergrid      : Grid of the periodic box (0 = open axis)
    """

    def __init__(self, cellsize, pbc_repetitions=(0, 0, 0), mastergrid=Grid((0, 0, 0))):
        if len(cellsize) != 3:
            raise ValueError("'cellsize' should have three dimensions.")
        if len(pbc_repetitions) != 3:
            raise ValueError("'pbc_repetitions' should have three dimensions.")
        if not all(isinstance(item, int) for item in pbc_repetitions):
            raise ValueError("All elements of `pbc_repetitions` must be integers.")
        self._impl = _cpp.World(cellsize, mastergrid._impl, pbc_repetitions)

    def __repr__(self):
        return f"World(cellsize={self.cellsize})"

    @classmethod
    def _from_impl(cls, impl):
        world = cls.__new__(cls)
        world._impl = impl
        return world

    @property
    def timesolver(self):
        return TimeSolver(self._impl.timesolver)

    def get_ferromagnet(self, name):
        from .ferromagnet import Ferromagnet
        impl = self._impl.get_ferromagnet(name)
        if impl is None:
            raise KeyError(f"No magnet named {name}")
        return Ferromagnet._from_impl(impl)

    @property
    def ferromagnets(self):
        from .ferromagnet import Ferromagnet
        return {k: Ferromagnet._from_impl(v) for k, v in self._impl.ferromagnets.items()}

    def get_antiferromagnet(self, name):
        from .antiferromagnet import Antiferromagnet
        impl = self._impl.get_antiferromagnet(name)
        if impl is None:
            raise KeyError(f"No magnet named {name}")
        return Antiferromagnet._from_impl(impl)

    @property
    def antiferromagnets(self):
        from .antiferromagnet import Antiferromagnet
        return {k: Antiferromagnet._from_impl(v) for k, v in self._impl.antiferromagnets.items()}

    @property
    def cellsize(self):
        return self._impl.cellsize

    @property
    def bias_magnetic_field(self):
        return self._impl.bias_magnetic_field

    @bias_magnetic_field.setter
    def bias_magnetic_field(self, value):
        self._impl.bias_magnetic_field = value

    @property
    def mastergrid(self):
        return Grid._from_impl(self._impl.mastergrid)

    @mastergrid.setter
    def mastergrid(self, grid):
        self._impl.mastergrid = grid._impl

    @property
    def pbc_repetitions(self):
        return self._impl.pbc_repetitions

    @pbc_repetitions.setter
    def pbc_repetitions(self, value):
        self._impl.pbc_repetitions = value

    @property
    def bounding_grid(self):
        return Grid._from_impl(self._impl.bounding_grid)

    def set_pbc(self, pbc_repetitions, mastergrid=None):
        """Set periodic boundary conditions; without a mastergrid the bounding
        box of all magnets is used along the repeated axes."""
        if mastergrid is None:
            self._impl.set_pbc(pbc_repetitions)
        else:
            self._impl.set_pbc(mastergrid._impl, pbc_repetitions)

    def unset_pbc(self):
        self._impl.unset_pbc()

    # ------------------------------------------------- snapshots, checkpoints
    def snapshot(self):
        """The dynamic state (magnetisations, elastic state, time solver); see
        :mod:`mumaxplus.checkpoint`."""
        from . import checkpoint
        return checkpoint.snapshot(self)

    def restore(self, snapshot):
        """Put a snapshot taken with :meth:`snapshot` back."""
        from . import checkpoint
        checkpoint.restore(self, snapshot)

    def save_checkpoint(self, path):
        """Write the dynamic state to an .npz file."""
        from . import checkpoint
        checkpoint.save(self, path)

    def load_checkpoint(self, path):
        """Restore the dynamic state from a file written by :meth:`save_checkpoint`."""
        from . import checkpoint
        checkpoint.load(self, path)

    def minimize(self, tol=1e-6, nsamples=10):
        """Minimise the total energy of every magnet of the world together."""
        self._impl.minimize(tol, nsamples)

    def relax(self, tol=1e-9):
        """Relax every magnet of the world (damping-only dynamics)."""
        if tol >= 1e-5:
            warnings.warn("The set tolerance is greater than or equal to the default value used for the "
                          "timesolver (1e-5). Using this value results in no torque minimization, only energy "
                          "minimization.", UserWarning)
        self._impl.relax(tol)

    @property
    def RelaxTorqueThreshold(self):
        """Torque threshold of `relax` (default -1)."""
        return self._impl.RelaxTorqueThreshold

    @RelaxTorqueThreshold.setter
    def RelaxTorqueThreshold(self, value):
        assert value != 0, "The relax threshold should not be zero."
        self._impl.RelaxTorqueThreshold = value
