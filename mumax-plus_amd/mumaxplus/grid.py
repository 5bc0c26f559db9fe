"""Grid (mumaxplus/grid.py of the reference)."""

from . import _mumaxpluscpp as _cpp


class Grid:
    """Create a Grid instance with a size and an origin (in cells)."""

    def __init__(self, size, origin=(0, 0, 0)):
        if len(size) != 3:
            raise ValueError("'size' should have three dimensions.")
        if len(origin) != 3:
            raise ValueError("'origin' should have three dimensions.")
        self._impl = _cpp.Grid(size, origin)

    def __repr__(self):
        return f"Grid(size={self.size}, origin={self.origin})"

    def __eq__(self, other):
        return self._impl == other._impl

    def __ne__(self, other):
        return self._impl != other._impl

    @classmethod
    def _from_impl(cls, impl):
        grid = cls.__new__(cls)
        grid._impl = impl
        return grid

    @property
    def size(self):
        return self._impl.size

    @property
    def origin(self):
        return self._impl.origin

    @property
    def shape(self):
        nx, ny, nz = self.size
        return (nz, ny, nx)

    @property
    def ncells(self):
        return self._impl.ncells
