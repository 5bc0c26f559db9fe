"""FieldQuantity (mumaxplus/fieldquantity.py of the reference)."""

import numpy as _np

from .grid import Grid


class FieldQuantity:
    """A functor representing a physical field quantity."""

    def __init__(self, impl):
        self._impl = impl

    def __repr__(self):
        return (f"FieldQuantity(grid={self.grid}, name='{self.name}', "
                f"ncomp={self.ncomp}, unit={self.unit})")

    @property
    def name(self):
        return self._impl.name

    @property
    def unit(self):
        return self._impl.unit

    @property
    def ncomp(self):
        return self._impl.ncomp

    @property
    def grid(self):
        return Grid._from_impl(self._impl.grid)

    @property
    def shape(self):
        return (self.ncomp, *self.grid.shape)

    def eval(self):
        return self._impl.eval()

    def __call__(self):
        return self.eval()

    def average(self):
        return self._impl.average()

    def get_rgb(self):
        """RGB colours (3, nz, ny, nx) of a vector quantity (fieldops.cu:316-385
        of the reference), computed on the host by the binding."""
        if self.ncomp != 3:
            raise ValueError("getRGB can only operate on vector fields with 3 components.")
        return self._impl.get_rgb().get()

    @property
    def meshgrid(self):
        nx, ny, nz = self.grid.size
        cellsize = self._impl.system.cellsize
        origin = self._impl.system.cell_position((0, 0, 0))
        mgrid_idx = _np.flip(_np.mgrid[0:nz, 0:ny, 0:nx], axis=0)
        mgrid = _np.zeros(mgrid_idx.shape, dtype=_np.float32)
        for c in range(3):
            mgrid[c] = origin[c] + mgrid_idx[c] * cellsize[c]
        return mgrid
