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
        """RGB colours (ncomp 3, nz, ny, nx) of a vector quantity (fieldops.cu:316-385
        of the reference): the field divided by its largest norm in the
        geometry, then hue = in-plane angle, saturation = norm and lightness
        = (1 + z) / 2, in FP32; grey outside the geometry.  Computed on the
        host from eval(): a display helper, not part of a time step."""
        if self.ncomp != 3:
            raise ValueError("getRGB can only operate on vector fields with 3 components.")
        v = _np.asarray(self.eval(), dtype=_np.float32)
        system = getattr(self._impl, "system", None)
        geom = system.geometry if system is not None else _np.ones(v.shape[1:], dtype=bool)
        norm = _np.sqrt((v * v).sum(axis=0))
        v = v * _np.float32(1.0 / norm[geom].max())
        x, y, z = v
        pi = _np.float32(3.1415926535897931)
        H = _np.arctan2(y, x)
        S = _np.sqrt(x * x + y * y + z * z)
        L = _np.float32(0.5) + _np.float32(0.5) * z
        Hp = _np.float32(3.0) * H / pi
        Hp = _np.where(Hp < 0, Hp + 6, _np.where(Hp >= 6, Hp - 6, Hp)).astype(_np.float32)
        C = _np.where(L <= 0.5, 2 * L * S, 2 * (1 - L) * S).astype(_np.float32)
        X = (C * (1 - _np.abs(_np.fmod(Hp, _np.float32(2)) - 1))).astype(_np.float32)
        m = L - C / 2
        sector = _np.floor(Hp).astype(int).clip(0, 5)
        # (R, G, B) additions of C and X per 60-degree sector of the hue
        addC = _np.array([[1, 0, 0], [0, 1, 0], [0, 1, 0], [0, 0, 1], [0, 0, 1], [1, 0, 0]], dtype=_np.float32)
        addX = _np.array([[0, 1, 0], [1, 0, 0], [0, 0, 1], [0, 1, 0], [1, 0, 0], [0, 0, 1]], dtype=_np.float32)
        rgb = m[None] + C[None] * _np.moveaxis(addC[sector], -1, 0) + X[None] * _np.moveaxis(addX[sector], -1, 0)
        rgb = _np.clip(rgb, 0, 1).astype(_np.float32)
        return _np.where(geom[None], rgb, _np.float32(0.5))

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
