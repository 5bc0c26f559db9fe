"""Ferromagnet (mumaxplus/ferromagnet.py and magnet.py of the reference)."""

import numpy as _np

from . import _mumaxpluscpp as _cpp
from .dmitensor import DmiTensor
from .fieldquantity import FieldQuantity
from .grid import Grid
from .parameter import Parameter
from .scalarquantity import ScalarQuantity
from .variable import Variable

_SCALAR_PARAMS = ("msat", "alpha", "aex", "ku1", "ku2", "kc1", "kc2", "kc3", "frozen_spins")
_VECTOR_PARAMS = ("anisU", "anisC1", "anisC2", "bias_magnetic_field")
_FIELD_QUANTITIES = ("exchange_field", "anisotropy_field", "external_field", "dmi_field", "demag_field",
                     "effective_field", "torque", "llg_torque", "damping_torque", "exchange_energy_density",
                     "anisotropy_energy_density", "zeeman_energy_density", "dmi_energy_density",
                     "demag_energy_density", "total_energy_density")
_SCALAR_QUANTITIES = ("max_torque", "exchange_energy", "anisotropy_energy", "zeeman_energy", "dmi_energy",
                      "demag_energy", "total_energy")


def _mask_array(value, grid, world, name):
    """Magnet._get_mask_array (magnet.py:52-76 of the reference)."""
    if value is None:
        return None
    T = bool if name == "geometry" else int
    if callable(value):
        nx, ny, nz = grid.size
        cs = world.cellsize
        idxs = _np.flip(_np.mgrid[0:nz, 0:ny, 0:nx], axis=0)
        x, y, z = [(grid.origin[i] + idxs[i]) * cs[i] for i in range(3)]
        return _np.vectorize(value, otypes=[T])(x, y, z)
    arr = _np.array(value, dtype=T)
    if arr.shape != grid.shape:
        raise ValueError(f"The dimensions of the {name} do not match the dimensions of the grid.")
    return arr


class Ferromagnet:
    """Create a ferromagnet in a world on a grid, optionally with a geometry
    (bool array of shape (nz, ny, nx) or a function of (x, y, z))."""

    def __init__(self, world, grid, name="", geometry=None, regions=None):
        geometry_array = _mask_array(geometry, grid, world, "geometry")
        regions_array = _mask_array(regions, grid, world, "regions")
        object.__setattr__(self, "_impl", world._impl.add_ferromagnet(grid._impl, geometry_array, regions_array,
                                                                      name))

    @classmethod
    def _from_impl(cls, impl):
        magnet = cls.__new__(cls)
        object.__setattr__(magnet, "_impl", impl)
        return magnet

    def __repr__(self):
        return f"Ferromagnet(grid={self.grid}, name='{self.name}')"

    # ----------------------------------------------------------- geometry
    @property
    def name(self):
        return self._impl.name

    @property
    def grid(self):
        return Grid._from_impl(self._impl.grid)

    @property
    def cellsize(self):
        return self._impl.system.cellsize

    @property
    def geometry(self):
        return self._impl.system.geometry

    @property
    def origin(self):
        return self._impl.system.origin

    @property
    def center(self):
        return self._impl.system.center

    @property
    def world(self):
        from .world import World
        return World._from_impl(self._impl.world)

    @property
    def meshgrid(self):
        nx, ny, nz = self.grid.size
        idx = _np.flip(_np.mgrid[0:nz, 0:ny, 0:nx], axis=0)
        mgrid = _np.zeros(idx.shape, dtype=_np.float32)
        for c in range(3):
            mgrid[c] = self.origin[c] + idx[c] * self.cellsize[c]
        return mgrid

    @property
    def is_sublattice(self):
        return False

    # -------------------------------------------------------- state / flags
    @property
    def magnetization(self):
        return Variable(self._impl.magnetization)

    @magnetization.setter
    def magnetization(self, value):
        self.magnetization.set(value)

    @property
    def enable_demag(self):
        return self._impl.enable_demag

    @enable_demag.setter
    def enable_demag(self, value):
        self._impl.enable_demag = value

    @property
    def enable_openbc(self):
        return self._impl.enable_openbc

    @enable_openbc.setter
    def enable_openbc(self, value):
        self._impl.enable_openbc = value

    @property
    def enable_as_stray_field_source(self):
        """Whether this magnet's stray field acts on other magnets
        (magnet.py of the reference).  Coupling between magnets is not built:
        evaluating a field it would enter raises RuntimeError."""
        return self._impl.enable_as_stray_field_source

    @enable_as_stray_field_source.setter
    def enable_as_stray_field_source(self, value):
        self._impl.enable_as_stray_field_source = value

    @property
    def enable_as_stray_field_destination(self):
        """Whether stray fields of other magnets act on this magnet."""
        return self._impl.enable_as_stray_field_destination

    @enable_as_stray_field_destination.setter
    def enable_as_stray_field_destination(self, value):
        self._impl.enable_as_stray_field_destination = value

    @property
    def dmi_tensor(self):
        return DmiTensor(self._impl.dmi_tensor)

    # ------------------------------------------- parameters and quantities
    def __getattr__(self, key):
        impl = object.__getattribute__(self, "_impl")
        if key in _SCALAR_PARAMS or key in _VECTOR_PARAMS:
            return Parameter(getattr(impl, key))
        if key in _FIELD_QUANTITIES:
            return FieldQuantity(impl.field_quantity(key))
        if key in _SCALAR_QUANTITIES:
            return ScalarQuantity(impl.scalar_quantity(key))
        raise AttributeError(f"'Ferromagnet' object has no attribute '{key}'")

    def __setattr__(self, key, value):
        if key in _SCALAR_PARAMS or key in _VECTOR_PARAMS:
            getattr(self, key).set(value)
        else:
            object.__setattr__(self, key, value)
