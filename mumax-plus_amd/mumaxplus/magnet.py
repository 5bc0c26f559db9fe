"""Magnet: members shared by ferromagnets and antiferromagnets (mumaxplus/magnet.py
of the reference): geometry, stray-field switches, elastodynamics with its
variables, parameters and quantities."""

import warnings

import numpy as _np

from .fieldquantity import FieldQuantity
from .grid import Grid
from .parameter import Parameter
from .scalarquantity import ScalarQuantity
from .variable import Variable

_ELASTIC_PARAMS = ("C11", "C12", "C44", "eta", "rho", "external_body_force")
_MAGNET_FIELD_QUANTITIES = ("strain_tensor", "stress_tensor", "elastic_force", "effective_body_force",
                            "elastic_damping", "elastic_acceleration", "kinetic_energy_density",
                            "elastic_energy_density", "poynting_vector")
_MAGNET_SCALAR_QUANTITIES = ("kinetic_energy", "elastic_energy")


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


class Magnet:
    """Base of Ferromagnet and Antiferromagnet; not created directly."""

    _FIELD_QUANTITIES = ()
    _SCALAR_QUANTITIES = ()
    _PARAMS = ()

    def _create(self, impl_function, world, grid, name, geometry, regions):
        geometry_array = _mask_array(geometry, grid, world, "geometry")
        regions_array = _mask_array(regions, grid, world, "regions")
        object.__setattr__(self, "_impl", impl_function(grid._impl, geometry_array, regions_array, name))

    @classmethod
    def _from_impl(cls, impl):
        magnet = cls.__new__(cls)
        object.__setattr__(magnet, "_impl", impl)
        return magnet

    def __eq__(self, other):
        return isinstance(other, Magnet) and self._impl == other._impl

    def __hash__(self):
        return hash(self._impl)

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
    def regions(self):
        return self._impl.system.regions

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

    # ------------------------------------------------------ stray fields
    @property
    def enable_as_stray_field_source(self):
        """Whether this magnet's stray field acts on other magnets.  Coupling
        between magnets is not built: evaluating a field it would enter
        raises RuntimeError."""
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

    def stray_field_from_magnet(self, source_magnet):
        """The stray field of `source_magnet` on this magnet (zero outside
        this magnet's geometry)."""
        from .strayfield import StrayField
        return StrayField._from_impl(self._impl.stray_field_from_magnet(source_magnet._impl))

    # ---------------------------------------------------- elastodynamics
    @property
    def enable_elastodynamics(self):
        """Enable/disable elastodynamic time evolution of the elastic
        displacement and velocity (zero when enabled).  Not allowed together
        with a rigid strain."""
        return self._impl.enable_elastodynamics

    @enable_elastodynamics.setter
    def enable_elastodynamics(self, value):
        self._impl.enable_elastodynamics = value

    @property
    def elastic_displacement(self):
        """Elastic displacement (m); exists only with elastodynamics."""
        return Variable(self._impl.elastic_displacement)

    @elastic_displacement.setter
    def elastic_displacement(self, value):
        self.elastic_displacement.set(value)

    @property
    def elastic_velocity(self):
        """Elastic velocity (m/s); exists only with elastodynamics."""
        return Variable(self._impl.elastic_velocity)

    @elastic_velocity.setter
    def elastic_velocity(self, value):
        self.elastic_velocity.set(value)

    @property
    def rigid_norm_strain(self):
        """Applied normal strain (exx, eyy, ezz), used by the magnetoelastic
        field instead of the strain of the displacement."""
        return Parameter(self._impl.rigid_norm_strain)

    @rigid_norm_strain.setter
    def rigid_norm_strain(self, value):
        if self.enable_elastodynamics:
            raise Exception("Can not use normal strain with elastodynamics enabled.")
        self.rigid_norm_strain.set(value)

    @property
    def rigid_shear_strain(self):
        """Applied shear strain (exy, exz, eyz)."""
        return Parameter(self._impl.rigid_shear_strain)

    @rigid_shear_strain.setter
    def rigid_shear_strain(self, value):
        if self.enable_elastodynamics:
            raise Exception("Can not use shear strain with elastodynamics enabled.")
        self.rigid_shear_strain.set(value)

    # ------------------------------------------------ relax and minimize
    _MINIMIZE_SAMPLES = 10

    def minimize(self, tol=1e-6, nsamples=None):
        """Minimise the total energy by steepest descent along the damping
        torque with Barzilai-Borwein steps, until `nsamples` consecutive
        maximum magnetisation changes are <= tol (10 by default, 20 for an
        antiferromagnet)."""
        self._impl.minimize(tol, self._MINIMIZE_SAMPLES if nsamples is None else nsamples)

    def relax(self, tol=1e-9):
        """Relax to an energy minimum by damping-only dynamics: first while
        the energy decreases, then while the torque decreases as the solver
        tolerance is lowered to `tol`."""
        if tol >= 1e-5:
            warnings.warn("The set tolerance is greater than or equal to the default value used for the "
                          "timesolver (1e-5). Using this value results in no torque minimization, only energy "
                          "minimization.", UserWarning)
        self._impl.relax(tol)

    # ----------------------------------------- parameters and quantities
    def _lookup(self, key):
        impl = object.__getattribute__(self, "_impl")
        if key in _ELASTIC_PARAMS or key in type(self)._PARAMS:
            return Parameter(getattr(impl, key))
        if key in _MAGNET_FIELD_QUANTITIES or key in type(self)._FIELD_QUANTITIES:
            return FieldQuantity(impl.field_quantity(key))
        if key in _MAGNET_SCALAR_QUANTITIES or key in type(self)._SCALAR_QUANTITIES:
            return ScalarQuantity(impl.scalar_quantity(key))
        raise AttributeError(f"'{type(self).__name__}' object has no attribute '{key}'")

    def __getattr__(self, key):
        return self._lookup(key)

    def _is_param(self, key):
        return key in _ELASTIC_PARAMS or key in type(self)._PARAMS
