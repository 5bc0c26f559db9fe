"""Antiferromagnet (mumaxplus/antiferromagnet.py of the reference): two
ferromagnetic sublattices on one grid, coupled by the homogeneous (afmex_cell)
and inhomogeneous (afmex_nn) antiferromagnetic exchange and an
inter-sublattice DMI tensor, sharing one demag field and one elastic
displacement."""

import warnings

from .dmitensor import DmiTensor, DmiTensorGroup
from .ferromagnet import Ferromagnet, _warn_positive
from .fieldquantity import FieldQuantity
from .interparameter import InterParameter
from .magnet import Magnet
from .parameter import Parameter
from .scalarquantity import ScalarQuantity


class Antiferromagnet(Magnet):
    """Create an antiferromagnet in a world on a grid, optionally with a
    geometry (bool array of shape (nz, ny, nx) or a function of (x, y, z)).

    Setting an attribute the antiferromagnet does not have but a ferromagnet
    does (msat, aex, alpha, anisotropy, B1, ...) sets it on both sublattices.
    """

    _PARAMS = ()
    _MINIMIZE_SAMPLES = 20
    _FIELD_QUANTITIES = ("neel_vector", "full_magnetization", "total_energy_density", "angle_field")
    _SCALAR_QUANTITIES = ("total_energy", "max_intracell_angle")

    def __init__(self, world, grid, name="", geometry=None, regions=None):
        self._create(world._impl.add_antiferromagnet, world, grid, name, geometry, regions)

    def __repr__(self):
        return f"Antiferromagnet(grid={self.grid}, name='{self.name}')"

    def __setattr__(self, name, value):
        """antiferromagnet.py:53-71 of the reference."""
        if name == "_impl" or hasattr(Antiferromagnet, name):
            object.__setattr__(self, name, value)  # own attribute or property setter
        elif self._is_param(name):
            getattr(self, name).set(value)  # elastic parameters of the magnet
        elif hasattr(Ferromagnet, name) or name in Ferromagnet._PARAMS:
            setattr(self.sub1, name, value)
            setattr(self.sub2, name, value)
        else:
            raise AttributeError(f'Both Antiferromagnet and Ferromagnet have no attribute "{name}".')

    # --------------------------------------------------------- sublattices
    @property
    def sub1(self):
        """First sublattice."""
        return Ferromagnet._from_impl(self._impl.sub1())

    @property
    def sub2(self):
        """Second sublattice."""
        return Ferromagnet._from_impl(self._impl.sub2())

    @property
    def sublattices(self):
        return (self.sub1, self.sub2)

    def other_sublattice(self, sub):
        """The sister sublattice of `sub`."""
        return Ferromagnet._from_impl(self._impl.other_sublattice(sub._impl))

    @property
    def bias_magnetic_field(self):
        """Uniform bias field (T) of both sublattices (read from the first)."""
        return self.sub1.bias_magnetic_field

    @bias_magnetic_field.setter
    def bias_magnetic_field(self, value):
        self.sub1.bias_magnetic_field.set(value)
        self.sub2.bias_magnetic_field.set(value)

    @property
    def enable_demag(self):
        """Demag switch of both sublattices (read from the first)."""
        return self.sub1.enable_demag

    @enable_demag.setter
    def enable_demag(self, value):
        self.sub1.enable_demag = value
        self.sub2.enable_demag = value

    # ---------------------------------------------------------- parameters
    @property
    def afmex_cell(self):
        """Homogeneous (intracell) antiferromagnetic exchange constant (J/m)."""
        return Parameter(self._impl.afmex_cell)

    @afmex_cell.setter
    def afmex_cell(self, value):
        self.afmex_cell.set(value)
        _warn_positive(self.afmex_cell, "The antiferromagnetic exchange constant afmex_cell is set to a "
                                        "positive value, instead of negative (or zero). Make sure this is "
                                        "intentional!")

    @property
    def afmex_nn(self):
        """Inhomogeneous (nearest-neighbour) antiferromagnetic exchange constant (J/m)."""
        return Parameter(self._impl.afmex_nn)

    @afmex_nn.setter
    def afmex_nn(self, value):
        self.afmex_nn.set(value)
        _warn_positive(self.afmex_nn, "The antiferromagnetic exchange constant afmex_nn is set to a "
                                      "positive value, instead of negative (or zero). Make sure this is "
                                      "intentional!")

    @property
    def inter_afmex_nn(self):
        """Antiferromagnetic exchange constant (J/m) between different regions;
        zero (the default) uses the harmonic mean of the two regions' afmex_nn
        (antiferromagnet.py:228-250 of the reference)."""
        return InterParameter(self._impl.inter_afmex_nn)

    @inter_afmex_nn.setter
    def inter_afmex_nn(self, value):
        if value > 0:
            warnings.warn("The antiferromagnetic exchange constant inter_afmex_nn is set to a positive value, "
                          "instead of negative (or zero). Make sure this is intentional!", UserWarning)
        self.inter_afmex_nn.set(value)

    @property
    def scale_afmex_nn(self):
        """Factor on the antiferromagnetic exchange between different regions (default 1)."""
        return InterParameter(self._impl.scale_afmex_nn)

    @scale_afmex_nn.setter
    def scale_afmex_nn(self, value):
        self.scale_afmex_nn.set(value)

    @property
    def latcon(self):
        """Lattice constant (m) of the homogeneous exchange; default 0.35 nm."""
        return Parameter(self._impl.latcon)

    @latcon.setter
    def latcon(self, value):
        self.latcon.set(value)

    @property
    def dmi_tensor(self):
        """Inter-sublattice DMI tensor."""
        return DmiTensor(self._impl.dmi_tensor)

    @property
    def dmi_tensors(self):
        """The inter-sublattice tensor and both sublattices' tensors, set together."""
        return DmiTensorGroup([self.dmi_tensor, self.sub1.dmi_tensor, self.sub2.dmi_tensor])

    # ---------------------------------------------------------- quantities
    @property
    def neel_vector(self):
        """(msat1 m1 - msat2 m2) / (msat1 + msat2)."""
        return FieldQuantity(self._impl.field_quantity("neel_vector"))

    @property
    def full_magnetization(self):
        """msat1 m1 + msat2 m2 (A/m)."""
        return FieldQuantity(self._impl.field_quantity("full_magnetization"))

    @property
    def total_energy_density(self):
        """Total energy density of both sublattices, plus the kinetic and
        elastic energy densities with elastodynamics (J/m3)."""
        return FieldQuantity(self._impl.field_quantity("total_energy_density"))

    @property
    def total_energy(self):
        return ScalarQuantity(self._impl.scalar_quantity("total_energy"))
