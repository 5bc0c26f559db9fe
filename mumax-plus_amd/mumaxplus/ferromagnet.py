"""Ferromagnet (mumaxplus/ferromagnet.py of the reference)."""

import warnings

import numpy as _np

from .dmitensor import DmiTensor
from .interparameter import InterParameter
from .poissonsystem import PoissonSystem
from .magnet import Magnet
from .variable import Variable

_SCALAR_PARAMS = ("msat", "alpha", "aex", "ku1", "ku2", "kc1", "kc2", "kc3", "frozen_spins", "xi", "Lambda",
                  "free_layer_thickness", "epsilon_prime", "pol", "applied_potential", "conductivity", "amr_ratio")
_VECTOR_PARAMS = ("anisU", "anisC1", "anisC2", "bias_magnetic_field", "fixed_layer", "jcur")
_FIELD_QUANTITIES = ("exchange_field", "anisotropy_field", "external_field", "dmi_field", "demag_field",
                     "magnetoelastic_field", "homogeneous_exchange_field", "inhomogeneous_exchange_field",
                     "effective_field", "torque", "llg_torque", "damping_torque", "spin_transfer_torque",
                     "thermal_noise", "magnetoelastic_force",
                     "full_magnetization", "exchange_energy_density", "anisotropy_energy_density",
                     "zeeman_energy_density", "dmi_energy_density", "demag_energy_density",
                     "magnetoelastic_energy_density", "homogeneous_exchange_energy_density",
                     "inhomogeneous_exchange_energy_density", "total_energy_density", "electrical_potential",
                     "conductivity_tensor")
_SCALAR_QUANTITIES = ("max_torque", "max_angle", "exchange_energy", "anisotropy_energy", "zeeman_energy", "dmi_energy",
                      "demag_energy", "magnetoelastic_energy", "homogeneous_exchange_energy",
                      "inhomogeneous_exchange_energy", "total_energy")


def _warn_positive(param, message):
    """The sign warnings of the reference's B1 / B2 / afmex setters."""
    warn = param.uniform_value > 0 if param.is_uniform else bool(_np.any(param.eval() > 0))
    if warn:
        warnings.warn(message, UserWarning)


class Ferromagnet(Magnet):
    """Create a ferromagnet in a world on a grid, optionally with a geometry
    (bool array of shape (nz, ny, nx) or a function of (x, y, z))."""

    _PARAMS = _SCALAR_PARAMS + _VECTOR_PARAMS
    _FIELD_QUANTITIES = _FIELD_QUANTITIES
    _SCALAR_QUANTITIES = _SCALAR_QUANTITIES

    def __init__(self, world, grid, name="", geometry=None, regions=None):
        self._create(world._impl.add_ferromagnet, world, grid, name, geometry, regions)

    def __repr__(self):
        return f"Ferromagnet(grid={self.grid}, name='{self.name}')"

    # -------------------------------------------------------- state / flags
    @property
    def magnetization(self):
        return Variable(self._impl.magnetization)

    @magnetization.setter
    def magnetization(self, value):
        self.magnetization.set(value)

    @property
    def is_sublattice(self):
        return self._impl.is_sublattice

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
    def dmi_tensor(self):
        return DmiTensor(self._impl.dmi_tensor)

    # ------------------------------------------ spin-transfer torque, noise
    @property
    def enable_zhang_li_torque(self):
        """Enable the Zhang-Li spin-transfer torque (default True)."""
        return self._impl.enable_zhang_li_torque

    @enable_zhang_li_torque.setter
    def enable_zhang_li_torque(self, value):
        self._impl.enable_zhang_li_torque = value

    @property
    def enable_slonczewski_torque(self):
        """Enable the Slonczewski spin-transfer torque (default True)."""
        return self._impl.enable_slonczewski_torque

    @enable_slonczewski_torque.setter
    def enable_slonczewski_torque(self, value):
        self._impl.enable_slonczewski_torque = value

    @property
    def fixed_layer_on_top(self):
        """Whether the Slonczewski fixed layer is above the free layer (default True)."""
        return self._impl.fixed_layer_on_top

    @fixed_layer_on_top.setter
    def fixed_layer_on_top(self, value):
        if not isinstance(value, bool):
            raise TypeError("fixed_layer_on_top should be a boolean")
        self._impl.fixed_layer_on_top = value

    @property
    def temperature(self):
        """Temperature (K); a non-zero value adds the thermal field to the time steps."""
        from .parameter import Parameter
        return Parameter(getattr(self._impl, "temperature"))

    @temperature.setter
    def temperature(self, value):
        # the reference's check for its random generator (ferromagnet.py:536-542),
        # kept so that the same scripts fail the same way
        if self.grid.ncells % 2:
            raise ValueError("The CUDA random number generator used to generate"
                             + " a random noise field only works for an even"
                             + " number of grid cells.\n"
                             + "The used number of grid cells is {}.".format(self.grid.ncells))
        self.temperature.set(value)

    @property
    def noise_seed(self):
        """32-bit key of this magnet's thermal field (reproducible random streams)."""
        return self._impl.noise_seed

    @noise_seed.setter
    def noise_seed(self, value):
        self._impl.noise_seed = value

    @property
    def noise_draws(self):
        """Thermal draws taken by time steps so far (one per step)."""
        return self._impl.noise_draws

    @property
    def RelaxTorqueThreshold(self):
        """Torque threshold of `relax` (default -1: relax until the torque
        stops decreasing; positive: until max torque <= threshold)."""
        return self._impl.RelaxTorqueThreshold

    @RelaxTorqueThreshold.setter
    def RelaxTorqueThreshold(self, value):
        assert value != 0, "The relax threshold should not be zero."
        self._impl.RelaxTorqueThreshold = value

    # --------------------------------------------------- magnetoelasticity
    @property
    def B1(self):
        """First magnetoelastic coupling constant (J/m3)."""
        return self._lookup_param("B1")

    @B1.setter
    def B1(self, value):
        self.B1.set(value)
        _warn_positive(self.B1, "The first magnetoelastic coupling constant B1 is set to a positive value, "
                                "instead of negative (or zero). Make sure this is intentional!")

    @property
    def B2(self):
        """Second magnetoelastic coupling constant (J/m3)."""
        return self._lookup_param("B2")

    @B2.setter
    def B2(self, value):
        self.B2.set(value)
        _warn_positive(self.B2, "The second magnetoelastic coupling constant B2 is set to a positive value, "
                                "instead of negative (or zero). Make sure this is intentional!")

    @property
    def poisson_system(self):
        """The solver of the electrical potential (ferromagnet.py:664-673 of the reference)."""
        return PoissonSystem(self._impl.poisson_system)

    @property
    def inter_exchange(self):
        """Exchange constant (J/m) between different regions; zero (the default)
        uses the harmonic mean of the two regions' aex (ferromagnet.py:248-270
        of the reference)."""
        return InterParameter(self._impl.inter_exchange)

    @inter_exchange.setter
    def inter_exchange(self, value):
        self.inter_exchange.set(value)

    @property
    def scale_exchange(self):
        """Factor on the exchange between different regions (default 1)."""
        return InterParameter(self._impl.scale_exchange)

    @scale_exchange.setter
    def scale_exchange(self, value):
        self.scale_exchange.set(value)

    def _lookup_param(self, key):
        from .parameter import Parameter
        return Parameter(getattr(self._impl, key))

    # ------------------------------------------- parameters and quantities
    def __setattr__(self, key, value):
        if self._is_param(key):
            getattr(self, key).set(value)
        else:
            object.__setattr__(self, key, value)
