"""Utilities: physical constants, geometry shapes, initial magnetisations and
characteristic quantities."""

from . import constants, shape  # noqa: F401
from .config import antivortex, blochskyrmion, neelskyrmion, twodomain, vortex  # noqa: F401
from .constants import GAMMALL, HBAR, KB, MU0, MUB, QE  # noqa: F401
from .formulary import (Km, exchange_length, helical_length, l_ex, magnetic_hardness,  # noqa: F401
                        magnetostatic_energy_density, wall_energy, wall_width)
from .shape import Circle, Ellipse, Rectangle, Shape, Square  # noqa: F401
