"""Characteristic micromagnetic quantities (mumaxplus/util/formulary.py of
the reference), with its conventions (no numerical prefactors in the wall
width, wall energy and helical length)."""

import numpy as _np

from .constants import MU0


def magnetostatic_energy_density(msat):
    """Km = mu0 Ms^2 / 2 (J/m3)."""
    return 0.5 * MU0 * msat ** 2


def Km(msat):
    return magnetostatic_energy_density(msat)


def exchange_length(aex, msat):
    """sqrt(A / Km) (m)."""
    return _np.sqrt(aex / Km(msat))


def l_ex(aex, msat):
    return exchange_length(aex, msat)


def wall_width(aex, K_eff):
    """sqrt(A / K_eff) (m)."""
    return _np.sqrt(aex / K_eff)


def wall_energy(aex, K_eff):
    """sqrt(A K_eff) (J/m2)."""
    return _np.sqrt(aex * K_eff)


def helical_length(aex, D):
    """A / D (m)."""
    return aex / D


def magnetic_hardness(K1, msat):
    """sqrt(|K1| / (mu0 Ms^2))."""
    return _np.sqrt(abs(K1) / (MU0 * msat ** 2))
