"""Physical constants (the values of src/physics/constants.hpp of the reference)."""

GAMMALL = 1.7595E11       # gyromagnetic ratio (rad/Ts)
MU0 = 1.25663706212E-6    # vacuum permeability (H/m)
KB = 1.38064852E-23       # Boltzmann constant (J/K)
QE = 1.60217646E-19       # electron charge (C)
MUB = 9.2740091523E-24    # Bohr magneton (J/T)
HBAR = 1.054571817E-34    # reduced Planck constant (Js)
