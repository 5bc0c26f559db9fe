"""Common initial magnetisations (mumaxplus/util/config.py of the reference).

Each function returns m(x, y, z) -> (mx, my, mz), to assign to a
magnetisation: ``magnet.magnetization = vortex(magnet.center, 20e-9, 1, 1)``.
"""

import numpy as _np


def _unit_sign(value, what):
    if value not in (1, -1):
        raise ValueError(f"The {what} should be 1 or -1.")


def twodomain(m1, mw, m2, wallposition, wallthickness=0.0):
    """Domain m1 for x < wallposition, m2 beyond it; with a positive
    wallthickness, mw is mixed in with the weight exp(-((x - wallposition) /
    wallthickness)^2) (a wall perpendicular to x)."""

    def m(x, y, z):
        side = m1 if x < wallposition else m2
        if wallthickness <= 0:
            return side
        w = _np.exp(-((x - wallposition) / wallthickness) ** 2)
        return tuple((1 - w) * side[c] + w * mw[c] for c in range(3))

    return m


def _swirl(position, diameter, circulation, polarization, anti):
    _unit_sign(circulation, "circulation")
    _unit_sign(polarization, "polarization")
    cx, cy = position[0], position[1]

    def m(x, y, z):
        dx, dy = x - cx, y - cy
        r2 = dx * dx + dy * dy
        if r2 == 0.0:
            return (0, 0, polarization)
        r = _np.sqrt(r2)
        # in-plane curl (vortex) or saddle (antivortex), and a Gaussian core
        ip = (-dx, dy) if anti else (-dy, dx)
        v = _np.array([circulation * ip[0] / r, circulation * ip[1] / r,
                       2 * polarization * _np.exp(-r2 / diameter ** 2)])
        v /= _np.linalg.norm(v)
        return tuple(v)

    return m


def vortex(position, diameter, circulation, polarization):
    """A vortex centred at position (x, y used), core diameter `diameter`,
    circulation and core polarization +1 or -1."""
    return _swirl(position, diameter, circulation, polarization, anti=False)


def antivortex(position, diameter, circulation, polarization):
    """An antivortex, as vortex."""
    return _swirl(position, diameter, circulation, polarization, anti=True)


def _skyrmion(position, radius, charge, polarization, bloch):
    _unit_sign(charge, "charge")
    _unit_sign(polarization, "polarization")
    cx, cy = position[0], position[1]

    def m(x, y, z):
        dx, dy = x - cx, y - cy
        r = _np.sqrt(dx * dx + dy * dy)
        if r == 0.0:
            return (0, 0, polarization)
        mz = 2 * polarization * (_np.exp(-(r / radius) ** 2) - 0.5)
        ip = 1 - abs(mz)
        ux, uy = (-dy / r, dx / r) if bloch else (dx / r, dy / r)  # tangential or radial
        return (charge * ux * ip, charge * uy * ip, mz)

    return m


def neelskyrmion(position, radius, charge, polarization):
    """A Néel skyrmion (radial in-plane rotation) of the given radius."""
    return _skyrmion(position, radius, charge, polarization, bloch=False)


def blochskyrmion(position, radius, charge, polarization):
    """A Bloch skyrmion (tangential in-plane rotation) of the given radius."""
    return _skyrmion(position, radius, charge, polarization, bloch=True)
