"""Minimal world / ferromagnet container for the oracle (test infrastructure only).

Mirrors the parameter set and defaults of the reference Ferromagnet
(src/physics/ferromagnet.cpp:18-61), the neighbour wrap against the world's
mastergrid (src/core/grid.hpp:75-83, src/core/system.hpp:207-213), the
"assured zero" rules that decide which terms enter the effective field
(exchange.cu:14-18, dmi.cu:80-85, anisotropy.cu:9-23, zeeman.cu:38-51,
demag.cpp:9-11) and the equation list of MumaxWorld::resetTimeSolverEquations
(src/physics/mumaxworld.cpp:159-205, one (m, LLG torque) equation per
ferromagnet).

Parameters are either Python scalars (uniform) or arrays of shape
(nz, ny, nx) (vector parameters: (3, nz, ny, nx)).
"""

import numpy as np

from . import demag as _demag
from . import fields as _fields
from . import rk as _rk

DMI_KEYS = ("xxy", "xyz", "xxz", "yxy", "yyz", "yxz", "zxy", "zyz", "zxz")

SCALAR_DEFAULTS = {"msat": 1.0, "aex": 0.0, "alpha": 0.0, "ku1": 0.0, "ku2": 0.0,
                   "kc1": 0.0, "kc2": 0.0, "kc3": 0.0, "frozen_spins": 0.0}
VECTOR_DEFAULTS = {"anisU": (0.0, 0.0, 0.0), "anisC1": (0.0, 0.0, 0.0),
                   "anisC2": (0.0, 0.0, 0.0), "bias": (0.0, 0.0, 0.0)}


class Ferromagnet:
    def __init__(self, size, cellsize, origin=(0, 0, 0), geometry=None,
                 mastergrid=(0, 0, 0), mastergrid_origin=(0, 0, 0), pbc=(0, 0, 0)):
        self.size = tuple(int(v) for v in size)
        self.cellsize = tuple(float(v) for v in cellsize)
        self.origin = tuple(int(v) for v in origin)
        nx, ny, nz = self.size
        self.shape = (nz, ny, nx)
        self.geometry = None if geometry is None else np.asarray(geometry, dtype=bool)
        self.mastergrid = tuple(int(v) for v in mastergrid)
        self.mastergrid_origin = tuple(int(v) for v in mastergrid_origin)
        self.pbc = tuple(int(v) for v in pbc)
        self.params = dict(SCALAR_DEFAULTS)
        self.params.update(VECTOR_DEFAULTS)
        self.dmi = {k: 0.0 for k in DMI_KEYS}
        self.enable_demag = True
        self.openbc = False
        self.m = np.zeros((3,) + self.shape)
        self.m[0] = 1.0
        self._plan = None
        self._nb_cache = {}

    # ----------------------------------------------------------- parameters
    def set(self, name, value):
        if name in DMI_KEYS:
            self.dmi[name] = value
        else:
            self.params[name] = value
        if name == "msat":
            self._plan = None

    def val(self, name):
        v = self.dmi[name] if name in DMI_KEYS else self.params[name]
        return np.broadcast_to(np.asarray(v, dtype=float), self.shape)

    def vec(self, name):
        v = np.asarray(self.params[name], dtype=float)
        if v.shape == (3,):
            return np.broadcast_to(v[:, None, None, None], (3,) + self.shape)
        return np.broadcast_to(v, (3,) + self.shape)

    def assured_zero(self, name):
        v = self.dmi[name] if name in DMI_KEYS else self.params[name]
        a = np.asarray(v, dtype=float)
        return a.ndim == 0 and a == 0 or (a.shape == (3,) and not a.any())

    def dmi_arrays(self):
        return {k: self.val(k) for k in DMI_KEYS}

    def exchange_assured_zero(self):
        return self.assured_zero("aex") or self.assured_zero("msat")

    def dmi_assured_zero(self):
        return self.assured_zero("msat") or all(self.assured_zero(k) for k in DMI_KEYS)

    def uniaxial_assured_zero(self):
        return ((self.assured_zero("ku1") and self.assured_zero("ku2"))
                or self.assured_zero("anisU") or self.assured_zero("msat"))

    def cubic_assured_zero(self):
        return ((self.assured_zero("kc1") and self.assured_zero("kc2") and self.assured_zero("kc3"))
                or (self.assured_zero("anisC1") and self.assured_zero("anisC2"))
                or self.assured_zero("msat"))

    # -------------------------------------------------------------- geometry
    def geom_mask(self):
        if self.geometry is None:
            return np.ones(self.shape, dtype=bool)
        return self.geometry

    def neighbor(self, d):
        """(in_geometry mask, clipped local index tuple) of the neighbour in direction d."""
        key = tuple(d)
        if key in self._nb_cache:
            return self._nb_cache[key]
        nz, ny, nx = self.shape
        Z, Y, X = np.indices(self.shape)
        loc = [X, Y, Z]
        n = [nx, ny, nz]
        glob = []
        for a in range(3):
            g = loc[a] + self.origin[a] + d[a]
            if self.mastergrid[a] > 0:
                mo = self.mastergrid_origin[a]
                g = mo + np.mod(g - mo, self.mastergrid[a])
            glob.append(g - self.origin[a])
        in_grid = np.ones(self.shape, dtype=bool)
        clipped = []
        for a in range(3):
            in_grid &= (glob[a] >= 0) & (glob[a] < n[a])
            clipped.append(np.clip(glob[a], 0, n[a] - 1))
        idx = (clipped[2], clipped[1], clipped[0])
        in_geom = in_grid & self.geom_mask()[idx]
        self._nb_cache[key] = (in_geom, idx)
        return in_geom, idx

    def demag_plan(self):
        if self._plan is None:
            self._plan = _demag.DemagPlan(self.size, self.cellsize, self.mastergrid, self.pbc)
        return self._plan

    # -------------------------------------------------------- magnetisation
    def set_m(self, m):
        m = np.array(m, dtype=float)
        if m.shape == (3,):
            m = np.broadcast_to(m[:, None, None, None], (3,) + self.shape).copy()
        self.m = normalize(m, self.geom_mask())


def normalize(m, geom):
    """k_normalize (src/core/fieldops.cu:232-247): m/|m| inside the geometry,
    zero outside (the geometry mask is applied when a variable is set)."""
    n = np.sqrt(m[0] ** 2 + m[1] ** 2 + m[2] ** 2)
    with np.errstate(divide="ignore", invalid="ignore"):
        out = m / n
    return np.where(geom[None], out, 0.0)


class World:
    """World with ferromagnets and a time solver (FP64)."""

    def __init__(self, cellsize, mastergrid=(0, 0, 0), pbc=(0, 0, 0)):
        self.cellsize = tuple(float(c) for c in cellsize)
        self.mastergrid = tuple(mastergrid)
        self.pbc = tuple(pbc)
        self.bias = (0.0, 0.0, 0.0)
        self.magnets = []
        self.solver = _rk.TimeSolver(self)

    def add_ferromagnet(self, size, origin=(0, 0, 0), geometry=None, defer_timestep=False):
        mag = Ferromagnet(size, self.cellsize, origin, geometry, self.mastergrid, (0, 0, 0), self.pbc)
        self.magnets.append(mag)
        if not defer_timestep:
            self.solver.equations_changed()
        return mag

    def torque(self, mag, m):
        h = _fields.effective_field(mag, m, self.bias)
        return _fields.llg_torque(mag, m, h)

    # interface used by rk.TimeSolver
    def get_states(self):
        return [mag.m for mag in self.magnets]

    def set_states(self, states):
        for mag, x in zip(self.magnets, states):
            mag.m = normalize(x, mag.geom_mask())

    def rhs(self, states):
        return [self.torque(mag, x) for mag, x in zip(self.magnets, states)]

    def geom_masks(self):
        return [mag.geom_mask() for mag in self.magnets]
