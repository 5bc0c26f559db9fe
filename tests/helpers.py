"""Builders that set up the same magnet in the product and in the oracle."""

import numpy as np

from oracle import model as omodel

import mumaxplus as mp

DMI_KEYS = ("xxy", "xyz", "xxz", "yxy", "yyz", "yxz", "zxy", "zyz", "zxz")


def random_unit(rng, shape):
    m = rng.normal(size=(3,) + shape)
    return m / np.linalg.norm(m, axis=0)


def make_pair(size, cellsize, params, m0, geometry=None, mastergrid=(0, 0, 0), pbc=(0, 0, 0), openbc=False,
              demag=True, world_bias=(0.0, 0.0, 0.0), dmi=None):
    """Return (product world, product magnet, oracle world, oracle magnet)."""
    nx, ny, nz = size
    world = mp.World(cellsize, pbc_repetitions=tuple(pbc), mastergrid=mp.Grid(tuple(mastergrid)))
    mag = mp.Ferromagnet(world, mp.Grid(size), geometry=geometry)
    ow = omodel.World(cellsize, mastergrid=tuple(mastergrid), pbc=tuple(pbc))
    om = ow.add_ferromagnet(size, geometry=geometry)
    mag.enable_demag = demag
    om.enable_demag = demag
    mag.enable_openbc = openbc
    om.openbc = openbc
    world.bias_magnetic_field = world_bias
    ow.bias = tuple(world_bias)
    for k, v in params.items():
        pv = v
        if isinstance(v, np.ndarray):
            pv = v.astype(np.float32)
        setattr(mag, k, pv)
        om.set("bias" if k == "bias_magnetic_field" else k, np.asarray(pv, dtype=np.float64)
               if isinstance(pv, np.ndarray) else pv)
    for k, v in (dmi or {}).items():
        pv = v.astype(np.float32) if isinstance(v, np.ndarray) else v
        setattr(mag.dmi_tensor, k, pv)
        om.set(k, np.asarray(pv, dtype=np.float64) if isinstance(pv, np.ndarray) else pv)
    m0 = np.asarray(m0, dtype=np.float32)
    mag.magnetization = m0
    om.set_m(np.asarray(mag.magnetization.get(), dtype=np.float64) if m0.ndim == 4 else m0.astype(np.float64))
    return world, mag, ow, om
