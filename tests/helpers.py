"""Builders that set up the same magnet in the product and in the oracle."""

import ctypes

import numpy as np

from oracle import model as omodel

import mumaxplus as mp

DMI_KEYS = ("xxy", "xyz", "xxz", "yxy", "yyz", "yxz", "zxy", "zyz", "zxz")


def random_unit(rng, shape):
    m = rng.normal(size=(3,) + shape)
    return m / np.linalg.norm(m, axis=0)


def make_pair(size, cellsize, params, m0, geometry=None, mastergrid=(0, 0, 0), pbc=(0, 0, 0), openbc=False,
              demag=True, world_bias=(0.0, 0.0, 0.0), dmi=None, regions=None):
    """Return (product world, product magnet, oracle world, oracle magnet)."""
    nx, ny, nz = size
    world = mp.World(cellsize, pbc_repetitions=tuple(pbc), mastergrid=mp.Grid(tuple(mastergrid)))
    mag = mp.Ferromagnet(world, mp.Grid(size), geometry=geometry, regions=regions)
    ow = omodel.World(cellsize, mastergrid=tuple(mastergrid), pbc=tuple(pbc))
    om = ow.add_ferromagnet(size, geometry=geometry)
    om.regions = None if regions is None else np.asarray(regions, dtype=int)
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


class DeviceBuffer:
    """Raw device memory through the HIP runtime the library itself loaded
    (libamdhip64.so.7 of /opt/rocm), for tests of the device-pointer entry points."""

    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            lib = ctypes.CDLL("libamdhip64.so.7")
            lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
            lib.hipFree.argtypes = [ctypes.c_void_p]
            lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
            lib.hipDeviceSynchronize.argtypes = []
            cls._hip = lib
        return cls._hip

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        rc = self.hip().hipMalloc(ctypes.byref(p), self.nbytes)
        if rc != 0:
            raise RuntimeError(f"hipMalloc failed ({rc})")
        self.ptr = p.value

    @classmethod
    def from_host(cls, arr):
        a = np.ascontiguousarray(arr)
        buf = cls(a.nbytes)
        rc = cls.hip().hipMemcpy(buf.ptr, a.ctypes.data, a.nbytes, 1)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy H2D failed ({rc})")
        return buf

    def to_host(self, dtype, shape):
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        self.hip().hipDeviceSynchronize()
        rc = self.hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy D2H failed ({rc})")
        return out

    def __del__(self):
        if getattr(self, "ptr", None):
            self.hip().hipFree(self.ptr)
            self.ptr = None
