"""Binding layer over the C ABI of libmumaxplus_amd.so (include/mumaxplus_amd.h).

This module plays the role of the reference's pybind11 extension
``_mumaxpluscpp`` (src/bindings/main.cpp:3-27): it exposes the classes
``Grid``, ``World``, ``Ferromagnet``, ``Antiferromagnet``, ``TimeSolver``, ``Parameter``,
``VectorParameter``, ``Variable``, ``FieldQuantity``, ``ScalarQuantity``,
``DmiTensor`` and ``System`` with the member names of src/bindings/wrap_*.cpp,
and the module function ``_demag_kernel`` (wrap_magnet.cpp:56-59).

All numerical work happens in the HIP library; this layer only converts
arguments (copies, like the reference's numpy conversion) and turns error
codes into Python exceptions (ValueError / RuntimeError / IndexError, the
translations pybind11 applies to the reference's C++ exceptions).  Importing
it fails loudly when the shared library is missing.
"""

import ctypes as _ct
import os as _os

import numpy as _np

_LIBNAME = "libmumaxplus_amd.so"
_LIBPATH = _os.path.join(_os.path.dirname(_os.path.abspath(__file__)), _LIBNAME)
if not _os.path.exists(_LIBPATH):
    raise ImportError(f"{_LIBNAME} is not built (expected at {_LIBPATH}); run the package build first")
_lib = _ct.CDLL(_LIBPATH)

_i3 = _ct.c_int * 3
_d3 = _ct.c_double * 3
_P = _ct.c_void_p
_c_char_p = _ct.c_char_p
_cint = _ct.c_int
_cdouble = _ct.c_double

# int fn(double t, void* ctx, double* out): a parameter's time-dependent term
_TIME_FN = _ct.CFUNCTYPE(_cint, _cdouble, _P, _ct.POINTER(_cdouble))

_SIGS = {
    "mpa_last_error": ([], _c_char_p),
    "mpa_abi_version": ([], _cint),
    "mpa_device_available": ([], _cint),
    "mpa_world_create": ([_P, _P, _P, _P, _ct.POINTER(_P)], _cint),
    "mpa_world_destroy": ([_P], _cint),
    "mpa_world_get_cellsize": ([_P, _P], _cint),
    "mpa_world_set_bias_magnetic_field": ([_P, _P], _cint),
    "mpa_world_get_bias_magnetic_field": ([_P, _P], _cint),
    "mpa_world_add_ferromagnet": ([_P, _P, _P, _P, _P, _c_char_p, _ct.POINTER(_P)], _cint),
    "mpa_world_num_ferromagnets": ([_P, _ct.POINTER(_cint)], _cint),
    "mpa_world_ferromagnet_at": ([_P, _cint, _ct.POINTER(_P)], _cint),
    "mpa_world_find_ferromagnet": ([_P, _c_char_p, _ct.POINTER(_P)], _cint),
    "mpa_world_set_pbc": ([_P, _P, _P, _P], _cint),
    "mpa_world_set_pbc_auto": ([_P, _P], _cint),
    "mpa_world_unset_pbc": ([_P], _cint),
    "mpa_world_get_pbc": ([_P, _P, _P, _P], _cint),
    "mpa_world_bounding_grid": ([_P, _P, _P], _cint),
    "mpa_world_stream": ([_P, _ct.POINTER(_P)], _cint),
    "mpa_fm_name": ([_P, _ct.c_char_p, _cint], _cint),
    "mpa_fm_grid": ([_P, _P, _P], _cint),
    "mpa_fm_geometry": ([_P, _P], _cint),
    "mpa_fm_cellsingeo": ([_P, _ct.POINTER(_cint)], _cint),
    "mpa_fm_set_flag": ([_P, _c_char_p, _cint], _cint),
    "mpa_fm_get_flag": ([_P, _c_char_p, _ct.POINTER(_cint)], _cint),
    "mpa_fm_get_noise_state": ([_P, _ct.POINTER(_ct.c_uint32), _ct.POINTER(_ct.c_uint64),
                                _ct.POINTER(_ct.c_uint64)], _cint),
    "mpa_fm_set_noise_state": ([_P, _ct.c_uint32, _ct.c_uint64, _ct.c_uint64], _cint),
    "mpa_fm_param_info": ([_P, _c_char_p, _ct.POINTER(_cint), _ct.POINTER(_cint), _c_char_p, _cint,
                           _c_char_p, _cint], _cint),
    "mpa_fm_param_set_uniform": ([_P, _c_char_p, _P], _cint),
    "mpa_fm_param_set_cells": ([_P, _c_char_p, _P], _cint),
    "mpa_fm_param_get": ([_P, _c_char_p, _P], _cint),
    "mpa_fm_param_get_uniform": ([_P, _c_char_p, _P], _cint),
    "mpa_fm_param_add_time_term": ([_P, _c_char_p, _TIME_FN, _P, _P], _cint),
    "mpa_fm_param_remove_time_terms": ([_P, _c_char_p], _cint),
    "mpa_fm_param_is_dynamic": ([_P, _c_char_p, _ct.POINTER(_cint)], _cint),
    "mpa_fm_poisson_init": ([_P], _cint),
    "mpa_fm_poisson_solve": ([_P, _P], _cint),
    "mpa_fm_poisson_step": ([_P], _cint),
    "mpa_fm_poisson_restart": ([_P], _cint),
    "mpa_fm_poisson_solver_solve": ([_P, _P], _cint),
    "mpa_fm_poisson_residual": ([_P, _ct.POINTER(_cdouble)], _cint),
    "mpa_fm_poisson_set_method": ([_P, _c_char_p], _cint),
    "mpa_fm_poisson_set_option": ([_P, _c_char_p, _cdouble], _cint),
    "mpa_fm_poisson_get_option": ([_P, _c_char_p, _ct.POINTER(_cdouble)], _cint),
    "mpa_fm_param_set_in_region": ([_P, _c_char_p, _ct.c_uint, _P], _cint),
    "mpa_fm_regions": ([_P, _P], _cint),
    "mpa_fm_unique_regions": ([_P, _P, _cint, _ct.POINTER(_cint)], _cint),
    "mpa_fm_inter_info": ([_P, _c_char_p, _ct.POINTER(_cint), _ct.POINTER(_cdouble), _ct.POINTER(_ct.c_longlong),
                            _c_char_p, _cint, _c_char_p, _cint], _cint),
    "mpa_fm_inter_set": ([_P, _c_char_p, _cdouble], _cint),
    "mpa_fm_inter_set_between": ([_P, _c_char_p, _ct.c_uint, _ct.c_uint, _cdouble], _cint),
    "mpa_fm_inter_get_between": ([_P, _c_char_p, _ct.c_uint, _ct.c_uint, _ct.POINTER(_cdouble)], _cint),
    "mpa_fm_inter_eval": ([_P, _c_char_p, _P], _cint),
    "mpa_fm_set_magnetization": ([_P, _P, _ct.c_longlong], _cint),
    "mpa_fm_get_magnetization": ([_P, _P], _cint),
    "mpa_fm_quantity_ncomp": ([_P, _c_char_p, _ct.POINTER(_cint)], _cint),
    "mpa_fm_eval": ([_P, _c_char_p, _P], _cint),
    "mpa_fm_average": ([_P, _c_char_p, _P], _cint),
    "mpa_fm_scalar": ([_P, _c_char_p, _ct.POINTER(_cdouble)], _cint),
    "mpa_fm_demag_kernel": ([_P, _P, _P], _cint),
    "mpa_fm_host": ([_P, _ct.POINTER(_P)], _cint),
    "mpa_fm_set_variable": ([_P, _c_char_p, _P, _ct.c_longlong], _cint),
    "mpa_fm_get_variable": ([_P, _c_char_p, _P], _cint),
    "mpa_world_add_antiferromagnet": ([_P, _P, _P, _P, _P, _c_char_p, _ct.POINTER(_P)], _cint),
    "mpa_world_num_antiferromagnets": ([_P, _ct.POINTER(_cint)], _cint),
    "mpa_world_antiferromagnet_at": ([_P, _cint, _ct.POINTER(_P)], _cint),
    "mpa_world_find_antiferromagnet": ([_P, _c_char_p, _ct.POINTER(_P)], _cint),
    "mpa_afm_name": ([_P, _ct.c_char_p, _cint], _cint),
    "mpa_afm_grid": ([_P, _P, _P], _cint),
    "mpa_afm_geometry": ([_P, _P], _cint),
    "mpa_afm_cellsingeo": ([_P, _ct.POINTER(_cint)], _cint),
    "mpa_afm_sublattice": ([_P, _cint, _ct.POINTER(_P)], _cint),
    "mpa_afm_set_flag": ([_P, _c_char_p, _cint], _cint),
    "mpa_afm_get_flag": ([_P, _c_char_p, _ct.POINTER(_cint)], _cint),
    "mpa_afm_param_info": ([_P, _c_char_p, _ct.POINTER(_cint), _ct.POINTER(_cint), _c_char_p, _cint,
                            _c_char_p, _cint], _cint),
    "mpa_afm_param_set_uniform": ([_P, _c_char_p, _P], _cint),
    "mpa_afm_param_set_cells": ([_P, _c_char_p, _P], _cint),
    "mpa_afm_param_get": ([_P, _c_char_p, _P], _cint),
    "mpa_afm_param_get_uniform": ([_P, _c_char_p, _P], _cint),
    "mpa_afm_param_add_time_term": ([_P, _c_char_p, _TIME_FN, _P, _P], _cint),
    "mpa_afm_param_remove_time_terms": ([_P, _c_char_p], _cint),
    "mpa_afm_param_is_dynamic": ([_P, _c_char_p, _ct.POINTER(_cint)], _cint),
    "mpa_afm_param_set_in_region": ([_P, _c_char_p, _ct.c_uint, _P], _cint),
    "mpa_afm_regions": ([_P, _P], _cint),
    "mpa_afm_unique_regions": ([_P, _P, _cint, _ct.POINTER(_cint)], _cint),
    "mpa_afm_inter_info": ([_P, _c_char_p, _ct.POINTER(_cint), _ct.POINTER(_cdouble), _ct.POINTER(_ct.c_longlong),
                            _c_char_p, _cint, _c_char_p, _cint], _cint),
    "mpa_afm_inter_set": ([_P, _c_char_p, _cdouble], _cint),
    "mpa_afm_inter_set_between": ([_P, _c_char_p, _ct.c_uint, _ct.c_uint, _cdouble], _cint),
    "mpa_afm_inter_get_between": ([_P, _c_char_p, _ct.c_uint, _ct.c_uint, _ct.POINTER(_cdouble)], _cint),
    "mpa_afm_inter_eval": ([_P, _c_char_p, _P], _cint),
    "mpa_afm_set_variable": ([_P, _c_char_p, _P, _ct.c_longlong], _cint),
    "mpa_afm_get_variable": ([_P, _c_char_p, _P], _cint),
    "mpa_afm_quantity_ncomp": ([_P, _c_char_p, _ct.POINTER(_cint)], _cint),
    "mpa_afm_eval": ([_P, _c_char_p, _P], _cint),
    "mpa_afm_average": ([_P, _c_char_p, _P], _cint),
    "mpa_afm_scalar": ([_P, _c_char_p, _ct.POINTER(_cdouble)], _cint),
    "mpa_fm_stray_field": ([_P, _P, _P, _cint, _P], _cint),
    "mpa_afm_stray_field": ([_P, _P, _P, _cint, _P], _cint),
    "mpa_world_relax": ([_P, _cdouble], _cint),
    "mpa_fm_relax": ([_P, _cdouble], _cint),
    "mpa_afm_relax": ([_P, _cdouble], _cint),
    "mpa_world_minimize": ([_P, _cdouble, _cint], _cint),
    "mpa_fm_minimize": ([_P, _cdouble, _cint], _cint),
    "mpa_afm_minimize": ([_P, _cdouble, _cint], _cint),
    "mpa_world_set_relax_torque_threshold": ([_P, _cdouble], _cint),
    "mpa_world_get_relax_torque_threshold": ([_P, _ct.POINTER(_cdouble)], _cint),
    "mpa_fm_set_relax_torque_threshold": ([_P, _cdouble], _cint),
    "mpa_fm_get_relax_torque_threshold": ([_P, _ct.POINTER(_cdouble)], _cint),
    "mpa_ts_get": ([_P, _c_char_p, _ct.POINTER(_cdouble)], _cint),
    "mpa_ts_set": ([_P, _c_char_p, _cdouble], _cint),
    "mpa_ts_sensible_timestep": ([_P, _ct.POINTER(_cdouble)], _cint),
    "mpa_ts_set_method": ([_P, _c_char_p], _cint),
    "mpa_ts_get_method": ([_P, _c_char_p, _cint], _cint),
    "mpa_ts_step": ([_P], _cint),
    "mpa_ts_steps": ([_P, _ct.c_longlong], _cint),
    "mpa_ts_run": ([_P, _cdouble], _cint),
    "mpa_demag_plan_create": ([_P, _P, _P, _P, _ct.POINTER(_P)], _cint),
    "mpa_demag_plan_destroy": ([_P], _cint),
    "mpa_demag_plan_exec": ([_P, _P, _P, _ct.c_float, _P, _P, _P], _cint),
    "mpa_demag_plan_exec_noalloc": ([_P, _P, _P, _ct.c_float, _P, _P, _P], _cint),
    "mpa_demag_plan_info": ([_P, _P, _ct.POINTER(_cdouble), _ct.POINTER(_cdouble)], _cint),
    "mpa_world_set_profiling": ([_P, _cint], _cint),
    "mpa_world_set_option": ([_P, _c_char_p, _cint], _cint),
    "mpa_world_get_option": ([_P, _c_char_p, _ct.POINTER(_cint)], _cint),
    "mpa_world_profile_reset": ([_P], _cint),
    "mpa_world_profile_report": ([_P, _cint, _P, _P, _P, _P, _ct.POINTER(_cint)], _cint),
    "mpa_batch_create": ([_ct.c_longlong, _c_char_p, _ct.POINTER(_P)], _cint),
    "mpa_batch_destroy": ([_P], _cint),
    "mpa_batch_add": ([_P, _P, _cint, _cdouble, _ct.POINTER(_cint)], _cint),
    "mpa_batch_remove": ([_P, _cint], _cint),
    "mpa_batch_advance": ([_P, _ct.c_longlong], _cint),
    "mpa_batch_finished": ([_P, _cint, _P, _ct.POINTER(_cint)], _cint),
    "mpa_batch_member": ([_P, _cint, _ct.POINTER(_cdouble), _ct.POINTER(_cdouble), _ct.POINTER(_ct.c_longlong),
                          _ct.POINTER(_ct.c_longlong), _ct.POINTER(_cint), _ct.POINTER(_ct.c_float),
                          _ct.POINTER(_ct.c_float)], _cint),
    "mpa_batch_info": ([_P, _ct.POINTER(_ct.c_longlong), _ct.POINTER(_ct.c_longlong), _ct.POINTER(_cint),
                        _ct.POINTER(_cdouble), _ct.POINTER(_cint)], _cint),
    "mpa_batch_set_option": ([_P, _c_char_p, _cint], _cint),
    "mpa_batch_profile_report": ([_P, _cint, _P, _P, _P, _P, _ct.POINTER(_cint)], _cint),
    "mpa_batch_step_path": ([_P, _ct.POINTER(_cint), _ct.POINTER(_cint)], _cint),
    "mpa_bench_dft8": ([_cint, _cint, _cint, _P, _P, _ct.POINTER(_cdouble)], _cint),
    "mpa_batch_tune_save": ([_ct.c_char_p], _cint),
    "mpa_batch_tune_load": ([_ct.c_char_p], _cint),
    "mpa_batch_tune_clear": ([], _cint),
    "mpa_field_tile_report": ([_cint, _P, _ct.POINTER(_cint)], _cint),
    "mpa_field_tile_clear": ([], _cint),
    "mpa_batch_tune_report": ([_P, _cint, _P, _ct.POINTER(_cint)], _cint),
}
for _name, (_args, _res) in _SIGS.items():
    _f = getattr(_lib, _name)
    _f.argtypes = _args
    _f.restype = _res

EXPORTED = tuple(_SIGS)
MPA_ERR_INVALID, MPA_ERR_RUNTIME, MPA_ERR_HIP, MPA_ERR_INDEX = 1, 2, 3, 4


# The exception a time-dependent term raised inside the library; the failing
# entry point re-raises it.
_term_error = None


def _check(rc):
    global _term_error
    if rc == 0:
        return
    msg = _lib.mpa_last_error().decode(errors="replace")
    if _term_error is not None:
        err, _term_error = _term_error, None
        raise err
    if rc == MPA_ERR_INVALID:
        raise ValueError(msg)
    if rc == MPA_ERR_INDEX:
        raise IndexError(msg)
    raise RuntimeError(msg)


def _enc(s):
    if not isinstance(s, str):
        raise TypeError(f"expected a string, got {type(s).__name__}")
    return s.encode()


def device_available():
    return bool(_lib.mpa_device_available())


def _as_i3(v, what):
    v = tuple(v)
    if len(v) != 3:
        raise ValueError(f"'{what}' should have three dimensions.")
    return _i3(*[int(x) for x in v])


def _as_d3(v, what):
    v = tuple(v)
    if len(v) != 3:
        raise ValueError(f"'{what}' should have three dimensions.")
    return _d3(*[float(x) for x in v])


def _ptr(a):
    return a.ctypes.data_as(_P)


def _call(owner, op, *args):
    """mpa_<fm|afm>_<op>(handle, *args) of a magnet handle, with error check."""
    _check(getattr(_lib, f"mpa_{owner._pfx}_{op}")(owner._h, *args))


# ---------------------------------------------------------------- Grid
class Grid:
    """Grid(size, origin=(0,0,0)) (wrap_grid.cpp:4-17)."""

    def __init__(self, size, origin=(0, 0, 0)):
        self._size = tuple(int(x) for x in _as_i3(size, "size"))
        self._origin = tuple(int(x) for x in _as_i3(origin, "origin"))

    @property
    def size(self):
        return self._size

    @property
    def origin(self):
        return self._origin

    @property
    def ncells(self):
        return self.size[0] * self.size[1] * self.size[2]

    def __eq__(self, other):
        return isinstance(other, Grid) and self.size == other.size and self.origin == other.origin

    def __ne__(self, other):
        return not self.__eq__(other)

    def __repr__(self):
        return f"Grid(size={self.size}, origin={self.origin})"


# -------------------------------------------------------------- System
class System:
    """Geometry information of a magnet (wrap_system.cpp:40-51)."""

    def __init__(self, magnet):
        self._fm = magnet

    @property
    def grid(self):
        return self._fm.grid

    @property
    def cellsize(self):
        return self._fm.world.cellsize

    @property
    def geometry(self):
        nx, ny, nz = self.grid.size
        out = _np.empty(nx * ny * nz, dtype=_np.uint8)
        _call(self._fm, "geometry", _ptr(out))
        return out.reshape(nz, ny, nx).astype(bool)

    @property
    def regions(self):
        nx, ny, nz = self.grid.size
        out = _np.empty(nx * ny * nz, dtype=_np.uint32)
        _call(self._fm, "regions", _ptr(out))
        return out.reshape(nz, ny, nx)

    @property
    def unique_regions(self):
        n = _cint()
        _call(self._fm, "unique_regions", None, 0, _ct.byref(n))
        out = _np.empty(max(n.value, 1), dtype=_np.uint32)
        _call(self._fm, "unique_regions", _ptr(out), n.value, _ct.byref(n))
        return [int(v) for v in out[:n.value]]

    def cell_position(self, idx):
        cs = self.cellsize
        o = self.grid.origin
        return tuple((o[a] + idx[a]) * cs[a] for a in range(3))

    @property
    def origin(self):
        return self.cell_position((0, 0, 0))

    @property
    def center(self):
        s = self.grid.size
        cs = self.cellsize
        o = self.origin
        return tuple(o[a] + (s[a] - 1) * cs[a] / 2 for a in range(3))

    @property
    def cellsingeo(self):
        v = _cint()
        _call(self._fm, "cellsingeo", _ct.byref(v))
        return v.value


# ------------------------------------------------------- FieldQuantity
class FieldQuantity:
    """A field quantity of a magnet (wrap_fieldquantity.cpp:8-22)."""

    _UNITS = {"magnetization": "", "full_magnetization": "A/m", "neel_vector": "", "angle_field": "rad",
              "electrical_potential": "V", "conductivity_tensor": "S/m",
              "exchange_field": "T",
              "anisotropy_field": "T", "external_field": "T", "dmi_field": "T", "demag_field": "T",
              "magnetoelastic_field": "T", "homogeneous_exchange_field": "T",
              "inhomogeneous_exchange_field": "T", "effective_field": "T", "torque": "/s", "llg_torque": "/s",
              "damping_torque": "/s", "magnetoelastic_force": "N/m3", "strain_tensor": "",
              "stress_tensor": "N/m2", "elastic_force": "N/m3", "effective_body_force": "N/m3",
              "elastic_damping": "N/m3", "elastic_acceleration": "m/s2", "elastic_velocity": "m/s",
              "elastic_displacement": "m", "poynting_vector": "W/m2"}

    def __init__(self, fm, qname):
        self._fm = fm
        self._q = qname
        nc = _cint()
        _call(fm, "quantity_ncomp", _enc(qname), _ct.byref(nc))
        self._ncomp = nc.value

    @property
    def ncomp(self):
        return self._ncomp

    @property
    def name(self):
        return f"{self._fm.name}:{self._q}"

    @property
    def unit(self):
        return self._UNITS.get(self._q, "J/m3" if self._q.endswith("density") else "")

    @property
    def grid(self):
        return self._fm.grid

    @property
    def system(self):
        return self._fm.system

    def eval(self):
        nx, ny, nz = self.grid.size
        out = _np.empty((self.ncomp, nz, ny, nx), dtype=_np.float32)
        _call(self._fm, "eval", _enc(self._q), _ptr(out))
        return out

    def exec(self):
        self.eval()

    def average(self):
        out = _np.zeros(self.ncomp, dtype=_np.float64)
        _call(self._fm, "average", _enc(self._q), _ptr(out))
        return [float(v) for v in out]

    def get_rgb(self):
        """The quantity's RGB colours as a Field (wrap_fieldquantity.cpp:20; fieldops.cu:316-385)."""
        system = getattr(self, "system", None)
        geom = system.geometry if system is not None else None
        return Field(self.grid, 3, _rgb(self.eval(), geom))


def _rgb(v, geom):
    """RGB colours (3, nz, ny, nx) of a vector field (fieldops.cu:316-385 of
    the reference): the field divided by its largest norm in the geometry,
    then hue = in-plane angle, saturation = norm and lightness = (1 + z) / 2,
    in FP32; grey outside the geometry.  A display helper on the host, not
    part of a time step."""
    v = _np.asarray(v, dtype=_np.float32)
    if v.shape[0] != 3:
        raise ValueError("getRGB can only operate on vector fields with 3 components.")
    if geom is None:
        geom = _np.ones(v.shape[1:], dtype=bool)
    norm = _np.sqrt((v * v).sum(axis=0))
    v = v * _np.float32(1.0 / norm[geom].max())
    x, y, z = v
    pi = _np.float32(3.1415926535897931)
    H = _np.arctan2(y, x)
    S = _np.sqrt(x * x + y * y + z * z)
    L = _np.float32(0.5) + _np.float32(0.5) * z
    Hp = _np.float32(3.0) * H / pi
    Hp = _np.where(Hp < 0, Hp + 6, _np.where(Hp >= 6, Hp - 6, Hp)).astype(_np.float32)
    C = _np.where(L <= 0.5, 2 * L * S, 2 * (1 - L) * S).astype(_np.float32)
    X = (C * (1 - _np.abs(_np.fmod(Hp, _np.float32(2)) - 1))).astype(_np.float32)
    m = L - C / 2
    sector = _np.floor(Hp).astype(int).clip(0, 5)
    # (R, G, B) additions of C and X per 60-degree sector of the hue
    addC = _np.array([[1, 0, 0], [0, 1, 0], [0, 1, 0], [0, 0, 1], [0, 0, 1], [1, 0, 0]], dtype=_np.float32)
    addX = _np.array([[0, 1, 0], [1, 0, 0], [0, 0, 1], [0, 1, 0], [1, 0, 0], [0, 0, 1]], dtype=_np.float32)
    rgb = m[None] + C[None] * _np.moveaxis(addC[sector], -1, 0) + X[None] * _np.moveaxis(addX[sector], -1, 0)
    rgb = _np.clip(rgb, 0, 1).astype(_np.float32)
    return _np.where(geom[None], rgb, _np.float32(0.5))


class Field:
    """A field of values on a grid (wrap_field.cpp:10-17), held on the host:
    ``get()`` returns a copy as (ncomp, nz, ny, nx), ``set()`` takes that
    shape or ncomp values for a uniform field."""

    def __init__(self, grid, ncomp, data=None):
        self._grid = Grid(grid.size, grid.origin)
        self._ncomp = int(ncomp)
        nx, ny, nz = self._grid.size
        self._data = _np.zeros((self._ncomp, nz, ny, nx), dtype=_np.float32)
        if data is not None:
            self.set(data)

    @property
    def grid(self):
        return self._grid

    @property
    def ncomp(self):
        return self._ncomp

    def get(self):
        return self._data.copy()

    def set(self, data):
        a = _np.asarray(data, dtype=_np.float32)
        if a.ndim == 1:
            if a.shape[0] != self._ncomp:
                raise ValueError(f"The number of components do not match, expected {a.shape[0]}, "
                                 f"got {self._ncomp}.")
            self._data[...] = a[:, None, None, None]
        elif a.shape == self._data.shape:
            self._data[...] = a
        else:
            raise ValueError(f"The shape of the data does not match the shape of the field, expected "
                             f"{self._data.shape}, got {a.shape}.")


class ScalarQuantity:
    """Scalar quantity of a magnet (wrap_scalarquantity.cpp)."""

    _UNITS = {"max_torque": "/s", "max_angle": "rad", "max_intracell_angle": "rad"}

    def __init__(self, fm, qname):
        self._fm = fm
        self._q = qname

    @property
    def name(self):
        return f"{self._fm.name}:{self._q}"

    @property
    def unit(self):
        return self._UNITS.get(self._q, "J")

    def eval(self):
        v = _cdouble()
        _call(self._fm, "scalar", _enc(self._q), _ct.byref(v))
        return v.value


# ----------------------------------------------------------- Variable
class Variable(FieldQuantity):
    """The magnetisation of a ferromagnet, or the elastic displacement or
    velocity of a magnet (wrap_variable.cpp:8-16)."""

    def __init__(self, fm, which="magnetization"):
        self._fm = fm
        self._q = which
        self._ncomp = 3

    def get(self):
        nx, ny, nz = self.grid.size
        out = _np.empty((3, nz, ny, nx), dtype=_np.float32)
        if self._q == "magnetization":
            _call(self._fm, "get_magnetization", _ptr(out))
        else:
            _call(self._fm, "get_variable", _enc(self._q), _ptr(out))
        return out

    def eval(self):
        return self.get()

    def set(self, value):
        a = _np.ascontiguousarray(_np.asarray(value, dtype=_np.float32))
        nx, ny, nz = self.grid.size
        if a.ndim == 1:
            if a.shape[0] != 3:
                raise ValueError(f"The number of components do not match, expected {a.shape[0]}, got 3.")
        elif a.shape != (3, nz, ny, nx):
            raise ValueError(f"The shape of the data does not match the shape of the field, expected "
                             f"{(3, nz, ny, nx)}, got {a.shape}.")
        if self._q == "magnetization":
            _call(self._fm, "set_magnetization", _ptr(a), a.size)
        else:
            _call(self._fm, "set_variable", _enc(self._q), _ptr(a), a.size)


# ---------------------------------------------------------- Parameter
class Parameter(FieldQuantity):
    """Material parameter (wrap_parameter.cpp:13-107)."""

    def __init__(self, fm, key):
        self._fm = fm
        self._q = key
        self._key = _enc(key)
        nc, uni = _cint(), _cint()
        nm = _ct.create_string_buffer(256)
        un = _ct.create_string_buffer(64)
        _call(fm, "param_info", self._key, _ct.byref(nc), _ct.byref(uni), nm, 256, un, 64)
        self._ncomp = nc.value
        self._name = nm.value.decode()
        self._unit = un.value.decode()

    @property
    def name(self):
        return self._name

    @property
    def unit(self):
        return self._unit

    @property
    def is_uniform(self):
        nc, uni = _cint(), _cint()
        _call(self._fm, "param_info", self._key, _ct.byref(nc), _ct.byref(uni), None, 0, None, 0)
        return bool(uni.value)

    @property
    def is_dynamic(self):
        out = _cint()
        _call(self._fm, "param_is_dynamic", self._key, _ct.byref(out))
        return bool(out.value)

    @property
    def uniform_value(self):
        out = _np.zeros(self.ncomp, dtype=_np.float64)
        _call(self._fm, "param_get_uniform", self._key, _ptr(out))
        return float(out[0]) if self.ncomp == 1 else tuple(float(v) for v in out)

    @uniform_value.setter
    def uniform_value(self, value):
        self.set(value)

    def eval(self):
        nx, ny, nz = self.grid.size
        out = _np.empty((self.ncomp, nz, ny, nx), dtype=_np.float32)
        _call(self._fm, "param_get", self._key, _ptr(out))
        return out

    def average(self):
        v = self.eval().astype(_np.float64)
        geom = self._fm.system.geometry
        n = max(int(geom.sum()), 1)
        return [float(v[c][geom].sum() / n) for c in range(self.ncomp)]

    def set(self, value):
        if _np.isscalar(value) or (isinstance(value, (tuple, list)) and len(value) == self.ncomp
                                   and all(_np.isscalar(v) for v in value)):
            vals = _np.atleast_1d(_np.asarray(value, dtype=_np.float64))
            if vals.size != self.ncomp:
                raise ValueError(f"expected {self.ncomp} components")
            _call(self._fm, "param_set_uniform", self._key, _ptr(_np.ascontiguousarray(vals)))
            return
        a = _np.asarray(value, dtype=_np.float32)
        nx, ny, nz = self.grid.size
        if a.shape == (nz, ny, nx) and self.ncomp == 1:
            a = a[None]
        if a.shape != (self.ncomp, nz, ny, nx):
            raise ValueError(f"The shape of the data does not match the shape of the parameter, expected "
                             f"{(self.ncomp, nz, ny, nx)}, got {a.shape}.")
        a = _np.ascontiguousarray(a)
        _call(self._fm, "param_set_cells", self._key, _ptr(a))

    def set_in_region(self, region_idx, value):
        vals = _np.ascontiguousarray(_np.atleast_1d(_np.asarray(value, dtype=_np.float64)))
        if vals.size != self.ncomp:
            raise ValueError(f"expected {self.ncomp} components")
        _call(self._fm, "param_set_in_region", self._key, int(region_idx), _ptr(vals))

    def add_time_term(self, term, mask=None):
        """term(t) -> ncomp values; mask None or an array of shape
        (ncomp, nz, ny, nx) (dynamic_parameter.hpp:20-48)."""
        nc = self.ncomp

        def call(t, ctx, out):
            global _term_error
            try:
                v = _np.asarray(term(t), dtype=_np.float64).reshape(-1)
                if v.size != nc:
                    raise ValueError(f"the time-dependent term returned {v.size} values, expected {nc}")
                for c in range(nc):
                    out[c] = float(v[c])
                return 0
            except BaseException as e:  # handed to the entry point that evaluated the term
                _term_error = e
                return 1

        fn = _TIME_FN(call)
        mptr = None
        if mask is not None:
            nx, ny, nz = self.grid.size
            a = _np.asarray(mask, dtype=_np.float32)
            if a.shape != (nc, nz, ny, nx):
                raise ValueError(f"The shape of the mask does not match the shape of the parameter, expected "
                                 f"{(nc, nz, ny, nx)}, got {a.shape}.")
            a = _np.ascontiguousarray(a)
            mptr = _ptr(a)
        _call(self._fm, "param_add_time_term", self._key, fn, None, mptr)
        # the library calls fn until the terms are removed or the world is
        # destroyed; magnet wrappers come and go, so the world that owns the
        # magnet keeps the callbacks (and drops them with itself)
        self._time_fns().setdefault(self._fn_key(), []).append(fn)

    def remove_time_terms(self):
        _call(self._fm, "param_remove_time_terms", self._key)
        self._time_fns().pop(self._fn_key(), None)

    def _time_fns(self):
        return self._fm._world._time_fns

    def _fn_key(self):
        return (self._fm._pfx, self._fm._h.value, self._q)


class VectorParameter(Parameter):
    pass


# ------------------------------------------------------ PoissonSystem
class LinSolver:
    """Iterative solver of a ferromagnet's potential system (wrap_linsolver.cpp:4-15)."""

    def __init__(self, fm):
        self._fm = fm

    def step(self):
        _call(self._fm, "poisson_step")

    def restart(self):
        _call(self._fm, "poisson_restart")

    def max_norm_residual(self):
        out = _cdouble()
        _call(self._fm, "poisson_residual", _ct.byref(out))
        return out.value

    def set_method(self, name):
        _call(self._fm, "poisson_set_method", _enc(name))

    def solve(self):
        nx, ny, nz = self._fm.grid.size
        out = _np.empty(nx * ny * nz, dtype=_np.float64)
        _call(self._fm, "poisson_solver_solve", _ptr(out))
        return out

    def _opt(self, key):
        out = _cdouble()
        _call(self._fm, "poisson_get_option", _enc(key), _ct.byref(out))
        return out.value

    @property
    def max_iter(self):
        return int(self._opt("max_iter"))

    @max_iter.setter
    def max_iter(self, value):
        _call(self._fm, "poisson_set_option", _enc("max_iter"), float(int(value)))

    @property
    def tol(self):
        return self._opt("tol")

    @tol.setter
    def tol(self, value):
        _call(self._fm, "poisson_set_option", _enc("tol"), float(value))


class PoissonSystem:
    """Potential system of a ferromagnet (wrap_poissonsystem.cpp:4-10)."""

    def __init__(self, fm):
        self._fm = fm

    def init(self):
        _call(self._fm, "poisson_init")

    def solve(self):
        nx, ny, nz = self._fm.grid.size
        out = _np.empty((1, nz, ny, nx), dtype=_np.float32)
        _call(self._fm, "poisson_solve", _ptr(out))
        return out

    @property
    def solver(self):
        return LinSolver(self._fm)


# ------------------------------------------------------ InterParameter
class InterParameter:
    """Value between pairs of regions (wrap_parameter.cpp:109-125)."""

    ncomp = 1

    def __init__(self, magnet, key):
        self._fm = magnet
        self._key = _enc(key)

    def _info(self):
        uni, val, size = _cint(), _cdouble(), _ct.c_longlong()
        nm = _ct.create_string_buffer(256)
        un = _ct.create_string_buffer(64)
        _call(self._fm, "inter_info", self._key, _ct.byref(uni), _ct.byref(val), _ct.byref(size), nm, 256, un, 64)
        return bool(uni.value), val.value, size.value, nm.value.decode(), un.value.decode()

    @property
    def name(self):
        return self._info()[3]

    @property
    def unit(self):
        return self._info()[4]

    @property
    def is_uniform(self):
        return self._info()[0]

    @property
    def uniform_value(self):
        uni, val, _, name, _ = self._info()
        if not uni:
            raise ValueError(f"Cannot get uniform value of non-uniform InterParameter {name}.")
        return val

    @uniform_value.setter
    def uniform_value(self, value):
        self.set(value)

    def set(self, value):
        _call(self._fm, "inter_set", self._key, float(value))

    def set_between(self, i, j, value):
        _call(self._fm, "inter_set_between", self._key, int(i), int(j), float(value))

    def get_between(self, i, j):
        out = _cdouble()
        _call(self._fm, "inter_get_between", self._key, int(i), int(j), _ct.byref(out))
        return out.value

    def eval(self):
        size = self._info()[2]
        out = _np.empty(max(size, 1), dtype=_np.float32)
        _call(self._fm, "inter_eval", self._key, _ptr(out))
        return [float(v) for v in out[:size]]

    @property
    def unique_regions(self):
        return self._fm.system.unique_regions

    @property
    def number_of_regions(self):
        return len(self.unique_regions)


class DmiTensor:
    """The nine DMI parameters (wrap_dmitensor.cpp; dmitensor.hpp:12-22)."""

    KEYS = ("xxy", "xyz", "xxz", "yxy", "yyz", "yxz", "zxy", "zyz", "zxz")

    def __init__(self, fm):
        self._elements = {k: Parameter(fm, "dmi_" + k) for k in self.KEYS}


def _element_property(key):
    return property(lambda self: self._elements[key], doc=f"DMI tensor element D_{key} (wrap_dmitensor.cpp)")


for _k in DmiTensor.KEYS:
    setattr(DmiTensor, _k, _element_property(_k))


# -------------------------------------------------------------- Magnets
class _Magnet:
    """Members every magnet has (wrap_magnet.cpp:17-54)."""

    _pfx = None
    _PARAMS = ()
    _VPARAMS = ()
    _INTER = ()
    _ELASTIC_PARAMS = ("C11", "C12", "C44", "eta", "rho")
    _ELASTIC_VPARAMS = ("external_body_force", "rigid_norm_strain", "rigid_shear_strain")

    def __init__(self, handle, world):
        object.__setattr__(self, "_h", handle)
        object.__setattr__(self, "_world", world)
        params = {k: Parameter(self, k) for k in self._PARAMS}
        params.update({k: VectorParameter(self, k) for k in self._VPARAMS})
        params.update({k: InterParameter(self, k) for k in self._INTER})
        if self._has_elastic_params():
            params.update({k: Parameter(self, k) for k in self._ELASTIC_PARAMS})
            params.update({k: VectorParameter(self, k) for k in self._ELASTIC_VPARAMS})
        object.__setattr__(self, "_params", params)

    def _has_elastic_params(self):
        return True

    def __getattr__(self, key):
        params = object.__getattribute__(self, "_params")
        if key in params:
            return params[key]
        raise AttributeError(key)

    def __eq__(self, other):
        return isinstance(other, _Magnet) and self._pfx == other._pfx and self._h.value == other._h.value

    def __hash__(self):
        return hash((self._pfx, self._h.value))

    @property
    def world(self):
        return self._world

    @property
    def name(self):
        buf = _ct.create_string_buffer(256)
        _call(self, "name", buf, 256)
        return buf.value.decode()

    @property
    def grid(self):
        s, o = _i3(), _i3()
        _call(self, "grid", s, o)
        return Grid(tuple(s), tuple(o))

    @property
    def system(self):
        return System(self)

    def _flag(self, key):
        v = _cint()
        _call(self, "get_flag", _enc(key), _ct.byref(v))
        return bool(v.value)

    def _set_flag(self, key, value):
        _call(self, "set_flag", _enc(key), int(bool(value)))

    def _int_flag(self, key):
        v = _cint()
        _call(self, "get_flag", _enc(key), _ct.byref(v))
        return int(v.value)

    def _set_int_flag(self, key, value):
        _call(self, "set_flag", _enc(key), int(value))

    @property
    def enable_as_stray_field_source(self):
        return self._flag("enable_as_stray_field_source")

    @enable_as_stray_field_source.setter
    def enable_as_stray_field_source(self, value):
        self._set_flag("enable_as_stray_field_source", value)

    @property
    def enable_as_stray_field_destination(self):
        return self._flag("enable_as_stray_field_destination")

    @enable_as_stray_field_destination.setter
    def enable_as_stray_field_destination(self, value):
        self._set_flag("enable_as_stray_field_destination", value)

    @property
    def enable_elastodynamics(self):
        return self._flag("enable_elastodynamics")

    @enable_elastodynamics.setter
    def enable_elastodynamics(self, value):
        self._set_flag("enable_elastodynamics", value)

    @property
    def elastic_displacement(self):
        if not self.enable_elastodynamics:
            raise ValueError("elasticDisplacement Variable does not exist yet. Enable elastodynamics first.")
        return Variable(self, "elastic_displacement")

    @property
    def elastic_velocity(self):
        if not self.enable_elastodynamics:
            raise ValueError("elasticVelocity Variable does not exist yet. Enable elastodynamics first.")
        return Variable(self, "elastic_velocity")

    def stray_field_from_magnet(self, source):
        """The stray field of `source` on this magnet's grid, zero outside this
        magnet's geometry (wrap_magnet.cpp:45-53)."""
        if source == self:
            raise RuntimeError("Can not compute the magnetic field of the magnet")
        return StrayField(source, self.grid, mask=self.system.geometry)

    def field_quantity(self, name):
        return FieldQuantity(self, name)

    def scalar_quantity(self, name):
        return ScalarQuantity(self, name)

    def relax(self, tol):
        _call(self, "relax", float(tol))

    def minimize(self, tol, nsamples):
        _call(self, "minimize", float(tol), int(nsamples))


class Ferromagnet(_Magnet):
    """Handle of a ferromagnet or sublattice owned by a World (wrap_ferromagnet.cpp:28-73)."""

    _pfx = "fm"
    _PARAMS = ("msat", "aex", "alpha", "ku1", "ku2", "kc1", "kc2", "kc3", "frozen_spins", "B1", "B2",
               "temperature", "xi", "Lambda", "free_layer_thickness", "epsilon_prime", "pol",
               "applied_potential", "conductivity", "amr_ratio")
    _VPARAMS = ("anisU", "anisC1", "anisC2", "bias_magnetic_field", "fixed_layer", "jcur")
    _INTER = ("inter_exchange", "scale_exchange")

    @property
    def poisson_system(self):
        return PoissonSystem(self)

    def __init__(self, handle, world):
        super().__init__(handle, world)
        object.__setattr__(self, "_dmi", DmiTensor(self))
        object.__setattr__(self, "_magnetization", Variable(self))

    @property
    def dmi_tensor(self):
        return self._dmi

    @property
    def magnetization(self):
        return self._magnetization

    def _has_elastic_params(self):
        return self._host_handle() is None

    def _host_handle(self):
        h = _P()
        _check(_lib.mpa_fm_host(self._h, _ct.byref(h)))
        return h if h.value else None

    @property
    def enable_demag(self):
        return self._flag("enable_demag")

    @enable_demag.setter
    def enable_demag(self, value):
        self._set_flag("enable_demag", value)

    @property
    def enable_openbc(self):
        return self._flag("enable_openbc")

    @enable_openbc.setter
    def enable_openbc(self, value):
        self._set_flag("enable_openbc", value)

    @property
    def enable_zhang_li_torque(self):
        return self._flag("enable_zhang_li_torque")

    @enable_zhang_li_torque.setter
    def enable_zhang_li_torque(self, value):
        self._set_flag("enable_zhang_li_torque", value)

    @property
    def enable_slonczewski_torque(self):
        return self._flag("enable_slonczewski_torque")

    @enable_slonczewski_torque.setter
    def enable_slonczewski_torque(self, value):
        self._set_flag("enable_slonczewski_torque", value)

    @property
    def fixed_layer_on_top(self):
        return self._flag("fixed_layer_on_top")

    @fixed_layer_on_top.setter
    def fixed_layer_on_top(self, value):
        self._set_flag("fixed_layer_on_top", value)

    @property
    def noise_seed(self):
        return self._int_flag("noise_seed") & 0xFFFFFFFF

    @noise_seed.setter
    def noise_seed(self, value):
        v = int(value) & 0xFFFFFFFF
        self._set_int_flag("noise_seed", v - (1 << 32) if v >= (1 << 31) else v)

    @property
    def noise_draws(self):
        return self.noise_state[1]

    @property
    def noise_state(self):
        """(seed, step draws, quantity draws) of the thermal-noise stream, 64-bit counters."""
        seed, a, b = _ct.c_uint32(), _ct.c_uint64(), _ct.c_uint64()
        _check(_lib.mpa_fm_get_noise_state(self._h, _ct.byref(seed), _ct.byref(a), _ct.byref(b)))
        return int(seed.value), int(a.value), int(b.value)

    @noise_state.setter
    def noise_state(self, value):
        seed, a, b = (int(v) for v in value)
        _check(_lib.mpa_fm_set_noise_state(self._h, seed & 0xFFFFFFFF, a, b))

    @property
    def is_sublattice(self):
        return self._host_handle() is not None

    @property
    def RelaxTorqueThreshold(self):
        v = _cdouble()
        _check(_lib.mpa_fm_get_relax_torque_threshold(self._h, _ct.byref(v)))
        return v.value

    @RelaxTorqueThreshold.setter
    def RelaxTorqueThreshold(self, value):
        _check(_lib.mpa_fm_set_relax_torque_threshold(self._h, float(value)))


class Antiferromagnet(_Magnet):
    """Handle of an antiferromagnet owned by a World (wrap_antiferromagnet.cpp:17-32)."""

    _pfx = "afm"
    _PARAMS = ("afmex_cell", "afmex_nn", "latcon")
    _INTER = ("inter_afmex_nn", "scale_afmex_nn")

    def __init__(self, handle, world):
        super().__init__(handle, world)
        object.__setattr__(self, "_dmi", DmiTensor(self))
        subs = []
        for i in range(2):
            h = _P()
            _check(_lib.mpa_afm_sublattice(self._h, i, _ct.byref(h)))
            subs.append(Ferromagnet(h, world))
        object.__setattr__(self, "_subs", tuple(subs))

    @property
    def dmi_tensor(self):
        """The inter-sublattice DMI tensor (wrap_antiferromagnet.cpp:29)."""
        return self._dmi

    def sub1(self):
        return self._subs[0]

    def sub2(self):
        return self._subs[1]

    def sublattices(self):
        return list(self._subs)

    def other_sublattice(self, sub):
        if sub == self._subs[0]:
            return self._subs[1]
        if sub == self._subs[1]:
            return self._subs[0]
        raise IndexError("Sublattice not found in Antiferromagnet.")


def _param_property(key):
    def get(self):
        params = object.__getattribute__(self, "_params")
        if key not in params:
            raise AttributeError(f"'{type(self).__name__}' object has no attribute '{key}'")
        return params[key]
    return property(get, doc=f"the '{key}' parameter (wrap_ferromagnet.cpp, wrap_antiferromagnet.cpp, "
                             f"wrap_magnet.cpp)")


for _k in _Magnet._ELASTIC_PARAMS + _Magnet._ELASTIC_VPARAMS:
    setattr(_Magnet, _k, _param_property(_k))
for _cls in (Ferromagnet, Antiferromagnet):
    for _k in _cls._PARAMS + _cls._VPARAMS + _cls._INTER:
        setattr(_cls, _k, _param_property(_k))

# the reference registers the common base as "Magnet" (wrap_magnet.cpp:17)
Magnet = _Magnet


class StrayField(FieldQuantity):
    """StrayField(magnet, grid) (wrap_strayfield.cpp:10-25): the field of a
    magnet on a grid; set_method("fft" | "brute")."""

    _METHODS = {"fft": 0, "brute": 1}

    def __init__(self, magnet, grid, mask=None):
        if isinstance(magnet, Ferromagnet) and magnet.is_sublattice:
            raise ValueError("the stray field source must be a magnet, not a sublattice")
        self._src = magnet
        self._grid = Grid(grid.size, grid.origin)
        self._method = 0
        self._mask = mask  # destination geometry (bool (nz, ny, nx)) or None
        self._ncomp = 3

    def set_method(self, method):
        if method not in self._METHODS:
            raise ValueError('Method should be "fft" or "brute"')
        self._method = self._METHODS[method]

    @property
    def name(self):
        return f"{self._src.name}:stray_field"

    @property
    def unit(self):
        return "T"

    @property
    def grid(self):
        return self._grid

    @property
    def system(self):
        return None

    def eval(self):
        nx, ny, nz = self._grid.size
        out = _np.empty((3, nz, ny, nx), dtype=_np.float32)
        _call(self._src, "stray_field", _i3(*self._grid.size), _i3(*self._grid.origin), self._method, _ptr(out))
        if self._mask is not None:
            out = _np.where(self._mask[None], out, _np.float32(0))
        return out

    def exec(self):
        self.eval()

    def average(self):
        v = self.eval().astype(_np.float64)
        if self._mask is not None:
            n = max(int(self._mask.sum()), 1)
            return [float(v[c][self._mask].sum() / n) for c in range(3)]
        return [float(v[c].mean()) for c in range(3)]


def _batch_tune_save(path):
    _check(_lib.mpa_batch_tune_save(_enc(str(path))))


def _batch_tune_load(path):
    _check(_lib.mpa_batch_tune_load(_enc(str(path))))


def _batch_tune_clear():
    _check(_lib.mpa_batch_tune_clear())


def _field_tile_report():
    """The tile-schedule autotuner's record: {(nx, ny, nz, variant): {"choice":
    zc, zc_candidate: microseconds per launch or None}}."""
    n = _cint(0)
    _check(_lib.mpa_field_tile_report(0, None, _ct.byref(n)))
    buf = _ct.create_string_buffer(n.value + 1)
    _check(_lib.mpa_field_tile_report(n.value + 1, buf, _ct.byref(n)))
    out = {}
    for line in buf.value.decode().splitlines():
        parts = [p.strip() for p in line.split("|")]
        dims, var = parts[0].split(" variant ")
        key = tuple(int(v) for v in dims.split("x")) + (int(var),)
        rec = {"choice": int(parts[1])}
        for p in parts[2:]:
            zc, t = p.split(" ", 1)
            rec[int(zc)] = None if t == "-" else float(t.split()[0])
        out[key] = rec
    return out


def _field_tile_clear():
    _check(_lib.mpa_field_tile_clear())


def _bench_dft8(variant, lines, rounds):
    """One radix-8 FFT stage applied `rounds` times in LDS to `lines` (n x 8
    complex64) by vector-ALU butterflies (variant 0) or a matrix-core product
    (variant 1): (output lines, kernel milliseconds)."""
    lines = _np.ascontiguousarray(lines, dtype=_np.complex64)
    if lines.ndim != 2 or lines.shape[1] != 8:
        raise ValueError("lines must have shape (n, 8)")
    out = _np.empty_like(lines)
    ms = _cdouble()
    _check(_lib.mpa_bench_dft8(int(variant), lines.shape[0], int(rounds), _ptr(lines), _ptr(out), _ct.byref(ms)))
    return out, ms.value


def _demag_kernel(fm):
    """The demag kernel on the reference kernel grid, shape (6, Kz, Ky, Kx)."""
    ks = _i3()
    _check(_lib.mpa_fm_demag_kernel(fm._h, ks, None))
    kx, ky, kz = tuple(ks)
    out = _np.empty((6, kz, ky, kx), dtype=_np.float32)
    _check(_lib.mpa_fm_demag_kernel(fm._h, ks, _ptr(out)))
    return out


# ---------------------------------------------------------- TimeSolver
class TimeSolver:
    """Time solver of a World (wrap_timesolver.cpp:15-44)."""

    def __init__(self, world):
        object.__setattr__(self, "_world", world)

    _KEYS = ("time", "timestep", "max_error", "headroom", "lower_bound", "upper_bound", "sensible_factor",
             "sensible_timestep_default")

    def _get(self, key):
        v = _cdouble()
        _check(_lib.mpa_ts_get(self._world._h, _enc(key), _ct.byref(v)))
        return v.value

    def _set(self, key, value):
        _check(_lib.mpa_ts_set(self._world._h, _enc(key), float(value)))

    def __getattr__(self, key):
        if key in ("nsteps", "nrejected", "nrhs"):
            return self._get(key)
        raise AttributeError(key)

    @property
    def adaptive_timestep(self):
        return self._get("adaptive_timestep") != 0.0

    @adaptive_timestep.setter
    def adaptive_timestep(self, value):
        self._set("adaptive_timestep", 1.0 if value else 0.0)

    @property
    def sensible_timestep(self):
        v = _cdouble()
        _check(_lib.mpa_ts_sensible_timestep(self._world._h, _ct.byref(v)))
        return v.value

    @property
    def method(self):
        buf = _ct.create_string_buffer(64)
        _check(_lib.mpa_ts_get_method(self._world._h, buf, 64))
        return buf.value.decode()

    def set_method(self, name):
        _check(_lib.mpa_ts_set_method(self._world._h, _enc(name)))

    def step(self):
        _check(_lib.mpa_ts_step(self._world._h))

    def steps(self, n):
        _check(_lib.mpa_ts_steps(self._world._h, int(n)))

    def run(self, duration):
        _check(_lib.mpa_ts_run(self._world._h, float(duration)))

    def run_while(self, condition):
        while condition():
            self.step()


def _solver_property(key):
    return property(lambda self: self._get(key), lambda self, value: self._set(key, value),
                    doc=f"TimeSolver.{key} (wrap_timesolver.cpp:15-44)")


for _k in TimeSolver._KEYS:
    setattr(TimeSolver, _k, _solver_property(_k))


# --------------------------------------------------------------- World
class World:
    """World(cellsize[, mastergrid, pbc_repetitions]) (wrap_world.cpp:36-119)."""

    def __init__(self, cellsize, mastergrid=None, pbc_repetitions=(0, 0, 0)):
        if mastergrid is None:
            mastergrid = Grid((0, 0, 0))
        h = _P()
        _check(_lib.mpa_world_create(_as_d3(cellsize, "cellsize"), _i3(*mastergrid.size), _i3(*mastergrid.origin),
                                     _as_i3(pbc_repetitions, "pbc_repetitions"), _ct.byref(h)))
        self._h = h
        self._magnets = {}
        self._afms = {}
        # (magnet kind, handle, parameter key) -> the time-term callbacks the library holds
        self._time_fns = {}
        self._timesolver = TimeSolver(self)

    @property
    def timesolver(self):
        return self._timesolver

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.mpa_world_destroy(h)
            self._h = None

    @property
    def cellsize(self):
        out = _d3()
        _check(_lib.mpa_world_get_cellsize(self._h, out))
        return tuple(out)

    @property
    def bias_magnetic_field(self):
        out = _d3()
        _check(_lib.mpa_world_get_bias_magnetic_field(self._h, out))
        return tuple(out)

    @bias_magnetic_field.setter
    def bias_magnetic_field(self, b):
        _check(_lib.mpa_world_set_bias_magnetic_field(self._h, _as_d3(b, "bias_magnetic_field")))

    @staticmethod
    def _region_words(grid, regions_array):
        if regions_array is None:
            return None
        nx, ny, nz = grid.size
        r = _np.asarray(regions_array)
        if r.size != nx * ny * nz:
            raise RuntimeError("The size of the region buffer does not match the size of the system.")
        if _np.any(r < 0):
            raise ValueError("region indices must be non-negative")
        return _np.ascontiguousarray(r.reshape(-1).astype(_np.uint32))

    @staticmethod
    def _geometry_bytes(grid, geometry_array, regions_array):
        if geometry_array is None:
            return None
        nx, ny, nz = grid.size
        g = _np.asarray(geometry_array, dtype=bool)
        if g.shape != (nz, ny, nx):
            raise ValueError("The dimensions of the geometry do not match the dimensions of the grid.")
        return _np.ascontiguousarray(g.astype(_np.uint8))

    def add_ferromagnet(self, grid, geometry_array=None, regions_array=None, name=""):
        geo = self._geometry_bytes(grid, geometry_array, regions_array)
        reg = self._region_words(grid, regions_array)
        h = _P()
        _check(_lib.mpa_world_add_ferromagnet(self._h, _i3(*grid.size), _i3(*grid.origin),
                                              None if geo is None else _ptr(geo),
                                              None if reg is None else _ptr(reg), _enc(name or ""),
                                              _ct.byref(h)))
        fm = Ferromagnet(h, self)
        self._magnets[fm.name] = fm
        return fm

    def add_antiferromagnet(self, grid, geometry_array=None, regions_array=None, name=""):
        geo = self._geometry_bytes(grid, geometry_array, regions_array)
        reg = self._region_words(grid, regions_array)
        h = _P()
        _check(_lib.mpa_world_add_antiferromagnet(self._h, _i3(*grid.size), _i3(*grid.origin),
                                                  None if geo is None else _ptr(geo),
                                                  None if reg is None else _ptr(reg), _enc(name or ""),
                                                  _ct.byref(h)))
        afm = Antiferromagnet(h, self)
        self._afms[afm.name] = afm
        return afm

    def get_ferromagnet(self, name):
        return self._magnets.get(name)

    @property
    def ferromagnets(self):
        return dict(self._magnets)

    @property
    def antiferromagnets(self):
        return dict(self._afms)

    def get_antiferromagnet(self, name):
        return self._afms.get(name)

    def _pbc(self):
        s, o, r = _i3(), _i3(), _i3()
        _check(_lib.mpa_world_get_pbc(self._h, s, o, r))
        return Grid(tuple(s), tuple(o)), tuple(r)

    @property
    def mastergrid(self):
        return self._pbc()[0]

    @mastergrid.setter
    def mastergrid(self, grid):
        self.set_pbc(grid, self.pbc_repetitions)

    @property
    def pbc_repetitions(self):
        return self._pbc()[1]

    @pbc_repetitions.setter
    def pbc_repetitions(self, reps):
        self.set_pbc(self.mastergrid, reps)

    def set_pbc(self, *args):
        if len(args) == 1:
            _check(_lib.mpa_world_set_pbc_auto(self._h, _as_i3(args[0], "pbc_repetitions")))
        elif len(args) == 2:
            grid, reps = args
            _check(_lib.mpa_world_set_pbc(self._h, _i3(*grid.size), _i3(*grid.origin),
                                          _as_i3(reps, "pbc_repetitions")))
        else:
            raise TypeError("set_pbc(pbc_repetitions) or set_pbc(mastergrid, pbc_repetitions)")

    def unset_pbc(self):
        _check(_lib.mpa_world_unset_pbc(self._h))

    @property
    def bounding_grid(self):
        s, o = _i3(), _i3()
        _check(_lib.mpa_world_bounding_grid(self._h, s, o))
        return Grid(tuple(s), tuple(o))

    def stream(self):
        v = _P()
        _check(_lib.mpa_world_stream(self._h, _ct.byref(v)))
        return v.value

    def set_option(self, key, value):
        """Stepper switches of the library (mpa_world_set_option), e.g. "fuse_stages"."""
        _check(_lib.mpa_world_set_option(self._h, _enc(key), int(value)))

    def get_option(self, key):
        v = _cint()
        _check(_lib.mpa_world_get_option(self._h, _enc(key), _ct.byref(v)))
        return v.value

    def set_profiling(self, enable):
        _check(_lib.mpa_world_set_profiling(self._h, int(bool(enable))))

    def profile_reset(self):
        _check(_lib.mpa_world_profile_reset(self._h))

    def profile_report(self):
        """{kernel name: (total ms, launches, algorithmic bytes)} since the last reset."""
        cap = 64
        names = _ct.create_string_buffer(32 * cap)
        ms = _np.zeros(cap)
        n = _np.zeros(cap, dtype=_np.int64)
        by = _np.zeros(cap)
        cnt = _cint()
        _check(_lib.mpa_world_profile_report(self._h, cap, names, _ptr(ms), _ptr(n), _ptr(by), _ct.byref(cnt)))
        raw = names.raw
        out = {}
        for i in range(min(cnt.value, cap)):
            nm = raw[32 * i:32 * (i + 1)].split(b"\0", 1)[0].decode()
            out[nm] = (float(ms[i]), int(n[i]), float(by[i]))
        return out

    def relax(self, tol):
        _check(_lib.mpa_world_relax(self._h, float(tol)))

    def minimize(self, tol, nsamples):
        _check(_lib.mpa_world_minimize(self._h, float(tol), int(nsamples)))

    @property
    def RelaxTorqueThreshold(self):
        v = _cdouble()
        _check(_lib.mpa_world_get_relax_torque_threshold(self._h, _ct.byref(v)))
        return v.value

    @RelaxTorqueThreshold.setter
    def RelaxTorqueThreshold(self, value):
        _check(_lib.mpa_world_set_relax_torque_threshold(self._h, float(value)))

    def enable_multi_gpu(self, enable=True):
        if enable:
            raise RuntimeError("this build runs on one GPU")

    def is_multi_gpu_enabled(self):
        return False


# ----------------------------------------------------------- DemagPlan
def _profile_report(fn, h):
    cap = 64
    names = _ct.create_string_buffer(32 * cap)
    ms = _np.zeros(cap)
    n = _np.zeros(cap, dtype=_np.int64)
    by = _np.zeros(cap)
    cnt = _cint()
    _check(fn(h, cap, names, _ptr(ms), _ptr(n), _ptr(by), _ct.byref(cnt)))
    raw = names.raw
    out = {}
    for i in range(min(cnt.value, cap)):
        nm = raw[32 * i:32 * (i + 1)].split(b"\0", 1)[0].decode()
        out[nm] = (float(ms[i]), int(n[i]), float(by[i]))
    return out


class Batch:
    """Independent ferromagnets stepped together (mpa_batch_*; no counterpart
    in the reference)."""

    MODES = {"steps": 0, "run": 1, "relax": 2}

    def __init__(self, capacity, method="DormandPrince"):
        h = _P()
        _check(_lib.mpa_batch_create(int(capacity), _enc(method), _ct.byref(h)))
        self._h = h
        self._members = {}  # id -> magnet binding (keeps its world alive)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.mpa_batch_destroy(h)
            self._h = None

    def add(self, magnet, mode, arg, keep_step=False):
        out = _cint()
        flags = self.MODES[mode] | (0x100 if keep_step else 0)
        _check(_lib.mpa_batch_add(self._h, magnet._h, flags, float(arg), _ct.byref(out)))
        if out.value < 0:
            return None
        self._members[out.value] = magnet
        return out.value

    def remove(self, mid):
        _check(_lib.mpa_batch_remove(self._h, int(mid)))
        return self._members.pop(mid)

    def advance(self, n):
        _check(_lib.mpa_batch_advance(self._h, int(n)))

    def finished(self):
        cnt = _cint()
        _check(_lib.mpa_batch_finished(self._h, 0, None, _ct.byref(cnt)))
        ids = _np.zeros(max(cnt.value, 1), dtype=_np.int32)
        _check(_lib.mpa_batch_finished(self._h, len(ids), _ptr(ids), _ct.byref(cnt)))
        return [int(i) for i in ids[:cnt.value]]

    def member(self, mid):
        t, dt = _cdouble(), _cdouble()
        st, rj = _ct.c_longlong(), _ct.c_longlong()
        dn = _cint()
        er, tq = _ct.c_float(), _ct.c_float()
        _check(_lib.mpa_batch_member(self._h, int(mid), _ct.byref(t), _ct.byref(dt), _ct.byref(st), _ct.byref(rj),
                                     _ct.byref(dn), _ct.byref(er), _ct.byref(tq)))
        return {"time": t.value, "timestep": dt.value, "steps": st.value, "rejected": rj.value,
                "done": bool(dn.value), "error": er.value, "max_torque": tq.value}

    def info(self):
        cap, free = _ct.c_longlong(), _ct.c_longlong()
        mem, lps = _cint(), _cint()
        by = _cdouble()
        _check(_lib.mpa_batch_info(self._h, _ct.byref(cap), _ct.byref(free), _ct.byref(mem), _ct.byref(by),
                                   _ct.byref(lps)))
        return {"capacity": cap.value, "free_cells": free.value, "members": mem.value, "step_bytes": by.value,
                "launches_per_step": lps.value}

    def set_option(self, key, value):
        _check(_lib.mpa_batch_set_option(self._h, _enc(key), int(value)))

    def profile_report(self):
        return _profile_report(_lib.mpa_batch_profile_report, self._h)

    def step_path(self):
        path, grid = _cint(), _cint()
        _check(_lib.mpa_batch_step_path(self._h, _ct.byref(path), _ct.byref(grid)))
        return {"last_path": ("passes", "persistent")[path.value], "persist_grid": grid.value}

    def tune_report(self):
        n = _cint()
        _check(_lib.mpa_batch_tune_report(self._h, 0, None, _ct.byref(n)))
        buf = _ct.create_string_buffer(n.value + 1)
        _check(_lib.mpa_batch_tune_report(self._h, n.value + 1, buf, _ct.byref(n)))
        return buf.value.decode()


class DemagPlan:
    """Stand-alone demag operator on device pointers (mpa_demag_plan_*)."""

    def __init__(self, size, cellsize, mastergrid=(0, 0, 0), pbc=(0, 0, 0)):
        h = _P()
        _check(_lib.mpa_demag_plan_create(_as_i3(size, "size"), _as_d3(cellsize, "cellsize"),
                                          _as_i3(mastergrid, "mastergrid"), _as_i3(pbc, "pbc"), _ct.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.mpa_demag_plan_destroy(h)
            self._h = None

    def exec(self, m_ptr, h_ptr, msat_uniform=1.0, msat_ptr=None, geom_ptr=None, stream=None):
        _check(_lib.mpa_demag_plan_exec(self._h, m_ptr, msat_ptr, float(msat_uniform), geom_ptr, h_ptr, stream))

    def exec_noalloc(self, m_ptr, h_ptr, msat_uniform=1.0, msat_ptr=None, geom_ptr=None, stream=None):
        """exec without any device allocation (the library's static scratch)."""
        _check(_lib.mpa_demag_plan_exec_noalloc(self._h, m_ptr, msat_ptr, float(msat_uniform), geom_ptr, h_ptr,
                                                stream))

    def info(self):
        ks = _i3()
        b, d = _cdouble(), _cdouble()
        _check(_lib.mpa_demag_plan_info(self._h, ks, _ct.byref(b), _ct.byref(d)))
        return tuple(ks), b.value, d.value


# ------------------------------------------- module-level quantity factories
# The reference registers one module function per quantity; each takes a
# magnet and returns a quantity object bound to it (wrap_ferromagnet.cpp:75-128,
# wrap_antiferromagnet.cpp:35-44, wrap_magnet.cpp:56-75).  The objects are the
# same FieldQuantity / ScalarQuantity as the package's, in the subclasses the
# reference registers (wrap_fieldquantity.cpp, wrap_scalarquantity.cpp).
class MagnetFieldQuantity(FieldQuantity):
    """M_FieldQuantity: a field quantity of any magnet."""


class FerromagnetFieldQuantity(FieldQuantity):
    """FM_FieldQuantity: a field quantity of a ferromagnet or sublattice."""


class AntiferromagnetFieldQuantity(FieldQuantity):
    """AFM_FieldQuantity: a field quantity of an antiferromagnet."""


class MagnetScalarQuantity(ScalarQuantity):
    """M_ScalarQuantity."""


class FerromagnetScalarQuantity(ScalarQuantity):
    """FM_ScalarQuantity."""


class AntiferromagnetScalarQuantity(ScalarQuantity):
    """AFM_ScalarQuantity."""


# wrap_ferromagnet.cpp:75-128
_FM_FIELD = ("torque", "llg_torque", "damping_torque", "spin_transfer_torque", "demag_energy_density",
             "anisotropy_field", "anisotropy_energy_density", "exchange_field", "exchange_energy_density",
             "inhomogeneous_exchange_field", "homogeneous_exchange_field", "inhomogeneous_exchange_energy_density",
             "homogeneous_exchange_energy_density", "dmi_field", "dmi_energy_density", "external_field",
             "zeeman_energy_density", "effective_field", "total_energy_density", "conductivity_tensor",
             "electrical_potential", "thermal_noise", "full_magnetization", "magnetoelastic_field",
             "magnetoelastic_energy_density", "magnetoelastic_force")
_FM_SCALAR = ("max_torque", "demag_energy", "anisotropy_energy", "exchange_energy", "max_angle",
              "inhomogeneous_exchange_energy", "homogeneous_exchange_energy", "dmi_energy", "zeeman_energy",
              "total_energy", "magnetoelastic_energy")
# wrap_antiferromagnet.cpp:35-44
_AFM_FIELD = ("neel_vector", "full_magnetization", "angle_field", "total_energy_density")
_AFM_SCALAR = ("max_intracell_angle", "total_energy")
# wrap_magnet.cpp:61-75
_MAGNET_FIELD = ("strain_tensor", "stress_tensor", "elastic_force", "effective_body_force", "elastic_damping",
                 "elastic_acceleration", "kinetic_energy_density", "elastic_energy_density", "poynting_vector")
_MAGNET_SCALAR = ("kinetic_energy", "elastic_energy")


def _quantity_factory(name):
    fm_cls = (FerromagnetFieldQuantity if name in _FM_FIELD else
              FerromagnetScalarQuantity if name in _FM_SCALAR else None)
    afm_cls = (AntiferromagnetFieldQuantity if name in _AFM_FIELD else
               AntiferromagnetScalarQuantity if name in _AFM_SCALAR else None)
    if name in _MAGNET_FIELD:
        fm_cls = afm_cls = MagnetFieldQuantity
    elif name in _MAGNET_SCALAR:
        fm_cls = afm_cls = MagnetScalarQuantity
    kinds = " or ".join(k for k, c in (("Ferromagnet", fm_cls), ("Antiferromagnet", afm_cls)) if c)

    def factory(magnet):
        if fm_cls is not None and isinstance(magnet, Ferromagnet):
            return fm_cls(magnet, name)
        if afm_cls is not None and isinstance(magnet, Antiferromagnet):
            return afm_cls(magnet, name)
        raise TypeError(f"{name}(): incompatible function arguments; expected a {kinds}, "
                        f"got {type(magnet).__name__}")

    factory.__name__ = factory.__qualname__ = name
    factory.__doc__ = f"The {name} quantity of a {kinds}."
    return factory


for _name in dict.fromkeys(_FM_FIELD + _FM_SCALAR + _AFM_FIELD + _AFM_SCALAR + _MAGNET_FIELD + _MAGNET_SCALAR):
    globals()[_name] = _quantity_factory(_name)
del _name, _k, _cls
