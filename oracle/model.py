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

from . import afm as _afm
from . import demag as _demag
from . import elastic as _elastic
from . import fields as _fields
from . import rk as _rk
from . import thermal as _thermal

DMI_KEYS = ("xxy", "xyz", "xxz", "yxy", "yyz", "yxz", "zxy", "zyz", "zxz")

SCALAR_DEFAULTS = {"msat": 1.0, "aex": 0.0, "alpha": 0.0, "ku1": 0.0, "ku2": 0.0,
                   "kc1": 0.0, "kc2": 0.0, "kc3": 0.0, "frozen_spins": 0.0,
                   # magnetoelastic constants (ferromagnet.cpp:24-61)
                   "B1": 0.0, "B2": 0.0,
                   # elastic parameters of an independent magnet (magnet.cpp:26-34)
                   "C11": 0.0, "C12": 0.0, "C44": 0.0, "eta": 0.0, "rho": 1.0,
                   # temperature and spin-transfer torque (ferromagnet.cpp:34-56)
                   "temperature": 0.0, "xi": 0.0, "Lambda": 0.0, "epsilon_prime": 0.0, "pol": 0.0,
                   # electrical potential (ferromagnet.cpp:53-55)
                   "applied_potential": float("nan"), "conductivity": 0.0, "amr_ratio": 0.0}
VECTOR_DEFAULTS = {"anisU": (0.0, 0.0, 0.0), "anisC1": (0.0, 0.0, 0.0),
                   "anisC2": (0.0, 0.0, 0.0), "bias": (0.0, 0.0, 0.0),
                   "external_body_force": (0.0, 0.0, 0.0), "rigid_norm_strain": (0.0, 0.0, 0.0),
                   "rigid_shear_strain": (0.0, 0.0, 0.0), "fixed_layer": (0.0, 0.0, 0.0),
                   "jcur": (0.0, 0.0, 0.0)}


class _Params:
    """Uniform (scalar or 3-vector) or per-cell parameters plus a DMI tensor."""

    def set(self, name, value):
        if name in DMI_KEYS:
            self.dmi[name] = value
        else:
            self.params[name] = value

    # Time-dependent terms (DynamicParameter::evalTimeDependentTerms,
    # dynamic_parameter.cpp:8-29; Parameter::eval, parameter.cu:65-91): the
    # value is static + sum_k fn_k(t) (* mask_k), t = clock(), which the world
    # binds to its time solver's (stage) time.
    def add_time_term(self, name, fn, mask=None):
        terms = self.__dict__.setdefault("time_terms", {})
        terms.setdefault(name, []).append((fn, None if mask is None else np.asarray(mask, dtype=float)))

    def remove_time_terms(self, name):
        self.__dict__.get("time_terms", {}).pop(name, None)

    def _dynamic(self, name, ncomp):
        terms = self.__dict__.get("time_terms", {}).get(name)
        if not terms:
            return None
        t = self.__dict__.get("clock", lambda: 0.0)()
        total = 0.0
        for fn, mask in terms:
            v = np.asarray(fn(t), dtype=float).reshape(-1)
            if ncomp == 1:
                total = total + (v[0] if mask is None else v[0] * mask.reshape(self.shape))
            else:
                vv = v[:, None, None, None]
                total = total + (vv if mask is None else vv * mask.reshape((3,) + self.shape))
        return total

    def val(self, name):
        v = self.dmi[name] if name in DMI_KEYS else self.params[name]
        base = np.broadcast_to(np.asarray(v, dtype=float), self.shape)
        d = self._dynamic(name, 1)
        return base if d is None else base + d

    def vec(self, name):
        v = np.asarray(self.params[name], dtype=float)
        if v.shape == (3,):
            base = np.broadcast_to(v[:, None, None, None], (3,) + self.shape)
        else:
            base = np.broadcast_to(v, (3,) + self.shape)
        d = self._dynamic(name, 3)
        return base if d is None else base + d

    def assured_zero(self, name):
        if self.__dict__.get("time_terms", {}).get(name):
            return False
        v = self.dmi[name] if name in DMI_KEYS else self.params[name]
        a = np.asarray(v, dtype=float)
        return a.ndim == 0 and a == 0 or (a.shape == (3,) and not a.any())

    def dmi_arrays(self):
        return {k: self.val(k) for k in DMI_KEYS}

    # Regions and values between pairs of regions (System::regions_,
    # system.cpp:13-36; InterParameter, inter_parameter.cu:7-82): `regions` is
    # None (every cell in region 0) or an int array of the grid's shape;
    # inter[name] is a scalar (every pair) or a symmetric matrix indexed by
    # region.
    def set_between(self, name, i, j, value):
        if i == j:
            raise ValueError("region indexes must be different")
        v = self.inter[name]
        if np.ndim(v) == 0:
            n = int(np.max(self.regions)) + 1
            v = np.full((n, n), float(v))
        v[i, j] = v[j, i] = value
        self.inter[name] = v

    def inter_assured_zero(self, name):
        v = self.inter[name]
        return np.ndim(v) == 0 and v == 0

    def region_link(self, inter_key, scale_key, nb_in, nb, hm):
        """Exchange constant of the links to in-geometry neighbours in
        different regions: the inter-region value when non-zero, else hm,
        times the inter-region scale (exchange.cu:79-99)."""
        if self.regions is None:
            return hm
        r = self.regions
        r_ = r[nb]
        diff = nb_in & (r != r_)

        def pair(v):
            return np.full(r.shape, float(v)) if np.ndim(v) == 0 else v[r, r_]

        iv, sv = pair(self.inter[inter_key]), pair(self.inter[scale_key])
        return np.where(diff, np.where(iv != 0, iv, hm) * sv, hm)

    # elasticity of the magnet that owns the displacement (magnet.cpp:26-34)
    def elasticity_assured_zero(self):
        """elastodynamics.cu:47-51."""
        return (not self.enable_elastodynamics
                or (self.assured_zero("C11") and self.assured_zero("C12") and self.assured_zero("C44")))

    def rigid_strain_applied(self):
        return not (self.assured_zero("rigid_norm_strain") and self.assured_zero("rigid_shear_strain"))


class Ferromagnet(_Params):
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
        self.regions = None
        self.inter = {"inter_exchange": 0.0, "scale_exchange": 1.0}
        self.enable_demag = True
        self.openbc = False
        # free layer thickness defaults to the grid's thickness, an FP32 product
        # as in the reference (ferromagnet.cpp:38-39)
        self.params["free_layer_thickness"] = float(np.float32(nz) * np.float32(self.cellsize[2]))
        self.enable_zhangli = True
        self.enable_slonc = True
        self.fixed_on_top = True
        # thermal field streams (see oracle.thermal): seed, draws of the time
        # steps (purpose 0) and of the thermal_noise quantity (purpose 1)
        self.noise_seed = 0
        self.noise_draws = [0, 0]
        self.step_draw = 0
        self.m = np.zeros((3,) + self.shape)
        self.m[0] = 1.0
        self._plan = None
        self._nb_cache = {}
        # sublattice of an antiferromagnet: the host holds the AFM and elastic state
        self.host = None
        # elastodynamics of an independent magnet (Magnet::setEnableElastodynamics,
        # magnet.cpp:177-213): displacement u and velocity v, zero when enabled
        self.enable_elastodynamics = False
        self.u = np.zeros((3,) + self.shape)
        self.v = np.zeros((3,) + self.shape)
        # Magnet::enableAsStrayFieldSource / Destination (magnet.cpp:23-24)
        self.stray_source = True
        self.stray_destination = True

    # ----------------------------------------------------------- parameters
    def set(self, name, value):
        super().set(name, value)
        if name == "msat":
            self._plan = None

    def exchange_assured_zero(self):
        """exchange.cu:14-18."""
        return ((self.assured_zero("aex") and self.inter_assured_zero("inter_exchange"))
                or (np.ndim(self.inter["scale_exchange"]) == 0 and self.inter["scale_exchange"] == 0)
                or self.assured_zero("msat"))

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


AFM_DEFAULTS = {"afmex_cell": 0.0, "afmex_nn": 0.0, "latcon": 0.35e-9,
                "C11": 0.0, "C12": 0.0, "C44": 0.0, "eta": 0.0, "rho": 1.0,
                "external_body_force": (0.0, 0.0, 0.0), "rigid_norm_strain": (0.0, 0.0, 0.0),
                "rigid_shear_strain": (0.0, 0.0, 0.0)}


class Antiferromagnet(_Params):
    """Two ferromagnetic sublattices on one grid (antiferromagnet.cpp:15-28):
    host parameters afmex_cell, afmex_nn, latcon (default 0.35 nm), an
    inter-sublattice DMI tensor, and the elastic state of the magnet."""

    def __init__(self, size, cellsize, origin=(0, 0, 0), geometry=None,
                 mastergrid=(0, 0, 0), mastergrid_origin=(0, 0, 0), pbc=(0, 0, 0)):
        self.sub1 = Ferromagnet(size, cellsize, origin, geometry, mastergrid, mastergrid_origin, pbc)
        self.sub2 = Ferromagnet(size, cellsize, origin, geometry, mastergrid, mastergrid_origin, pbc)
        for sub in (self.sub1, self.sub2):
            sub.host = self
        self.shape = self.sub1.shape
        self.cellsize = self.sub1.cellsize
        self.params = dict(AFM_DEFAULTS)
        self.dmi = {k: 0.0 for k in DMI_KEYS}
        self.inter = {"inter_afmex_nn": 0.0, "scale_afmex_nn": 1.0}
        self.enable_elastodynamics = False
        self.u = np.zeros((3,) + self.shape)
        self.v = np.zeros((3,) + self.shape)
        self._plan = None
        self.origin = self.sub1.origin
        self.size = self.sub1.size
        self.stray_source = True
        self.stray_destination = True

    @property
    def sublattices(self):
        return (self.sub1, self.sub2)

    @property
    def regions(self):
        return self.sub1.regions

    @regions.setter
    def regions(self, r):
        # the sublattices share the antiferromagnet's regions
        for sub in (self.sub1, self.sub2):
            sub.regions = None if r is None else np.asarray(r, dtype=int)

    def neighbor(self, d):
        return self.sub1.neighbor(d)

    def geom_mask(self):
        return self.sub1.geom_mask()

    def demag_plan(self):
        if self._plan is None:
            s = self.sub1
            self._plan = _demag.DemagPlan(s.size, s.cellsize, s.mastergrid, s.pbc)
        return self._plan


class World:
    """World with ferromagnets, antiferromagnets and a time solver (FP64).

    Equations, in the order of MumaxWorld::resetTimeSolverEquations
    (mumaxworld.cpp:159-205): (m, LLG torque) per ferromagnet, per
    antiferromagnet sublattice, then (u, v) and (v, a) per magnet with
    elastodynamics.  Only the magnetisations are normalised when a stage
    state is set (NormalizedVariable, variable.cpp:70-87)."""

    def __init__(self, cellsize, mastergrid=(0, 0, 0), pbc=(0, 0, 0)):
        self.cellsize = tuple(float(c) for c in cellsize)
        self.mastergrid = tuple(mastergrid)
        self.pbc = tuple(pbc)
        self.bias = (0.0, 0.0, 0.0)
        self.magnets = []  # independent ferromagnets
        self.afms = []
        self.solver = _rk.TimeSolver(self)
        # set by oracle.relax while a relaxation runs: the ferromagnets whose
        # damping-only torque are the equations, and whether (u, v) stay in
        self.relax_fms = None
        self.relax_elastic = False
        # set by the time solver while a step with thermal noise runs: the dt
        # the noise term is divided by the square root of
        self.noise_dt = None

    def _clock(self):
        return self.solver.time

    def add_ferromagnet(self, size, origin=(0, 0, 0), geometry=None, defer_timestep=False):
        mag = Ferromagnet(size, self.cellsize, origin, geometry, self.mastergrid, (0, 0, 0), self.pbc)
        mag.clock = self._clock
        self.magnets.append(mag)
        if not defer_timestep:
            self.solver.equations_changed()
        return mag

    def add_antiferromagnet(self, size, origin=(0, 0, 0), geometry=None, defer_timestep=False):
        afm = Antiferromagnet(size, self.cellsize, origin, geometry, self.mastergrid, (0, 0, 0), self.pbc)
        for obj in (afm, afm.sub1, afm.sub2):
            obj.clock = self._clock
        self.afms.append(afm)
        if not defer_timestep:
            self.solver.equations_changed()
        return afm

    # --------------------------------------------------------- stray fields
    def _source_M(self, src, m_of):
        """Ms m of a source magnet (both sublattices for an antiferromagnet),
        zero outside its geometry; None when its msat is assured zero
        (strayfield.cpp:88-95)."""
        if isinstance(src, Antiferromagnet):
            if src.sub1.assured_zero("msat") and src.sub2.assured_zero("msat"):
                return None
            M = src.sub1.val("msat") * m_of(src.sub1) + src.sub2.val("msat") * m_of(src.sub2)
        else:
            if src.assured_zero("msat"):
                return None
            M = src.val("msat") * m_of(src)
        return np.where(src.geom_mask()[None], M, 0.0)

    def stray_field_on(self, dst, m_of=None):
        """Sum of the other magnets' stray fields on magnet `dst` (an
        independent ferromagnet or an antiferromagnet), or None when no
        source acts (zeeman.cu:13-34, 70-88)."""
        from . import stray as _stray
        m_of = m_of or (lambda f: f.m)
        if not dst.stray_destination:
            return None
        H = None
        for src in self.magnets + self.afms:
            if src is dst or not src.stray_source:
                continue
            M = self._source_M(src, m_of)
            if M is None:
                continue
            h = _stray.stray_field(M, src.origin, dst.size, dst.origin, self.cellsize, self.mastergrid, self.pbc)
            H = h if H is None else H + h
        return H

    # -------------------------------------------------------------- physics
    def effective_field(self, mag, m, m2=None, u=None, m_of=None):
        """evalEffectiveField (effectivefield.cpp:15-28) of a ferromagnet or a
        sublattice; m2 is the sister sublattice's magnetisation, u the
        displacement of the magnet that owns the elastic state."""
        host = mag.host
        if host is not None and not mag.openbc:
            h = _afm.exchange_field_sublattice(mag, m, m2)
        else:
            h = _fields.exchange_field(mag, m)
        h = h + _fields.anisotropy_field(mag, m)
        hs = self.stray_field_on(self._owner(mag), m_of)
        if not _fields.external_assured_zero(mag, self.bias) or hs is not None:
            h = h + _fields.external_field(mag, m, self.bias)
            if hs is not None:
                h = h + np.where(mag.geom_mask()[None], hs, 0.0)
        if host is None:
            h = h + _fields.dmi_field(mag, m)
        else:
            h = h + _afm.dmi_field_afm(mag, m, m2)
        if mag.enable_demag and not mag.assured_zero("msat"):
            if host is None:
                h = h + _fields.demag_field(mag, m)
            else:
                s1, s2 = host.sub1, host.sub2
                m1_, m2_ = (m, m2) if mag is s1 else (m2, m)
                M = s1.val("msat") * m1_ + s2.val("msat") * m2_
                h = h + host.demag_plan().field(M, np.ones(host.shape), host.geom_mask())
        h = h + self.magnetoelastic_field(mag, m, u)
        if host is not None:
            if not _afm.inhomo_assured_zero(mag):
                h = h + _afm.afm_nn_field(mag, m, m2)
            if not _afm.homo_assured_zero(mag):
                h = h + _afm.afm_site_field(mag, m2)
        return h

    @staticmethod
    def _owner(mag):
        return mag.host if mag.host is not None else mag

    def magnetoelastic_field(self, mag, m, u):
        """evalMagnetoelasticField (magnetoelasticfield.cu:10-42,105-150)."""
        own = self._owner(mag)
        no_b = mag.assured_zero("msat") or (mag.assured_zero("B1") and mag.assured_zero("B2"))
        if not no_b and own.rigid_strain_applied():
            return _elastic.magnetoelastic_field_rigid(mag, m, own.vec("rigid_norm_strain"),
                                                       own.vec("rigid_shear_strain"))
        if no_b or not own.enable_elastodynamics:
            return np.zeros_like(m)
        return _elastic.magnetoelastic_field_dynamic(mag, m, _elastic.strain_tensor(own, u))

    @staticmethod
    def magnetoelastic_assured_zero(fm):
        own = fm.host if fm.host is not None else fm
        no_b = fm.assured_zero("msat") or (fm.assured_zero("B1") and fm.assured_zero("B2"))
        return no_b or not (own.enable_elastodynamics or own.rigid_strain_applied())

    def body_force(self, own, u, ms):
        """evalEffectiveBodyForce (elastodynamics.cu:23-45); ms: magnetisation
        of each ferromagnet of the magnet (one, or the two sublattices)."""
        f = np.zeros((3,) + own.shape)
        if not own.elasticity_assured_zero():
            f = f + _elastic.elastic_force(own, _elastic.stress_tensor(own, _elastic.strain_tensor(own, u)))
        if not own.assured_zero("external_body_force"):
            f = f + own.vec("external_body_force")
        fms = own.sublattices if isinstance(own, Antiferromagnet) else (own,)
        for fm, m in zip(fms, ms):
            if not self.magnetoelastic_assured_zero(fm):
                f = f + _elastic.magnetoelastic_force(fm, m)
        # Field += skips cells outside the geometry (k_addFields,
        # src/core/fieldops.cu:1-14), so they keep the initial zero
        return np.where(own.geom_mask()[None], f, 0.0)

    def acceleration(self, own, u, v, ms):
        """evalElasticAcceleration (elastodynamics.cu:54-84)."""
        a = self.body_force(own, u, ms)
        if own.enable_elastodynamics and not own.assured_zero("eta"):
            a = a + _elastic.elastic_damping(own, v)
        geom = own.geom_mask()
        if not own.assured_zero("rho"):
            rho = own.val("rho")
            with np.errstate(divide="ignore", invalid="ignore"):
                a = np.where(rho != 0, a / rho, 0.0)
        return np.where(geom[None], a, 0.0)

    # ----------------------------------------------------- equation plumbing
    def _equations(self):
        """[(kind, object)]: 'm' for a ferromagnet or sublattice, 'u'/'v' for
        the elastic state of a magnet."""
        if self.relax_fms is not None:
            eqs = [("m", mag) for mag in self.relax_fms]
            if not self.relax_elastic:
                return eqs
        else:
            eqs = [("m", mag) for mag in self.magnets]
            for afm in self.afms:
                eqs += [("m", afm.sub1), ("m", afm.sub2)]
        for own in self.magnets + self.afms:
            if own.enable_elastodynamics:
                eqs += [("u", own), ("v", own)]
        return eqs

    def get_states(self):
        return [getattr(obj, kind) for kind, obj in self._equations()]

    def set_states(self, states):
        for (kind, obj), x in zip(self._equations(), states):
            if kind == "m":
                obj.m = normalize(x, obj.geom_mask())
            else:
                setattr(obj, kind, np.array(x))

    def rhs(self, states):
        eqs = self._equations()
        cur = {(id(obj), kind): x for (kind, obj), x in zip(eqs, states)}

        def m_of(fm):
            return cur.get((id(fm), "m"), fm.m)

        def uv_of(own, kind):
            return cur.get((id(own), kind), getattr(own, kind))

        out = []
        for kind, obj in eqs:
            if kind == "m":
                own = self._owner(obj)
                m2 = None
                if obj.host is not None:
                    m2 = m_of(obj.host.sub2 if obj is obj.host.sub1 else obj.host.sub1)
                h = self.effective_field(obj, m_of(obj), m2, uv_of(own, "u"), m_of)
                if self.relax_fms is not None:
                    t = _fields.damping_torque(obj, m_of(obj), h)
                else:
                    # evalTorque (torque.cu:14-19): LLG plus spin-transfer torque
                    t = _fields.llg_torque(obj, m_of(obj), h) + _fields.spin_transfer_torque(obj, m_of(obj))
                if self.noise_dt is not None and not obj.assured_zero("temperature"):
                    # the stepper's noise term (rungekutta.cpp:95-106)
                    t = t + _thermal.thermal_field(obj, 0, obj.step_draw) / np.sqrt(self.noise_dt)
                out.append(t)
            elif kind == "u":
                out.append(uv_of(obj, "v"))
            else:
                fms = obj.sublattices if isinstance(obj, Antiferromagnet) else (obj,)
                out.append(self.acceleration(obj, uv_of(obj, "u"), uv_of(obj, "v"), [m_of(f) for f in fms]))
        return out

    def geom_masks(self):
        return [obj.geom_mask() for kind, obj in self._equations()]

    def torque(self, mag, m):
        """LLG torque of an independent ferromagnet at magnetisation m."""
        return _fields.llg_torque(mag, m, self.effective_field(mag, m, None, mag.u))
