"""Antiferromagnets, elastodynamics and magnetoelastic coupling on the GPU.

Two kinds of check, both through the public Python API over the C ABI:

* parity with the FP64 oracle on the same seeded inputs (oracle/afm.py,
  oracle/elastic.py, oracle/model.py): per field term, per elastic quantity
  and over fixed time steps of the coupled (m1, m2, u, v) system;
  tolerance max |x - x_oracle| / max |x_oracle| <= 1e-5 for single
  evaluations (FP32 against FP64), stated per test for time stepping;
* the closed forms the reference's own tests use (test/test_exchange_afm.py,
  test_strain.py, test_full_magnetization.py, test_is_sublattice.py,
  test_magnetoelasticity_afm.py), with the reference's tolerances.
"""

import math
import warnings
import zlib

import numpy as np
import pytest

from conftest import rel_max_err
from helpers import random_unit
from oracle import afm as oafm
from oracle import elastic as oel
from oracle import fields as of
from oracle import model as omodel

import mumaxplus as mp

pytestmark = pytest.mark.gpu

TOL = 1e-5
SIZE = (16, 8, 3)
CS = (2e-9, 2.5e-9, 3e-9)


def _f32(v):
    return v.astype(np.float32) if isinstance(v, np.ndarray) else v


def _f64(v):
    return np.asarray(v, dtype=np.float64) if isinstance(v, np.ndarray) else v


def _set_both(prod, orc, key, value):
    """Set one parameter on a product magnet and its oracle twin."""
    value = _f32(value)
    setattr(prod, key, value)
    orc.set("bias" if key == "bias_magnetic_field" else key, _f64(value))


class AfmPair:
    """The same antiferromagnet in the product and in the oracle."""

    def __init__(self, rng, size=SIZE, cs=CS, geometry=None, mastergrid=(0, 0, 0), pbc=(0, 0, 0), openbc=False,
                 demag=True, afm=None, subs=(None, None), inter_dmi=None, sub_dmi=(None, None), elastic=None,
                 uniform_m=False):
        nx, ny, nz = size
        shape = (nz, ny, nx)
        self.world = mp.World(cs, pbc_repetitions=tuple(pbc), mastergrid=mp.Grid(tuple(mastergrid)))
        self.afm = mp.Antiferromagnet(self.world, mp.Grid(size), geometry=geometry)
        self.ow = omodel.World(cs, mastergrid=tuple(mastergrid), pbc=tuple(pbc))
        self.oa = self.ow.add_antiferromagnet(size, geometry=geometry, defer_timestep=True)
        self.subs = (self.afm.sub1, self.afm.sub2)
        self.osubs = (self.oa.sub1, self.oa.sub2)
        with warnings.catch_warnings():
            warnings.simplefilter("ignore", UserWarning)  # sign warnings of afmex / B1 / B2
            for k, v in (afm or {}).items():
                _set_both(self.afm, self.oa, k, v)
            for k, v in (inter_dmi or {}).items():
                _set_both(self.afm.dmi_tensor, self.oa, k, v)
            for sub, osub, p, d in zip(self.subs, self.osubs, subs, sub_dmi):
                sub.enable_demag = demag
                osub.enable_demag = demag
                sub.enable_openbc = openbc
                osub.openbc = openbc
                for k, v in (p or {}).items():
                    _set_both(sub, osub, k, v)
                for k, v in (d or {}).items():
                    _set_both(sub.dmi_tensor, osub, k, v)
        for sub, osub in zip(self.subs, self.osubs):
            m0 = np.array([0.3, -0.2, 0.9]) if uniform_m else random_unit(rng, shape)
            sub.magnetization = m0.astype(np.float32)
            osub.set_m(sub.magnetization.get().astype(np.float64))
        if elastic is not None:
            self.afm.enable_elastodynamics = True
            self.oa.enable_elastodynamics = True
            for k, v in elastic.items():
                if k in ("u", "v"):
                    continue
                _set_both(self.afm, self.oa, k, v)
            self.sync_elastic(elastic.get("u"), elastic.get("v"))
        self.ow.solver.equations_changed()

    def sync_elastic(self, u=None, v=None):
        if u is not None:
            self.afm.elastic_displacement = u.astype(np.float32)
        if v is not None:
            self.afm.elastic_velocity = v.astype(np.float32)
        self.oa.u = self.afm.elastic_displacement.get().astype(np.float64)
        self.oa.v = self.afm.elastic_velocity.get().astype(np.float64)

    def oracle_ms(self):
        return [s.m for s in self.osubs]

    def oracle_demag(self):
        s1, s2 = self.osubs
        M = s1.val("msat") * s1.m + s2.val("msat") * s2.m
        return self.oa.demag_plan().field(M, np.ones(self.oa.shape), self.oa.geom_mask())


def _sub_params(rng, shape, nonuniform):
    p = {"msat": 5.66e5, "aex": 2.48e-12, "alpha": 0.1, "ku1": 6.4e4, "anisU": (0.0, 0.3, 1.0),
         "bias_magnetic_field": (0.05, -0.02, 0.3)}
    if nonuniform:
        p["msat"] = 4e5 + 3e5 * rng.random(shape)
        p["aex"] = 2e-12 + 1e-12 * rng.random(shape)
    return p


def _afm_params(rng, shape, nonuniform):
    p = {"afmex_cell": -4e-12, "afmex_nn": -5e-12, "latcon": 0.35e-9}
    if nonuniform:
        p["afmex_nn"] = -4e-12 - 2e-12 * rng.random(shape)
        p["afmex_cell"] = -3e-12 - 2e-12 * rng.random(shape)
    return p


CASES = {
    "open": dict(),
    "geometry": dict(geometry=True),
    "openbc": dict(openbc=True),
    "pbc": dict(mastergrid=(16, 8, 0), pbc=(1, 1, 0)),
    "nonuniform": dict(nonuniform=True, geometry=True),
}


def _build_case(name, elastic=False):
    rng = np.random.default_rng(zlib.crc32(name.encode()) + (7 if elastic else 0))
    kw = dict(CASES[name])
    nonuniform = kw.pop("nonuniform", False)
    nx, ny, nz = SIZE
    shape = (nz, ny, nx)
    if kw.pop("geometry", False):
        kw["geometry"] = rng.random(shape) > 0.2
    el = None
    if elastic:
        el = {"C11": 283e9, "C12": 166e9, "C44": 58e9, "rho": 8900.0, "eta": 1e12,
              "external_body_force": (1e14, -2e14, 5e13),
              "u": 1e-12 * rng.normal(size=(3,) + shape), "v": rng.normal(size=(3,) + shape)}
    sub_p = [_sub_params(rng, shape, nonuniform), _sub_params(rng, shape, nonuniform)]
    if elastic:
        for p in sub_p:
            p["B1"] = -8.8e6
            p["B2"] = -8.8e6 if not nonuniform else -8e6 - 1e6 * rng.random(shape)
    pair = AfmPair(rng, afm=_afm_params(rng, shape, nonuniform), subs=sub_p,
                   inter_dmi={"xxz": 1e-3, "yyz": 1e-3, "xyz": 4e-4, "yxz": -4e-4},
                   sub_dmi=({"xxz": 2e-3, "yyz": 2e-3}, {"xxz": -1e-3, "yyz": 1.5e-3}), elastic=el, **kw)
    return pair


@pytest.mark.parametrize("case", sorted(CASES))
def test_sublattice_field_terms_match_oracle(case):
    """Every term of the sublattice effective field, the torque and the
    energy densities of both sublattices against the FP64 oracle."""
    p = _build_case(case)
    hd = p.oracle_demag()
    for i, (sub, osub) in enumerate(zip(p.subs, p.osubs)):
        m, m2 = osub.m, p.osubs[1 - i].m
        want = {
            "exchange_field": (of.exchange_field(osub, m) if osub.openbc
                               else oafm.exchange_field_sublattice(osub, m, m2)),
            "dmi_field": oafm.dmi_field_afm(osub, m, m2),
            "homogeneous_exchange_field": oafm.afm_site_field(osub, m2),
            "inhomogeneous_exchange_field": oafm.afm_nn_field(osub, m, m2),
            "anisotropy_field": of.anisotropy_field(osub, m),
            "external_field": of.external_field(osub, m, p.ow.bias),
            "demag_field": hd,
        }
        for term, w in want.items():
            err = rel_max_err(getattr(sub, term).eval(), w)
            assert err < TOL, (case, i, term, err)
        heff = p.ow.effective_field(osub, m, m2, p.oa.u)
        assert rel_max_err(sub.effective_field.eval(), heff) < TOL
        assert rel_max_err(sub.torque.eval(), of.llg_torque(osub, m, heff)) < TOL
        ms = osub.val("msat")
        for term, field in (("homogeneous_exchange_energy_density", want["homogeneous_exchange_field"]),
                            ("inhomogeneous_exchange_energy_density", want["inhomogeneous_exchange_field"])):
            e = -0.5 * ms * np.sum(m * field, axis=0)
            assert rel_max_err(getattr(sub, term).eval()[0], e) < 1e-4, (case, i, term)


def test_neel_vector_and_full_magnetization():
    p = _build_case("nonuniform")
    s1, s2 = p.osubs
    M1, M2 = s1.val("msat") * s1.m, s2.val("msat") * s2.m
    geom = p.oa.geom_mask()
    with np.errstate(divide="ignore", invalid="ignore"):
        tot = s1.val("msat") + s2.val("msat")
        neel = np.where(geom[None] & (tot > 0), (M1 - M2) / tot, 0.0)
    assert rel_max_err(p.afm.full_magnetization.eval(), np.where(geom[None], M1 + M2, 0.0)) < TOL
    assert rel_max_err(p.afm.neel_vector.eval(), neel) < TOL
    assert rel_max_err(p.afm.sub1.full_magnetization.eval(), np.where(geom[None], M1, 0.0)) < TOL


@pytest.mark.parametrize("case", sorted(CASES))
def test_elastic_quantities_match_oracle(case):
    """Strain, stress, elastic / magnetoelastic / body forces, damping,
    acceleration, magnetoelastic field, energies and energy flux of an
    antiferromagnet with elastodynamics against the FP64 oracle."""
    p = _build_case(case, elastic=True)
    oa = p.oa
    u, v = oa.u, oa.v
    strain = oel.strain_tensor(oa, u)
    stress = oel.stress_tensor(oa, strain)
    geom = oa.geom_mask()[None]
    checks = {
        "strain_tensor": strain,
        "stress_tensor": stress,
        "elastic_force": oel.elastic_force(oa, stress),
        "effective_body_force": p.ow.body_force(oa, u, p.oracle_ms()),
        "elastic_damping": oel.elastic_damping(oa, v),
        "elastic_acceleration": p.ow.acceleration(oa, u, v, p.oracle_ms()),
        "kinetic_energy_density": np.where(geom, 0.5 * oa.val("rho") * np.sum(v * v, axis=0), 0.0),
        "elastic_energy_density": np.where(
            geom, 0.5 * np.sum(stress[:3] * strain[:3], axis=0) + np.sum(stress[3:] * strain[3:], axis=0), 0.0),
    }
    sig = np.empty((3, 3) + oa.shape)
    for i in range(3):
        for j in range(3):
            sig[i, j] = stress[i if i == j else i + j + 2]
    checks["poynting_vector"] = np.where(geom, -np.einsum("ij...,j...->i...", sig, v), 0.0)
    for q, want in checks.items():
        err = rel_max_err(getattr(p.afm, q).eval(), want)
        assert err < TOL, (case, q, err)
    for i, (sub, osub) in enumerate(zip(p.subs, p.osubs)):
        assert rel_max_err(sub.magnetoelastic_field.eval(), p.ow.magnetoelastic_field(osub, osub.m, u)) < TOL
        assert rel_max_err(sub.magnetoelastic_force.eval(), oel.magnetoelastic_force(osub, osub.m)) < TOL
        m2 = p.osubs[1 - i].m
        heff = p.ow.effective_field(osub, osub.m, m2, u)
        assert rel_max_err(sub.effective_field.eval(), heff) < TOL, (case, i)


def test_ferromagnet_elastodynamics_and_rigid_strain():
    """An independent ferromagnet: elastic quantities with elastodynamics,
    then the rigid-strain magnetoelastic field without it."""
    rng = np.random.default_rng(11)
    nx, ny, nz = SIZE
    shape = (nz, ny, nx)
    geom = rng.random(shape) > 0.15
    world = mp.World(CS, pbc_repetitions=(1, 0, 0), mastergrid=mp.Grid((16, 0, 0)))
    mag = mp.Ferromagnet(world, mp.Grid(SIZE), geometry=geom)
    ow = omodel.World(CS, mastergrid=(16, 0, 0), pbc=(1, 0, 0))
    om = ow.add_ferromagnet(SIZE, geometry=geom, defer_timestep=True)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", UserWarning)
        for k, val in {"msat": 1.2e6, "aex": 1e-11, "alpha": 0.05, "B1": -4.4e6,
                       "B2": -8.8e6 - 1e6 * rng.random(shape)}.items():
            _set_both(mag, om, k, val)
    mag.magnetization = random_unit(rng, shape).astype(np.float32)
    om.set_m(mag.magnetization.get().astype(np.float64))
    mag.enable_elastodynamics = True
    om.enable_elastodynamics = True
    for k, val in {"C11": 283e9, "C12": 58e9 + 1e9 * rng.random(shape), "C44": 166e9, "rho": 8900.0,
                   "eta": 3e11}.items():
        _set_both(mag, om, k, val)
    mag.elastic_displacement = (1e-12 * rng.normal(size=(3,) + shape)).astype(np.float32)
    mag.elastic_velocity = rng.normal(size=(3,) + shape).astype(np.float32)
    om.u = mag.elastic_displacement.get().astype(np.float64)
    om.v = mag.elastic_velocity.get().astype(np.float64)
    strain = oel.strain_tensor(om, om.u)
    stress = oel.stress_tensor(om, strain)
    assert rel_max_err(mag.strain_tensor.eval(), strain) < TOL
    assert rel_max_err(mag.stress_tensor.eval(), stress) < TOL
    assert rel_max_err(mag.elastic_acceleration.eval(), ow.acceleration(om, om.u, om.v, [om.m])) < TOL
    assert rel_max_err(mag.magnetoelastic_field.eval(), ow.magnetoelastic_field(om, om.m, om.u)) < TOL
    heff = ow.effective_field(om, om.m, None, om.u)
    assert rel_max_err(mag.effective_field.eval(), heff) < TOL
    assert rel_max_err(mag.torque.eval(), of.llg_torque(om, om.m, heff)) < TOL
    # kinetic and elastic energy densities enter the total energy density
    e_total = (mag.total_energy_density.eval() - mag.kinetic_energy_density.eval()
               - mag.elastic_energy_density.eval())
    e_mag = sum(getattr(mag, t).eval() for t in ("exchange_energy_density", "magnetoelastic_energy_density",
                                                 "demag_energy_density"))
    assert rel_max_err(e_total, e_mag) < 1e-4
    # rigid strain (no elastodynamics): magnetoelasticfield.cu:72-103
    with pytest.raises(Exception):
        mag.rigid_norm_strain = (1e-3, 0, 0)
    mag.enable_elastodynamics = False
    om.enable_elastodynamics = False
    _set_both(mag, om, "rigid_norm_strain", (1e-3, -2e-3, 5e-4))
    _set_both(mag, om, "rigid_shear_strain", (2e-4, 0.0, -1e-3))
    with pytest.raises(ValueError):
        mag.enable_elastodynamics = True
    assert rel_max_err(mag.magnetoelastic_field.eval(), ow.magnetoelastic_field(om, om.m, None)) < TOL
    assert rel_max_err(mag.magnetoelastic_force.eval(), oel.magnetoelastic_force(om, om.m)) < TOL
    assert np.all(mag.strain_tensor.eval() == 0)
    with pytest.raises(ValueError):
        mag.elastic_displacement


@pytest.mark.parametrize("method", ["DormandPrince", "Fehlberg"])
def test_coupled_fixed_steps_match_oracle(method):
    """(m1, m2, u, v) of an antiferromagnet with afm exchange, DMI, demag,
    magnetoelastic coupling and elastodynamics, 40 fixed steps of 2 fs
    against the FP64 oracle: magnetisations within 2e-4 per component,
    displacement and velocity within 2e-3 of their maxima."""
    rng = np.random.default_rng(21)
    nx, ny, nz = 16, 8, 2
    shape = (nz, ny, nx)
    sub = {"msat": 5.66e5, "aex": 2.48e-12, "alpha": 0.1, "B1": -8.8e6, "B2": -8.8e6, "ku1": 6.4e4,
           "anisU": (0.0, 0.0, 1.0)}
    p = AfmPair(rng, size=(nx, ny, nz), cs=(2e-9, 2e-9, 2e-9), mastergrid=(nx, ny, 0), pbc=(1, 1, 0),
                afm={"afmex_cell": -4e-12, "afmex_nn": -5e-12},
                subs=(dict(sub), dict(sub)), inter_dmi={"xxz": 5e-4, "yyz": 5e-4},
                elastic={"C11": 283e9, "C12": 166e9, "C44": 58e9, "rho": 8900.0,
                         "u": 1e-11 * rng.normal(size=(3,) + shape), "v": 10 * rng.normal(size=(3,) + shape)})
    ts = p.world.timesolver
    ts.adaptive_timestep = False
    ts.set_method(method)
    ts.timestep = 2e-15
    p.ow.solver.adaptive = False
    p.ow.solver.set_method(method)
    p.ow.solver.dt = 2e-15
    ts.steps(40)
    p.ow.solver.steps(40)
    for sub_, osub in zip(p.subs, p.osubs):
        assert np.abs(sub_.magnetization.get() - osub.m).max() < 2e-4
    assert rel_max_err(p.afm.elastic_displacement.get(), p.oa.u) < 2e-3
    assert rel_max_err(p.afm.elastic_velocity.get(), p.oa.v) < 2e-3
    assert abs(ts.time - p.ow.solver.time) < 1e-25


@pytest.mark.parametrize("options", [{"split_elastic": 1}, {"split_demag": 1}, {"split_elastic": 1, "split_demag": 1}])
def test_streams_beside_the_demag_passes_change_nothing(options):
    """The strain and acceleration launches (world option split_elastic) and
    the sublattice fields without demag (split_demag) run on other streams
    beside the demag passes: adaptive Dormand-Prince steps of the coupled
    antiferromagnet give the same m1, m2, u, v and time bit for bit with the
    options on and off."""
    out = []
    for opts in ({}, options):
        rng = np.random.default_rng(22)
        nx, ny, nz = 32, 16, 1
        shape = (nz, ny, nx)
        sub = {"msat": 5.66e5, "aex": 2.48e-12, "alpha": 0.05, "B1": -8.8e6, "B2": -8.8e6}
        p = AfmPair(rng, size=(nx, ny, nz), cs=(2e-9, 2e-9, 2e-9), mastergrid=(nx, ny, 0), pbc=(1, 1, 0),
                    afm={"afmex_cell": -4e-12}, subs=(dict(sub), dict(sub)), inter_dmi={"xxz": 5e-4, "yyz": 5e-4},
                    elastic={"C11": 283e9, "C12": 166e9, "C44": 58e9, "rho": 8900.0, "eta": 1e10,
                             "u": 1e-11 * rng.normal(size=(3,) + shape), "v": 10 * rng.normal(size=(3,) + shape)})
        for k, v in opts.items():
            p.world._impl.set_option(k, v)
            assert p.world._impl.get_option(k) == v
        ts = p.world.timesolver
        ts.set_method("DormandPrince")
        ts.steps(25)
        out.append(([s_.magnetization.get() for s_ in p.subs], p.afm.elastic_displacement.get(),
                    p.afm.elastic_velocity.get(), ts.time))
    (m_a, u_a, v_a, t_a), (m_b, u_b, v_b, t_b) = out
    for x, y in zip(m_a, m_b):
        assert np.array_equal(x, y)
    assert np.array_equal(u_a, u_b) and np.array_equal(v_a, v_b)
    assert t_a == t_b


def test_sensible_timestep_counts_elastic_equations():
    """The initial time step is sensible_factor / the largest rhs norm over
    all equations, the elastic accelerations included (timesolver.cpp:37-47)."""
    p = _build_case("open", elastic=True)
    got = p.world.timesolver.sensible_timestep
    rhs = p.ow.rhs(p.ow.get_states())
    mx = max(float(np.sqrt((r ** 2).sum(axis=0)).max()) for r in rhs)
    assert abs(got - 0.01 / mx) <= 1e-4 * (0.01 / mx)


# ------------------------------------------------ closed forms of the reference's tests

def _laplacian_neumann(m, cs):
    """Sum over axes of (m[i+1] - m[i]) + (m[i-1] - m[i]) over existing
    neighbours, divided by the cell size squared (test_exchange_afm.py:7-31)."""
    h = np.zeros_like(m)
    for axis, c in zip((3, 2, 1), cs):
        n = m.shape[axis]
        for step in (1, -1):
            diff = np.roll(m, -step, axis=axis) - m
            sl = [slice(None)] * 4
            sl[axis] = slice(0, n - 1) if step == 1 else slice(1, n)
            h[tuple(sl)] += diff[tuple(sl)] / c ** 2
    return h


def test_afm_exchange_closed_forms():
    """test/test_exchange_afm.py: sublattice exchange, 4 A_c m2 / (a^2 Ms),
    A_nn Laplacian(m2) / Ms on a 16x16x4 grid, relative error < 1e-3."""
    cs = (1e3, 2e3, 3e3)
    world = mp.World(cs)
    magnet = mp.Antiferromagnet(world, mp.Grid((16, 16, 4)))
    magnet.aex = 3.2e7
    magnet.msat = 5.4
    for sub in magnet.sublattices:
        want = 2 * 3.2e7 * _laplacian_neumann(sub.magnetization.get().astype(np.float64), cs) / 5.4
        got = sub.exchange_field.eval()
        assert np.max(np.abs(got - want) / np.abs(want)) < 1e-3
    world = mp.World(cs)
    magnet = mp.Antiferromagnet(world, mp.Grid((16, 16, 4)))
    magnet.msat = 5.4
    with warnings.catch_warnings():
        warnings.simplefilter("error", UserWarning)
        magnet.afmex_cell = -10e4
        magnet.afmex_nn = -10e4
    latcon = magnet.latcon.average()[0]
    for i, sub in enumerate(magnet.sublattices):
        other = magnet.other_sublattice(sub)
        assert other == magnet.sublattices[1 - i]
        m2 = other.magnetization.get().astype(np.float64)
        want = 4 * -10e4 * m2 / (latcon ** 2 * 5.4)
        assert np.max(np.abs(sub.homogeneous_exchange_field() - want) / np.abs(want)) < 1e-3
        want = -10e4 * _laplacian_neumann(m2, cs) / 5.4
        assert np.max(np.abs(sub.inhomogeneous_exchange_field() - want) / np.abs(want)) < 1e-3


@pytest.mark.parametrize("i,j", [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2), (1, 0), (2, 0), (2, 1)])
def test_strain_of_sine_displacement(i, j):
    """test/test_strain.py: u_j = A sin(k x_i) on a periodic 1-D magnet along
    i; e_ij = k A cos(k x_i) (halved off the diagonal), error < 1e-4 of max."""
    N, A, P = 128, 4.0, 2
    cs = (1.5e-9, 2e-9, 2.5e-9)
    mg, size, reps = [0, 0, 0], [1, 1, 1], [0, 0, 0]
    mg[i], size[i], reps[i] = N, N, 1
    world = mp.World(cs, mastergrid=mp.Grid(tuple(mg)), pbc_repetitions=tuple(reps))
    magnet = mp.Ferromagnet(world, mp.Grid(tuple(size)))
    magnet.enable_elastodynamics = True
    k = P * 2 * math.pi / (N * cs[i])

    def disp(x, y, z):
        u = [0.0, 0.0, 0.0]
        u[j] = A * math.sin(k * (x, y, z)[i])
        return tuple(u)

    magnet.elastic_displacement = disp
    got = magnet.strain_tensor.eval()
    want = np.zeros_like(got, dtype=np.float64)
    c = (A * k * np.cos(k * cs[i] * np.arange(N))).reshape([N if a == i else 1 for a in (2, 1, 0)])
    if i == j:
        want[j] = c
    else:
        want[2 + i + j] = 0.5 * c
    err = np.max(np.linalg.norm(got - want, axis=0)) / np.max(np.abs(want))
    assert err < 1e-4


def test_full_magnetization_sums_sublattices():
    """test/test_full_magnetization.py (31x120x60 grid, random msat)."""
    rng = np.random.default_rng(3)
    nx, ny, nz = 31, 120, 60
    world = mp.World(cellsize=(1e-9, 1e-9, 1e-9))
    magnet = mp.Antiferromagnet(world, mp.Grid((nx, ny, nz)))
    magnet.sub1.msat = 800e3 * rng.normal(loc=1, scale=0.05, size=(1, nz, ny, nx))
    magnet.sub2.msat = 750e3 * rng.normal(loc=1, scale=0.05, size=(1, nz, ny, nx))
    want = magnet.sub1.full_magnetization.eval() + magnet.sub2.full_magnetization.eval()
    got = magnet.full_magnetization.eval()
    err = np.max(np.linalg.norm(got - want, axis=0) / np.linalg.norm(want, axis=0))
    assert err < 1e-5


def test_is_sublattice_and_names():
    world = mp.World(cellsize=(3e-9, 2e-9, 1e-9))
    fm = mp.Ferromagnet(world, mp.Grid((16, 8, 2)))
    afm = mp.Antiferromagnet(world, mp.Grid((8, 16, 2), origin=(0, 8, 0)), name="afm")
    fm.enable_as_stray_field_source = False
    afm.enable_as_stray_field_source = False
    assert not fm.is_sublattice
    assert afm.sub1.is_sublattice and afm.sub2.is_sublattice
    assert afm.sub1.name == "afm:sublattice_1" and afm.sub2.name == "afm:sublattice_2"
    assert world.get_antiferromagnet("afm") == afm
    assert set(world.antiferromagnets) == {"afm"}
    with pytest.raises(KeyError):
        world.get_antiferromagnet("nope")
    with pytest.raises(ValueError):
        afm.sub1.enable_elastodynamics = True
    with pytest.raises(IndexError):  # overlaps (mumaxworld.cpp:42-47)
        mp.Antiferromagnet(world, mp.Grid((8, 16, 2)), name="other")
    with pytest.raises(RuntimeError):  # the name is taken (mumaxworld.cpp:50-53)
        mp.Antiferromagnet(world, mp.Grid((8, 16, 2), origin=(40, 0, 0)), name="afm")


def test_afm_magnetoelastic_sum_rules():
    """test/test_magnetoelasticity_afm.py: the body force holds both
    sublattices' magnetoelastic forces, each sublattice's magnetoelastic
    field is the closed form of its own m and the host strain, and the total
    energy density is the sublattices' plus the kinetic and elastic ones."""
    rng = np.random.default_rng(5)
    nx, ny, nz = 256, 128, 1
    msat, B1, B2 = 1.2e6, -4.4e6, -8.8e6
    world = mp.World((1.5e-9, 2e-9, 2.5e-9))
    magnet = mp.Antiferromagnet(world, mp.Grid((nx, ny, nz)))
    magnet.msat = msat
    magnet.enable_elastodynamics = True
    magnet.C11 = 283e9
    magnet.C12 = 58e9
    magnet.C44 = 166e9
    magnet.B1 = B1
    magnet.B2 = B2
    magnet.elastic_displacement = 1e-12 * rng.random((3, nz, ny, nx))
    magnet.elastic_velocity = rng.random((3, nz, ny, nx))
    for sub in magnet.sublattices:
        sub.magnetization = rng.random((3, nz, ny, nx))

    def semirel(got, want):
        return np.max(np.linalg.norm(got - want, axis=0)) / np.max(np.abs(want))

    force = (magnet.elastic_force.eval() + magnet.sub1.magnetoelastic_force.eval()
             + magnet.sub2.magnetoelastic_force.eval())
    assert semirel(magnet.effective_body_force.eval(), force) < 1e-5
    assert not np.all(magnet.sub1.magnetoelastic_force.eval() == magnet.sub2.magnetoelastic_force.eval())
    strain = magnet.strain_tensor.eval()
    for sub in magnet.sublattices:
        m = sub.magnetization.eval()
        want = np.zeros_like(m)
        for i in range(3):
            i1, i2 = (i + 1) % 3, (i + 2) % 3
            want[i] = -2 / msat * (B1 * strain[i] * m[i] + B2 * (strain[i + i1 + 2] * m[i1]
                                                                 + strain[i + i2 + 2] * m[i2]))
        assert semirel(sub.magnetoelastic_field.eval(), want) < 1e-5
    edens = (magnet.sub1.total_energy_density.eval() + magnet.sub2.total_energy_density.eval()
             + magnet.kinetic_energy_density.eval() + magnet.elastic_energy_density.eval())
    assert semirel(magnet.total_energy_density.eval(), edens) < 1e-5
    assert magnet.total_energy.eval() == pytest.approx(
        magnet.total_energy_density.average()[0] * nx * ny * nz * 1.5e-9 * 2e-9 * 2.5e-9, rel=1e-6)


@pytest.mark.parametrize("method,adaptive,elastic,geometry",
                         [("DormandPrince", True, True, False), ("DormandPrince", False, True, True),
                          ("Fehlberg", True, True, False), ("DormandPrince", True, False, False)])
def test_stage_states_formed_in_the_field_and_acceleration_launches(method, adaptive, elastic, geometry):
    """World option fuse_stages (default on): for a lone antiferromagnet the
    field launch of each stage but the last also writes both sublattices'
    next stage states, and the acceleration launch those of u and v (the
    displacement's next derivative included), with the combination launch's
    arithmetic and term order; m1, m2, u, v and the time equal the separate
    combination's bit for bit, over fixed and adaptive steps, with an FSAL
    tableau and without, with a geometry mask, and without elastodynamics.
    With a fixed step and an FSAL tableau the step's last evaluation also
    forms the next step's first stage state, so no combination launch runs
    at all (the first step evaluates its first stage)."""
    out = []
    for fuse in (0, 1):
        rng = np.random.default_rng(31)
        nx, ny, nz = 24, 20, 2
        shape = (nz, ny, nx)
        geom = None
        if geometry:
            x = (np.arange(nx) + 0.5) / nx - 0.5
            y = (np.arange(ny) + 0.5) / ny - 0.5
            geom = np.broadcast_to((x[None, :] ** 2 + y[:, None] ** 2) <= 0.25, shape).copy()
        sub = {"msat": 5.66e5, "aex": 2.48e-12, "alpha": 0.05, "B1": -8.8e6, "B2": -8.8e6}
        el = None
        if elastic:
            el = {"C11": 283e9, "C12": 166e9, "C44": 58e9, "rho": 8900.0, "eta": 1e10,
                  "u": 1e-11 * rng.normal(size=(3,) + shape), "v": 10 * rng.normal(size=(3,) + shape)}
        p = AfmPair(rng, size=(nx, ny, nz), cs=(2e-9, 2e-9, 2e-9), geometry=geom, afm={"afmex_cell": -4e-12},
                    subs=(dict(sub), dict(sub)), inter_dmi={"xxz": 5e-4, "yyz": 5e-4}, elastic=el)
        p.world._impl.set_option("fuse_stages", fuse)
        ts = p.world.timesolver
        ts.set_method(method)
        if not adaptive:
            ts.adaptive_timestep = False
            ts.timestep = 2e-14
        p.world._impl.set_profiling(True)
        ts.steps(12)
        launches = p.world._impl.profile_report()
        p.world._impl.set_profiling(False)
        assert ("k_field_pair_next" in launches) == bool(fuse), launches.keys()
        if elastic:
            assert ("k_elastic_next" in launches) == bool(fuse), launches.keys()
        if fuse and not adaptive and method == "DormandPrince":
            assert "k_combine" not in launches, launches["k_combine"]
        res = [s_.magnetization.get() for s_ in p.subs]
        if elastic:
            res += [p.afm.elastic_displacement.get(), p.afm.elastic_velocity.get()]
        out.append((res, ts.time, ts.timestep))
    (ra, ta, da), (rb, tb, db) = out
    for x, y in zip(ra, rb):
        assert np.array_equal(x, y)
    assert ta == tb and da == db
