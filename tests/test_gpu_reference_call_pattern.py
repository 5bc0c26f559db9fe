"""The binding driven the way the reference's own package drives its pybind11 module.

The reference's mumaxplus/*.py never computes anything itself: each class holds
an `_impl` from `_mumaxpluscpp` and every quantity is a module-level factory
called on it, e.g. `FieldQuantity(_cpp.exchange_field(self._impl))`
(ferromagnet.py:766).  These tests make the same calls on this build's
binding, with no use of this build's package, and check the values against
the FP64 oracle (per-term 1e-5) and against closed forms:

* World(cellsize, Grid, pbc_repetitions) (world.py:50), add_ferromagnet /
  add_antiferromagnet(grid_impl, geometry, regions, name) (magnet.py:47),
  Parameter.set / uniform_value / isinstance checks (parameter.py:93-211),
  Variable.set / get (variable.py), dmi_tensor elements (dmitensor.py);
* every module-level factory of wrap_ferromagnet.cpp, wrap_antiferromagnet.cpp
  and wrap_magnet.cpp, with the quantity members the package reads
  (name, unit, ncomp, grid, system, eval, exec, average, get_rgb);
* TimeSolver driven as timesolver.py drives it, against oracle steps.
"""

import math

import numpy as np
import pytest

from conftest import rel_max_err
from helpers import random_unit
from oracle import elastic as oel
from oracle import fields as of
from oracle import model as omodel

import mumaxplus._mumaxpluscpp as _cpp

pytestmark = pytest.mark.gpu

TOL = 1e-5
SIZE = (16, 8, 2)
CS = (2e-9, 2.5e-9, 3e-9)
FM_PARAMS = {"msat": 8e5, "aex": 1.3e-11, "alpha": 0.1, "ku1": 4e5, "anisU": (0.3, 0.2, 1.0),
             "bias_magnetic_field": (0.01, -0.02, 0.03)}


def _fm_pair(seed=0, elastic=False):
    rng = np.random.default_rng(seed)
    nx, ny, nz = SIZE
    world = _cpp.World(CS, _cpp.Grid((0, 0, 0)), (0, 0, 0))
    world.bias_magnetic_field = (0.0, 0.0, 0.02)
    fm = world.add_ferromagnet(_cpp.Grid(SIZE), None, None, "film")
    for k, v in FM_PARAMS.items():
        getattr(fm, k).set(v)
    fm.dmi_tensor.xxz.set(3e-3)
    fm.dmi_tensor.yyz.set(3e-3)
    fm.magnetization.set(random_unit(rng, (nz, ny, nx)).astype(np.float32))
    ow = omodel.World(CS)
    ow.bias = (0.0, 0.0, 0.02)
    om = ow.add_ferromagnet(SIZE, defer_timestep=True)
    for k, v in FM_PARAMS.items():
        om.set("bias" if k == "bias_magnetic_field" else k, v)
    om.set("xxz", 3e-3)
    om.set("yyz", 3e-3)
    if elastic:
        fm.enable_elastodynamics = True
        om.enable_elastodynamics = True
        for k, v in {"C11": 283e9, "C12": 166e9, "C44": 58e9, "rho": 8900.0, "B1": -8.8e6, "B2": -8.8e6}.items():
            getattr(fm, k).set(v)
            om.set(k, v)
        u = (1e-11 * rng.normal(size=(3, nz, ny, nx))).astype(np.float32)
        v = (10 * rng.normal(size=(3, nz, ny, nx))).astype(np.float32)
        fm.elastic_displacement.set(u)
        fm.elastic_velocity.set(v)
        om.u, om.v = u.astype(np.float64), v.astype(np.float64)
    om.set_m(fm.magnetization.get().astype(np.float64))
    ow.solver.equations_changed()
    return world, fm, ow, om


def test_world_and_magnet_surface():
    world, fm, ow, om = _fm_pair()
    assert world.cellsize == pytest.approx(CS)
    assert world.get_ferromagnet("film") == fm and world.get_ferromagnet("none") is None
    assert set(world.ferromagnets) == {"film"} and world.antiferromagnets == {}
    assert fm.name == "film" and fm.world is world and not fm.is_sublattice
    assert fm.system.grid.size == SIZE and fm.system.cellsize == pytest.approx(CS)
    assert fm.system.geometry.shape == SIZE[::-1] and fm.system.geometry.all()
    assert isinstance(fm.msat, _cpp.Parameter) and isinstance(fm.anisU, _cpp.VectorParameter)
    assert fm.msat.is_uniform and not fm.msat.is_dynamic and fm.msat.uniform_value == pytest.approx(8e5)
    assert fm.anisU.uniform_value == pytest.approx((0.3, 0.2, 1.0))
    assert fm.magnetization.get().shape == (3,) + SIZE[::-1]
    assert world.bounding_grid.size == SIZE
    world.set_pbc((1, 0, 0))
    assert world.pbc_repetitions == (1, 0, 0) and world.mastergrid.size == (SIZE[0], 0, 0)
    world.unset_pbc()
    assert world.mastergrid.size == (0, 0, 0)


FM_FIELD_TERMS = {
    "exchange_field": lambda ow, om: of.exchange_field(om, om.m),
    "dmi_field": lambda ow, om: of.dmi_field(om, om.m),
    "anisotropy_field": lambda ow, om: of.anisotropy_field(om, om.m),
    "external_field": lambda ow, om: of.external_field(om, om.m, ow.bias),
    "effective_field": lambda ow, om: of.effective_field(om, om.m, ow.bias),
    "torque": lambda ow, om: of.llg_torque(om, om.m, of.effective_field(om, om.m, ow.bias)),
    "llg_torque": lambda ow, om: of.llg_torque(om, om.m, of.effective_field(om, om.m, ow.bias)),
    "damping_torque": lambda ow, om: of.damping_torque(om, om.m, of.effective_field(om, om.m, ow.bias)),
    "full_magnetization": lambda ow, om: 8e5 * om.m,
}


@pytest.mark.parametrize("name", sorted(FM_FIELD_TERMS))
def test_ferromagnet_field_factories_match_oracle(name):
    world, fm, ow, om = _fm_pair(1)
    q = getattr(_cpp, name)(fm)
    assert isinstance(q, _cpp.FerromagnetFieldQuantity)
    assert q.ncomp == 3 and q.grid.size == SIZE and q.system.grid.size == SIZE
    assert q.name.endswith(name) and q.unit == ("A/m" if name == "full_magnetization" else
                                                "/s" if "torque" in name else "T")
    got = q.eval()
    want = FM_FIELD_TERMS[name](ow, om)
    assert rel_max_err(got, want) < TOL
    assert np.allclose(q.average(), got.reshape(3, -1).astype(np.float64).mean(axis=1), rtol=1e-5,
                       atol=1e-6 * np.abs(got).max())
    q.exec()


def test_energy_factories_are_consistent():
    """Each energy density is -prefactor * Ms m . h of its field (energy.cu:18-50)
    and each energy is its density integrated over the magnet."""
    world, fm, ow, om = _fm_pair(2)
    V = CS[0] * CS[1] * CS[2]
    m = fm.magnetization.get().astype(np.float64)
    for term, pref in (("exchange", 0.5), ("dmi", 0.5), ("zeeman", 1.0)):
        field = "external_field" if term == "zeeman" else f"{term}_field"
        h = getattr(_cpp, field)(fm).eval().astype(np.float64)
        want = -pref * 8e5 * np.sum(m * h, axis=0)
        dens = getattr(_cpp, f"{term}_energy_density")(fm)
        assert isinstance(dens, _cpp.FerromagnetFieldQuantity) and dens.ncomp == 1 and dens.unit == "J/m3"
        got = dens.eval()[0]
        assert np.abs(got - want).max() <= 1e-5 * np.abs(want).max()
        e = getattr(_cpp, f"{term}_energy")(fm)
        assert isinstance(e, _cpp.FerromagnetScalarQuantity) and e.unit == "J"
        assert e.eval() == pytest.approx(got.astype(np.float64).sum() * V, rel=1e-5)
    tot = _cpp.total_energy(fm).eval()
    parts = sum(getattr(_cpp, f"{t}_energy")(fm).eval() for t in ("exchange", "dmi", "zeeman", "anisotropy",
                                                                   "demag"))
    assert tot == pytest.approx(parts, rel=1e-5)
    tq = _cpp.max_torque(fm).eval()
    assert tq == pytest.approx(np.linalg.norm(_cpp.torque(fm).eval().astype(np.float64), axis=0).max(), rel=1e-5)
    assert _cpp.max_angle(fm).eval() > 0


def test_remaining_ferromagnet_factories_evaluate():
    world, fm, ow, om = _fm_pair(3)
    for name in ("spin_transfer_torque", "demag_energy_density", "anisotropy_energy_density",
                 "inhomogeneous_exchange_field", "homogeneous_exchange_field",
                 "inhomogeneous_exchange_energy_density", "homogeneous_exchange_energy_density",
                 "total_energy_density", "conductivity_tensor", "electrical_potential", "thermal_noise",
                 "magnetoelastic_field", "magnetoelastic_energy_density", "magnetoelastic_force"):
        q = getattr(_cpp, name)(fm)
        v = q.eval()
        assert v.shape == (q.ncomp,) + SIZE[::-1] and np.isfinite(v).all(), name
    for name in ("demag_energy", "anisotropy_energy", "inhomogeneous_exchange_energy",
                 "homogeneous_exchange_energy", "magnetoelastic_energy"):
        assert math.isfinite(getattr(_cpp, name)(fm).eval()), name
    # not a sublattice: no sister, no homogeneous AFM exchange
    assert np.all(_cpp.homogeneous_exchange_field(fm).eval() == 0)
    rgb = _cpp.effective_field(fm).get_rgb()
    assert isinstance(rgb, _cpp.Field) and rgb.get().shape == (3,) + SIZE[::-1]


def test_magnet_factories_match_oracle():
    world, fm, ow, om = _fm_pair(4, elastic=True)
    strain = _cpp.strain_tensor(fm)
    assert isinstance(strain, _cpp.MagnetFieldQuantity) and strain.ncomp == 6
    assert rel_max_err(strain.eval(), oel.strain_tensor(om, om.u)) < TOL
    assert rel_max_err(_cpp.stress_tensor(fm).eval(), oel.stress_tensor(om, oel.strain_tensor(om, om.u))) < TOL
    acc = ow.rhs(ow.get_states())[2]
    assert rel_max_err(_cpp.elastic_acceleration(fm).eval(), acc) < TOL
    for name in ("elastic_force", "effective_body_force", "elastic_damping", "kinetic_energy_density",
                 "elastic_energy_density", "poynting_vector"):
        assert np.isfinite(getattr(_cpp, name)(fm).eval()).all(), name
    ke = _cpp.kinetic_energy(fm)
    assert isinstance(ke, _cpp.MagnetScalarQuantity)
    V = CS[0] * CS[1] * CS[2]
    assert ke.eval() == pytest.approx(0.5 * 8900.0 * np.sum(om.v ** 2) * V, rel=1e-5)
    assert _cpp.elastic_energy(fm).eval() > 0


def test_antiferromagnet_factories():
    rng = np.random.default_rng(5)
    nx, ny, nz = SIZE
    world = _cpp.World(CS, _cpp.Grid((0, 0, 0)), (0, 0, 0))
    afm = world.add_antiferromagnet(_cpp.Grid(SIZE), None, None, "afm")
    for sub, ms in zip(afm.sublattices(), (5e5, 3e5)):
        sub.msat.set(ms)
        sub.aex.set(1e-11)
        sub.magnetization.set(random_unit(rng, (nz, ny, nx)).astype(np.float32))
    afm.afmex_cell.set(-1e-12)
    afm.dmi_tensor.xyz.set(1e-3)
    assert afm.sub1().is_sublattice and afm.other_sublattice(afm.sub1()) == afm.sub2()
    m1, m2 = (s.magnetization.get().astype(np.float64) for s in afm.sublattices())
    n = _cpp.neel_vector(afm)
    assert isinstance(n, _cpp.AntiferromagnetFieldQuantity)
    assert np.allclose(n.eval(), (5e5 * m1 - 3e5 * m2) / 8e5, atol=1e-6)
    assert np.allclose(_cpp.full_magnetization(afm).eval(), 5e5 * m1 + 3e5 * m2, rtol=1e-6, atol=1e-1)
    assert np.allclose(_cpp.full_magnetization(afm.sub1()).eval(), 5e5 * m1, rtol=1e-6, atol=1e-1)
    ang = _cpp.angle_field(afm).eval()[0]
    # acos(sign(afmex_cell) m1 . m2): the angle from antiparallel for a
    # negative homogeneous exchange (afmexchange.cu:280-282)
    want = np.arccos(np.clip(-np.sum(m1 * m2, axis=0), -1, 1))
    assert np.abs(ang - want).max() < 1e-3
    assert _cpp.max_intracell_angle(afm).eval() == pytest.approx(want.max(), abs=1e-3)
    parts = _cpp.total_energy(afm.sub1()).eval() + _cpp.total_energy(afm.sub2()).eval()
    assert _cpp.total_energy(afm).eval() == pytest.approx(parts, rel=1e-5)
    # the homogeneous exchange of a sublattice (afmexchange.cu:28-48)
    want_h = 4 * -1e-12 * m2 / (0.35e-9 ** 2) / 3e5
    assert rel_max_err(_cpp.homogeneous_exchange_field(afm.sub1()).eval(), want_h) < TOL
    with pytest.raises(TypeError):
        _cpp.neel_vector(afm.sub1())


def test_timesolver_as_the_package_drives_it():
    """timesolver.py:61-152: set_method, adaptive_timestep, timestep, steps, run, time."""
    world, fm, ow, om = _fm_pair(6)
    ts = world.timesolver
    ts.set_method("DormandPrince")
    ts.adaptive_timestep = False
    ts.timestep = 5e-14
    assert not ts.adaptive_timestep and ts.timestep == pytest.approx(5e-14)
    ts.steps(20)
    ow.solver.set_method("DormandPrince")
    ow.solver.adaptive = False
    ow.solver.dt = 5e-14
    ow.solver.steps(20)
    assert ts.time == pytest.approx(20 * 5e-14, rel=1e-12)
    assert np.abs(fm.magnetization.get() - om.m).max() < 2e-5
    for k in ("headroom", "lower_bound", "upper_bound", "max_error", "sensible_factor", "sensible_timestep_default"):
        assert getattr(ts, k) > 0
    assert ts.sensible_timestep > 0
    with pytest.raises((TypeError, ValueError)):
        ts.set_method("NotAMethod")
    calls = []
    ts.run_while(lambda: len(calls) < 3 and not calls.append(1))
    assert ts.time == pytest.approx(23 * 5e-14, rel=1e-12)
