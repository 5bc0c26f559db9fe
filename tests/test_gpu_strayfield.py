"""Stray fields between magnets on the GPU.

* The reference's own checks (test/test_strayfield.py): FFT against the
  direct sum for every grid pair (relative error < 1e-4), the enclosing-box
  identity, and the source / destination switches over all 16 source
  combinations of four magnets;
* parity with the FP64 oracle (oracle/stray.py, oracle/model.py): the field of
  one magnet on another grid, every magnet's external and effective field in a
  world of two ferromagnets and an antiferromagnet, and 30 fixed
  Dormand-Prince steps of that world; max |x - x_oracle| / max |x_oracle|
  <= 1e-5 for single evaluations, stated per test for time stepping.
"""

import warnings

import numpy as np
import pytest

from conftest import rel_max_err
from helpers import random_unit
from oracle import fields as of
from oracle import model as omodel
from oracle import stray as ost

import mumaxplus as mp

pytestmark = pytest.mark.gpu

CS = (1e-9, 2e-9, 3.1e-9)

CASES = [
    ((17, 4, 1), (3, 6, 0), (17, 4, 1), (3, 6, 0)),
    ((8, 10, 1), (2, 15, 0), (6, 7, 2), (-5, 0, 2)),
    ((8, 10, 1), (2, 15, 0), (32, 32, 1), (-5, 0, 0)),
    ((8, 10, 1), (2, 15, -1), (32, 32, 1), (-5, 0, 0)),
    ((8, 10, 2), (2, 15, -1), (32, 32, 1), (-5, 0, 1)),
    ((8, 10, 2), (2, 15, 2), (32, 32, 5), (0, 0, 1)),
]


def _max_rel(result, wanted):
    err = np.linalg.norm(result - wanted, axis=0)
    return float(np.max(err / np.linalg.norm(wanted, axis=0)))


@pytest.mark.parametrize("pbc", [None, (1, 2, 0)])
@pytest.mark.parametrize("case", CASES)
def test_fft_against_brute(case, pbc):
    sn, so, hn, ho = case
    world = mp.World(CS)
    magnet = mp.Ferromagnet(world, mp.Grid(sn, so))
    magnet.msat = 800e3
    magnet.magnetization = (1, 0, 0)
    if pbc is not None:
        world.set_pbc(pbc)
    sf = mp.StrayField(magnet, mp.Grid(hn, ho))
    sf.set_method("fft")
    result = sf.eval()
    sf.set_method("brute")
    wanted = sf.eval()
    assert _max_rel(result, wanted) < 1e-4


@pytest.mark.parametrize("pbc", [None, (1, 2, 1)])
@pytest.mark.parametrize("case", CASES)
def test_field_is_the_box_demag_field(case, pbc):
    """The field on hgrid equals the demag field of the box holding both
    grids with the magnetisation only inside mgrid (test_strayfield.py:86-136)."""
    sn, so, hn, ho = case
    bo = [min(so[a], ho[a]) for a in range(3)]
    bn = [max(so[a] + sn[a], ho[a] + hn[a]) - bo[a] for a in range(3)]
    box = mp.Grid(tuple(bn), tuple(bo))
    reps = (0, 0, 0) if pbc is None else pbc
    mgrid = mp.Grid((0, 0, 0)) if pbc is None else box
    world = mp.World(CS, reps, mgrid)
    magnet = mp.Ferromagnet(world, mp.Grid(sn, so))
    magnet.msat = 800e3
    magnet.magnetization = (1, -1, 2)
    hfield = mp.StrayField(magnet, mp.Grid(hn, ho)).eval()
    world2 = mp.World(CS, reps, mgrid)
    magnet2 = mp.Ferromagnet(world2, box)
    s0 = [so[a] - bo[a] for a in range(3)]
    h0 = [ho[a] - bo[a] for a in range(3)]
    ms = np.zeros((1, bn[2], bn[1], bn[0]))
    ms[:, s0[2]:s0[2] + sn[2], s0[1]:s0[1] + sn[1], s0[0]:s0[0] + sn[0]] = 800e3
    magnet2.msat = ms
    m = np.ones((3, bn[2], bn[1], bn[0]))
    m[:, s0[2]:s0[2] + sn[2], s0[1]:s0[1] + sn[1], s0[0]:s0[0] + sn[0]] = magnet.magnetization.eval()
    magnet2.magnetization = m
    wanted = magnet2.demag_field.eval()[:, h0[2]:h0[2] + hn[2], h0[1]:h0[1] + hn[1], h0[0]:h0[0] + hn[0]]
    assert _max_rel(hfield, wanted) < 1e-4


@pytest.mark.parametrize("pbc", [None, (1, 1, 0)])
def test_field_on_a_grid_matches_oracle(pbc):
    rng = np.random.default_rng(4)
    sn, so = (9, 6, 2), (3, -2, 1)
    hn, ho = (12, 5, 3), (14, 0, 0)
    geom = rng.random((sn[2], sn[1], sn[0])) > 0.25
    world = mp.World(CS)
    magnet = mp.Ferromagnet(world, mp.Grid(sn, so), geometry=geom)
    ms = (5e5 + 5e5 * rng.random((sn[2], sn[1], sn[0]))).astype(np.float32)
    magnet.msat = ms
    magnet.magnetization = random_unit(rng, (sn[2], sn[1], sn[0])).astype(np.float32)
    mg, reps = (0, 0, 0), (0, 0, 0)
    if pbc is not None:
        world.set_pbc(pbc)
        mg = world.mastergrid.size
        reps = pbc
    got = mp.StrayField(magnet, mp.Grid(hn, ho)).eval()
    M = np.where(geom[None], ms.astype(np.float64) * magnet.magnetization.get().astype(np.float64), 0.0)
    want = ost.stray_field(M, so, hn, ho, world.cellsize, mg, reps)
    assert rel_max_err(got, want) < 1e-5


def _coupled_world(rng):
    """Two ferromagnets and an antiferromagnet near each other, in the product
    and in the oracle."""
    cs = (2e-9, 2e-9, 2e-9)
    world = mp.World(cs)
    ow = omodel.World(cs)
    mags, omags = [], []
    specs = [((10, 6, 2), (0, 0, 0)), ((6, 8, 1), (12, 2, 1))]
    for n, o in specs:
        mag = mp.Ferromagnet(world, mp.Grid(n, o))
        om = ow.add_ferromagnet(n, o, defer_timestep=True)
        for k, v in {"msat": 8e5, "aex": 1.3e-11, "alpha": 0.1, "ku1": 2e5, "anisU": (0.0, 1.0, 0.3)}.items():
            setattr(mag, k, v)
            om.set("bias" if k == "bias_magnetic_field" else k, v)
        mag.magnetization = random_unit(rng, (n[2], n[1], n[0])).astype(np.float32)
        om.set_m(mag.magnetization.get().astype(np.float64))
        mags.append(mag)
        omags.append(om)
    afm = mp.Antiferromagnet(world, mp.Grid((5, 5, 2), (3, 8, 0)))
    oa = ow.add_antiferromagnet((5, 5, 2), (3, 8, 0), defer_timestep=True)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", UserWarning)
        afm.afmex_cell = -4e-12
    oa.set("afmex_cell", -4e-12)
    for i, (sub, osub) in enumerate(zip(afm.sublattices, oa.sublattices)):
        for k, v in {"msat": 5.66e5 if i == 0 else 4e5, "aex": 2.48e-12, "alpha": 0.1}.items():
            setattr(sub, k, v)
            osub.set(k, v)
        sub.magnetization = random_unit(rng, (2, 5, 5)).astype(np.float32)
        osub.set_m(sub.magnetization.get().astype(np.float64))
    world.bias_magnetic_field = (0.0, 0.01, 0.0)
    ow.bias = (0.0, 0.01, 0.0)
    ow.solver.equations_changed()
    return world, mags, afm, ow, omags, oa


def test_world_fields_include_the_other_magnets():
    rng = np.random.default_rng(12)
    world, mags, afm, ow, omags, oa = _coupled_world(rng)
    for mag, om in zip(mags, omags):
        hs = ow.stray_field_on(om)
        want_ext = of.external_field(om, om.m, ow.bias) + np.where(om.geom_mask()[None], hs, 0.0)
        assert rel_max_err(mag.external_field.eval(), want_ext) < 1e-5
        heff = ow.effective_field(om, om.m, None, om.u)
        assert rel_max_err(mag.effective_field.eval(), heff) < 1e-5
        assert rel_max_err(mag.torque.eval(), of.llg_torque(om, om.m, heff)) < 1e-5
    for i, (sub, osub) in enumerate(zip(afm.sublattices, oa.sublattices)):
        m2 = oa.sublattices[1 - i].m
        heff = ow.effective_field(osub, osub.m, m2, oa.u)
        assert rel_max_err(sub.effective_field.eval(), heff) < 1e-5
    # destination switch of the antiferromagnet and source switch of a ferromagnet
    afm.enable_as_stray_field_destination = False
    oa.stray_destination = False
    mags[0].enable_as_stray_field_source = False
    omags[0].stray_source = False
    for mag, om in zip(mags, omags):
        heff = ow.effective_field(om, om.m, None, om.u)
        assert rel_max_err(mag.effective_field.eval(), heff) < 1e-5
    for i, (sub, osub) in enumerate(zip(afm.sublattices, oa.sublattices)):
        heff = ow.effective_field(osub, osub.m, oa.sublattices[1 - i].m, oa.u)
        assert rel_max_err(sub.effective_field.eval(), heff) < 1e-5
    # zeeman energy includes the stray field
    om = omags[1]
    e = -1.0 * om.val("msat") * np.sum(om.m * (of.external_field(om, om.m, ow.bias)
                                               + ow.stray_field_on(om)), axis=0)
    assert rel_max_err(mags[1].zeeman_energy_density.eval()[0], np.where(om.geom_mask(), e, 0.0)) < 1e-5


def test_coupled_world_fixed_steps_match_oracle():
    rng = np.random.default_rng(13)
    world, mags, afm, ow, omags, oa = _coupled_world(rng)
    ts = world.timesolver
    ts.adaptive_timestep = False
    ts.set_method("DormandPrince")
    ts.timestep = 5e-15
    ow.solver.adaptive = False
    ow.solver.set_method("DormandPrince")
    ow.solver.dt = 5e-15
    ts.steps(30)
    ow.solver.steps(30)
    for mag, om in zip(mags, omags):
        assert np.abs(mag.magnetization.get() - om.m).max() < 2e-4
    for sub, osub in zip(afm.sublattices, oa.sublattices):
        assert np.abs(sub.magnetization.get() - osub.m).max() < 2e-4


@pytest.mark.parametrize("pbc", [None, (1, 1, 0), (0, 0, 1)])
def test_source_and_destination_switches(pbc):
    """test_strayfield.py:141-206: every magnet's external field is the sum of
    the pairwise stray fields of the enabled sources, for all 16 source
    combinations of four magnets; none when it is not a destination."""
    rng = np.random.default_rng(5)
    world = mp.World(CS)
    magnets = []
    for i in range(4):
        size = tuple(int(v) for v in rng.integers(2, 16, 3))
        mag = mp.Ferromagnet(world, mp.Grid(size, tuple(16 * i for _ in range(3))))
        mag.aex = 13e-12
        mag.msat = 800e3
        magnets.append(mag)
    if pbc is not None:
        world.set_pbc(pbc)
    for mag in magnets:
        mag.enable_as_stray_field_destination = False
        assert np.all(np.abs(mag.external_field.eval()) < 1e-15)
        mag.enable_as_stray_field_destination = True
    pair = {}
    for src in magnets:
        for dst in magnets:
            if src is not dst:
                pair[(src.name, dst.name)] = mp.StrayField(src, dst.grid).eval()
    for i in range(16):
        enable = [bool((i >> k) & 1) for k in range(4)]
        for on, mag in zip(enable, magnets):
            mag.enable_as_stray_field_source = on
        for dst in magnets:
            got = dst.external_field.eval()
            want = np.zeros_like(got)
            size = np.zeros_like(got)
            for on, src in zip(enable, magnets):
                if on and src is not dst:
                    want += pair[(src.name, dst.name)]
                    size += np.abs(pair[(src.name, dst.name)])
            # fields of different sources may cancel: rounding scales with
            # the size of the terms, not of their sum
            scale = max(float(size.max()), 1e-30)
            assert float(np.abs(got - want).max()) <= 1e-5 * scale + 1e-12
