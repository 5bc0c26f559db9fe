"""Regions, inter-region exchange and Parameter.set_in_region on the GPU
(System regions, system.cpp:13-36; InterParameter, inter_parameter.cu;
exchange.cu:61-101, 150-199; afmexchange.cu:96-143; parameter.cu:41-47).

* exchange field, effective field and torque with random regions (indices
  with gaps), per-pair inter_exchange and scale_exchange tables, per-cell
  aex, and geometry / periodic / open-boundary / DMI cases, against the FP64
  oracle (pinned by the link-energy gradient): max |x - x_o| / max |x_o|
  <= 1e-5;
* an antiferromagnet with regions: the sublattice exchange and the
  neighbour antiferromagnetic exchange with inter_afmex_nn / scale_afmex_nn;
* fixed-step Dormand-Prince with regions against the oracle (1e-4 / 1e-5);
* the InterParameter, regions and set_in_region API, and its errors.
"""

import warnings

import numpy as np
import pytest

from conftest import rel_max_err
from helpers import make_pair, random_unit
from oracle import afm as oafm
from oracle import fields as of
from oracle import model as omodel

import mumaxplus as mp

pytestmark = pytest.mark.gpu

SIZE = (12, 9, 3)
CS = (2e-9, 2.5e-9, 2e-9)
IDX = (0, 2, 5)  # region indices with gaps


def _regions(rng, shape):
    return np.asarray(IDX)[rng.integers(0, len(IDX), size=shape)]


def _tables(mag, om):
    """Some pairs with an inter value, some with a scale; the rest default."""
    for key, i, j, v in (("inter_exchange", 0, 2, 2.4e-11), ("inter_exchange", 2, 5, 0.5e-11),
                         ("scale_exchange", 0, 5, 0.25), ("scale_exchange", 2, 5, 1.8)):
        getattr(mag, key).set_between(i, j, v)
        om.set_between(key, i, j, v)


@pytest.mark.parametrize("case", ["plain", "geometry", "periodic", "openbc", "dmi"])
def test_exchange_with_regions_matches_oracle(case):
    rng = np.random.default_rng(10 + len(case))
    nx, ny, nz = SIZE
    shape = (nz, ny, nx)
    params = {"msat": 8e5, "aex": 1e-11 * (0.5 + rng.random(shape)), "alpha": 0.05, "ku1": 1e5,
              "anisU": (0.0, 0.0, 1.0)}
    kw = {}
    if case == "geometry":
        kw["geometry"] = rng.random(shape) > 0.2
    if case == "periodic":
        kw.update(mastergrid=(nx, ny, 0), pbc=(1, 1, 0))
    if case == "openbc":
        kw["openbc"] = True
        kw["geometry"] = rng.random(shape) > 0.2
    if case == "dmi":
        kw["dmi"] = {"xxz": 1e-3, "yyz": 1e-3}
    reg = _regions(rng, shape)
    world, mag, ow, om = make_pair(SIZE, CS, params, random_unit(rng, shape), regions=reg, **kw)
    _tables(mag, om)
    want = of.exchange_field(om, om.m)
    assert rel_max_err(mag.exchange_field.eval(), want) < 1e-5
    h = ow.effective_field(om, om.m, None, om.u)
    assert rel_max_err(mag.effective_field.eval(), h) < 1e-5
    assert rel_max_err(mag.torque.eval(), of.llg_torque(om, om.m, h)) < 1e-5


def test_antiferromagnet_with_regions_matches_oracle():
    rng = np.random.default_rng(21)
    nx, ny, nz = SIZE
    shape = (nz, ny, nx)
    reg = _regions(rng, shape)
    world = mp.World(CS)
    afm = mp.Antiferromagnet(world, mp.Grid(SIZE), regions=reg)
    ow = omodel.World(CS)
    oa = ow.add_antiferromagnet(SIZE, defer_timestep=True)
    oa.regions = reg
    ann = -1e-12 * (0.5 + rng.random(shape))
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", UserWarning)
        afm.afmex_nn = ann.astype(np.float32)
        afm.afmex_cell = -2e-12
    oa.set("afmex_nn", ann.astype(np.float32).astype(np.float64))
    oa.set("afmex_cell", -2e-12)
    for sub, osub in ((afm.sub1, oa.sub1), (afm.sub2, oa.sub2)):
        sub.msat, sub.aex, sub.alpha = 6e5, 1.2e-11, 0.02
        osub.set("msat", 6e5)
        osub.set("aex", 1.2e-11)
        osub.set("alpha", 0.02)
        sub.enable_demag = False
        osub.enable_demag = False
        sub.magnetization = random_unit(rng, shape).astype(np.float32)
        osub.set_m(sub.magnetization.get().astype(np.float64))
    for key, i, j, v in (("inter_afmex_nn", 0, 5, -2e-12), ("scale_afmex_nn", 2, 5, 0.3),
                         ("scale_afmex_nn", 0, 2, 1.5)):
        getattr(afm, key).set_between(i, j, v)
        oa.set_between(key, i, j, v)
    afm.sub1.inter_exchange.set_between(0, 2, 2e-11)
    oa.sub1.set_between("inter_exchange", 0, 2, 2e-11)
    afm.sub2.scale_exchange.set_between(2, 5, 0.1)
    oa.sub2.set_between("scale_exchange", 2, 5, 0.1)
    s1, s2 = oa.sub1, oa.sub2
    for sub, a, b in ((afm.sub1, s1, s2), (afm.sub2, s2, s1)):
        assert rel_max_err(sub.inhomogeneous_exchange_field.eval(), oafm.afm_nn_field(a, a.m, b.m)) < 1e-5
        assert rel_max_err(sub.exchange_field.eval(), oafm.exchange_field_sublattice(a, a.m, b.m)) < 1e-5
    # inter_exchange set on the antiferromagnet reaches both sublattices
    afm.inter_exchange = 3e-11
    assert afm.sub1.inter_exchange.uniform_value == pytest.approx(3e-11)
    assert afm.sub2.inter_exchange.uniform_value == pytest.approx(3e-11)


def test_fixed_steps_with_regions_match_the_oracle():
    rng = np.random.default_rng(31)
    size = (16, 8, 2)
    shape = (2, 8, 16)
    reg = np.zeros(shape, dtype=int)
    reg[:, :, 8:] = 2
    reg[:, 4:, :] += 5
    params = {"msat": 8e5, "aex": 1.3e-11, "alpha": 0.02, "ku1": 3e5, "anisU": (0.0, 0.0, 1.0)}
    m0 = np.array([1.0, 0.2, 0.1])[:, None, None, None] + 0.05 * rng.normal(size=(3,) + shape)
    world, mag, ow, om = make_pair(size, CS, params, m0 / np.linalg.norm(m0, axis=0), regions=reg,
                                   world_bias=(0.0, 0.02, 0.0))
    for key, i, j, v in (("inter_exchange", 0, 2, 0.4e-11), ("scale_exchange", 0, 5, 0.2),
                         ("scale_exchange", 5, 7, 0.0)):
        getattr(mag, key).set_between(i, j, v)
        om.set_between(key, i, j, v)
    world.timesolver.set_method("DormandPrince")
    world.timesolver.adaptive_timestep = False
    world.timesolver.timestep = 1e-13
    ow.solver.set_method("DormandPrince", defer_timestep=True)
    ow.solver.adaptive = False
    ow.solver.dt = 1e-13
    world.timesolver.steps(150)
    ow.solver.steps(150)
    assert np.abs(mag.magnetization.get() - om.m).max() < 1e-4
    assert np.abs(np.array(mag.magnetization.average()) - om.m.reshape(3, -1).mean(axis=1)).max() < 1e-5


def test_regions_api_and_errors():
    world = mp.World(CS)
    reg = np.array([[[0, 0, 3, 3, 7]]] * 2).reshape(2, 1, 5)
    mag = mp.Ferromagnet(world, mp.Grid((5, 1, 2)), regions=reg)
    assert np.array_equal(mag.regions, reg)
    ie = mag.inter_exchange
    assert ie.is_uniform and ie.uniform_value == 0.0 and ie.number_of_regions == 3
    assert ie.region_indices == [0, 3, 7]
    ie.set_between(7, 3, 2e-11)
    assert not ie.is_uniform and ie.get_between(3, 7) == pytest.approx(2e-11)
    with pytest.raises(ValueError):
        ie.uniform_value
    M = ie.eval()
    assert M.shape == (8, 8) and M[3, 7] == M[7, 3] == pytest.approx(2e-11) and M[0, 3] == 0
    with pytest.raises(ValueError):
        ie.set_between(3, 3, 1.0)
    with pytest.raises(ValueError):
        ie.set_between(1, 3, 1.0)  # 1 is not a region of the magnet
    mag.scale_exchange = 0.5
    assert mag.scale_exchange.uniform_value == 0.5 and mag.scale_exchange.get_between(0, 7) == 0.5
    # set_in_region: scalar and vector parameters, static values
    mag.ku1 = 1.0
    mag.ku1.set_in_region(3, 4.0)
    assert np.array_equal(mag.ku1.eval()[0], np.where(reg == 3, 4.0, 1.0))
    mag.anisU = (0.0, 0.0, 1.0)
    mag.anisU.set_in_region(7, (1.0, 0.0, 0.0))
    u = mag.anisU.eval()
    assert np.array_equal(u[0], np.where(reg == 7, 1.0, 0.0)) and np.array_equal(u[2], np.where(reg == 7, 0.0, 1.0))
    with pytest.raises(ValueError):
        mag.ku1.set_in_region(4, 1.0)
    # a magnet without regions: every cell in region 0, no region defined
    plain = mp.Ferromagnet(world, mp.Grid((3, 1, 1), (20, 0, 0)))
    assert np.array_equal(plain.regions, np.zeros((1, 1, 3)))
    with pytest.raises(ValueError):
        plain.ku1.set_in_region(0, 1.0)
    with pytest.raises(ValueError):
        plain.inter_exchange.eval()
    # a batch takes the regions only when they change nothing
    w2 = mp.World(CS)
    m2 = mp.Ferromagnet(w2, mp.Grid((8, 8, 1)), regions=np.arange(64).reshape(1, 8, 8) % 2)
    m2.msat, m2.aex, m2.alpha = 8e5, 1.3e-11, 0.1
    w2.timesolver.adaptive_timestep = False
    w2.timesolver.timestep = 1e-13
    b = mp.Batch(1024)
    mid = b.add(m2, steps=1)
    assert mid is not None
    b.run_until_done()  # takes the finished member out
    m2.scale_exchange.set_between(0, 1, 0.5)
    with pytest.raises(ValueError):
        b.add(m2, steps=1)
