"""Spin-transfer torques and the thermal field on the GPU.

* Zhang-Li and Slonczewski torques, the torque with them and the LLG torque
  without them, against the FP64 oracle (oracle/fields.py): max |x - x_o| /
  max |x_o| <= 1e-5, in set-ups with geometry, per-cell parameters, periodic
  wrap and both fixed-layer positions;
* test/test_free_layer_thickness.py (empty layers do not change the
  Slonczewski torque once the free layer thickness is set);
* a domain wall under a Zhang-Li current with xi = alpha moves by -u t,
  u = muB P j / (e Ms (1 + xi^2)) (the torque then reduces to (u . grad) m):
  within 5 % on a 1 nm lattice;
* the thermal field: the same Philox stream as the oracle (oracle/thermal.py)
  cell by cell (1e-5), a new draw per evaluation, the variance closed form,
  and noisy fixed-step Dormand-Prince trajectories against the oracle, which
  draws once per step and evaluates every stage afresh.
"""

import numpy as np
import pytest

from conftest import rel_max_err
from helpers import make_pair, random_unit
from oracle import fields as of
from oracle import thermal as oth

import mumaxplus as mp

pytestmark = pytest.mark.gpu


def _zl_pair(rng, case):
    size = (12, 7, 3)
    cs = (2e-9, 3e-9, 1.5e-9)
    geom = None
    pbc = (0, 0, 0)
    mg = (0, 0, 0)
    params = {"msat": 8e5, "aex": 1.3e-11, "alpha": 0.03, "pol": 0.6, "xi": 0.07, "jcur": (2e12, -1e12, 5e11)}
    if case == "geometry":
        geom = rng.random((3, 7, 12)) > 0.2
    if case == "percell":
        ms = 6e5 + 4e5 * rng.random((3, 7, 12))
        ms[0, 0, :3] = 0.0
        params["msat"] = ms
        params["pol"] = 0.3 + 0.5 * rng.random((3, 7, 12))
    if case == "periodic":
        mg, pbc = (12, 7, 0), (1, 1, 0)
    m0 = random_unit(rng, (3, 7, 12))
    return make_pair(size, cs, params, m0, geometry=geom, mastergrid=mg, pbc=pbc)


@pytest.mark.parametrize("case", ["plain", "geometry", "percell", "periodic"])
def test_zhang_li_matches_oracle(case):
    rng = np.random.default_rng(3)
    world, mag, ow, om = _zl_pair(rng, case)
    want = of.zhang_li_torque(om, om.m)
    assert rel_max_err(mag.spin_transfer_torque.eval(), want) < 1e-5
    h = ow.effective_field(om, om.m, None, om.u)
    llg = of.llg_torque(om, om.m, h)
    assert rel_max_err(mag.llg_torque.eval(), llg) < 1e-5
    assert rel_max_err(mag.torque.eval(), llg + want) < 1e-5
    mag.enable_zhang_li_torque = False
    assert np.abs(mag.spin_transfer_torque.eval()).max() == 0


@pytest.mark.parametrize("on_top", [True, False])
@pytest.mark.parametrize("percell", [False, True])
def test_slonczewski_matches_oracle(on_top, percell):
    rng = np.random.default_rng(4)
    size, cs = (10, 6, 2), (3e-9, 3e-9, 2e-9)
    params = {"msat": 8e5, "aex": 1.3e-11, "alpha": 0.01, "pol": 0.5669, "Lambda": 2.0, "epsilon_prime": 1.0,
              "jcur": (0.0, 0.0, -6e-3 / (160e-9 * 80e-9)), "fixed_layer": (np.cos(0.35), np.sin(0.35), 0.0)}
    if percell:
        params["free_layer_thickness"] = 1e-9 + 4e-9 * rng.random((2, 6, 10))
        params["epsilon_prime"] = rng.random((2, 6, 10))
    world, mag, ow, om = make_pair(size, cs, params, random_unit(rng, (2, 6, 10)))
    mag.fixed_layer_on_top = on_top
    om.fixed_on_top = on_top
    want = of.slonczewski_torque(om, om.m)
    assert rel_max_err(mag.spin_transfer_torque.eval(), want) < 1e-5
    h = ow.effective_field(om, om.m, None, om.u)
    assert rel_max_err(mag.torque.eval(), of.llg_torque(om, om.m, h) + want) < 1e-5


def test_free_layer_thickness_and_empty_layers():
    """test/test_free_layer_thickness.py."""
    m0 = np.array([0.2, 0.7, 0.4])
    m0 /= np.linalg.norm(m0)

    def build(size, geometry=None):
        world = mp.World(cellsize=(1e-9, 1e-9, 1e-9))
        magnet = mp.Ferromagnet(world, mp.Grid(size), geometry=geometry)
        magnet.msat = 4.3
        magnet.jcur = (0, 0, 5.4)
        magnet.pol = 0.6
        magnet.fixed_layer = (1, 0, 1)
        magnet.free_layer_thickness = 1e-9
        magnet.magnetization = tuple(m0)
        return world, magnet

    w1, a = build((1, 1, 2))
    w2, b = build((1, 1, 4), geometry=np.array([[[True]], [[True]], [[False]], [[False]]]))
    ta = a.spin_transfer_torque.eval()[:, 0, 0, 0]
    tb = b.spin_transfer_torque.eval()[:, 0, 0, 0]
    assert np.linalg.norm(tb - ta) / np.linalg.norm(ta) < 1e-5
    # the default free layer thickness is the grid's thickness
    w3 = mp.World(cellsize=(1e-9, 1e-9, 1.5e-9))
    c = mp.Ferromagnet(w3, mp.Grid((2, 2, 3)))
    assert c.free_layer_thickness.uniform_value == pytest.approx(4.5e-9, rel=1e-6)


def test_zhang_li_domain_wall_moves_with_the_drift_velocity():
    nx, cs = 512, 1e-9
    ms, aex, ku, alpha, pol, j = 8e5, 1.3e-11, 5e5, 0.02, 0.5, 1e12
    world = mp.World((cs, cs, cs))
    mag = mp.Ferromagnet(world, mp.Grid((nx, 1, 1)))
    mag.enable_demag = False
    mag.msat, mag.aex, mag.ku1, mag.anisU, mag.alpha = ms, aex, ku, (1, 0, 0), alpha
    x = (np.arange(nx) - nx / 2 + 0.5) * cs
    delta = np.sqrt(aex / ku)
    th = 2 * np.arctan(np.exp(x / delta))
    m0 = np.stack([np.cos(th), np.sin(th), 0 * th])[:, None, None, :]
    mag.magnetization = m0.astype(np.float32)

    def centre():
        mx = mag.magnetization.get()[0, 0, 0]
        i = int(np.argmax(np.diff(np.sign(mx)) != 0))
        return (i + mx[i] / (mx[i] - mx[i + 1]) - nx / 2 + 0.5) * cs

    world.timesolver.set_method("DormandPrince")
    mag.relax()
    x0 = centre()
    mag.xi = alpha
    mag.pol = pol
    mag.jcur = (j, 0, 0)
    t = 0.5e-9
    world.timesolver.run(t)
    u = 9.2740091523e-24 * pol * j / (1.60217646e-19 * ms * (1 + alpha * alpha))
    shift = centre() - x0
    assert shift < 0
    assert abs(shift / (-u * t) - 1) < 0.05


def _noisy_pair(seed):
    rng = np.random.default_rng(seed)
    size, cs = (16, 8, 2), (2e-9, 2e-9, 2e-9)
    params = {"msat": 8e5, "aex": 1.3e-11, "alpha": 0.1, "ku1": 2e5, "anisU": (0.0, 0.0, 1.0),
              "temperature": 300.0}
    world, mag, ow, om = make_pair(size, cs, params, random_unit(rng, (2, 8, 16)))
    mag.noise_seed = 0xC0FFEE + seed
    om.noise_seed = 0xC0FFEE + seed
    return world, mag, ow, om


def test_thermal_field_matches_oracle_and_redraws():
    world, mag, ow, om = _noisy_pair(1)
    mag.alpha = np.linspace(0.05, 0.2, 16 * 8 * 2).reshape(2, 8, 16).astype(np.float32)
    om.set("alpha", mag.alpha.eval()[0].astype(np.float64))
    a = mag.thermal_noise.eval()
    b = mag.thermal_noise.eval()
    assert rel_max_err(a, oth.thermal_field(om, 1, 0)) < 1e-5
    assert rel_max_err(b, oth.thermal_field(om, 1, 1)) < 1e-5
    assert not np.allclose(a, b)


def test_thermal_field_variance():
    world = mp.World((2e-9, 3e-9, 1e-9))
    mag = mp.Ferromagnet(world, mp.Grid((128, 64, 8)))
    mag.msat, mag.alpha, mag.temperature = 8e5, 0.1, 300.0
    h = mag.thermal_noise.eval()
    want = 2 * 1.38064852e-23 * 1.7595e11 * 0.1 * 300.0 / ((1 + 0.01) * 8e5 * 6e-27)
    n = h[0].size
    for c in range(3):
        assert abs(h[c].astype(np.float64).var() / want - 1) < 10 / np.sqrt(n)
        assert abs(h[c].astype(np.float64).mean()) < 5 * np.sqrt(want / n)
    with pytest.raises(ValueError):
        odd = mp.Ferromagnet(world, mp.Grid((3, 3, 1), (500, 0, 0)))
        odd.temperature = 10.0


@pytest.mark.parametrize("method", ["DormandPrince", "Heun"])
def test_noisy_fixed_steps_match_oracle(method):
    world, mag, ow, om = _noisy_pair(2)
    for ts in (world.timesolver,):
        ts.set_method(method)
        ts.adaptive_timestep = False
        ts.timestep = 1e-13
    ow.solver.set_method(method, defer_timestep=True)
    ow.solver.adaptive = False
    ow.solver.dt = 1e-13
    world.timesolver.steps(20)
    ow.solver.steps(20)
    assert mag.noise_draws == 20 and om.noise_draws[0] == 20
    assert np.abs(mag.magnetization.get() - om.m).max() < 1e-4
    assert np.abs(np.array(mag.magnetization.average()) - om.m.reshape(3, -1).mean(axis=1)).max() < 1e-5


def test_batch_refuses_noise_and_currents():
    world, mag, ow, om = _noisy_pair(3)
    world.timesolver.adaptive_timestep = False
    world.timesolver.timestep = 1e-13
    b = mp.Batch(4096)
    with pytest.raises(ValueError):
        b.add(mag, steps=1)
