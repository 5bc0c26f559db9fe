"""Per-term effective field, torque and energy density on the GPU vs the FP64 oracle.

Tolerance (per term): max |h - h_oracle| / max |h_oracle| <= 1e-5 in FP32.
"""

import zlib

import numpy as np
import pytest

from conftest import rel_max_err
from helpers import make_pair, random_unit
from oracle import fields as of

pytestmark = pytest.mark.gpu

TOL = 1e-5
SIZE = (16, 8, 3)
CS = (2e-9, 2.5e-9, 3e-9)


def _setup(rng, openbc=False, geometry=True, nonuniform=True, dmi="interfacial", mg=(0, 0, 0), pbc=(0, 0, 0),
           size=SIZE, cs=CS):
    nx, ny, nz = size
    geom = (rng.random((nz, ny, nx)) > 0.2) if geometry else None
    shape = (nz, ny, nx)
    params = {"msat": (5e5 + 8e5 * rng.random(shape)) if nonuniform else 8e5,
              "aex": (1e-11 + 1e-11 * rng.random(shape)) if nonuniform else 1.3e-11,
              "alpha": 0.1 + 0.2 * rng.random(shape) if nonuniform else 0.05,
              "ku1": 5e5 if not nonuniform else 4e5 + 4e5 * rng.random(shape),
              "ku2": 1e5,
              "anisU": (0.3, 0.2, 1.0),
              "bias_magnetic_field": (0.01, -0.02, 0.03)}
    if nonuniform:
        params["msat"][rng.random(shape) < 0.1] = 0.0
    D = {}
    if dmi == "interfacial":
        D = {"xxz": 3e-3, "yyz": 3e-3}
    elif dmi == "bulk":
        D = {"xyz": 2e-3, "yxz": -2e-3, "zxy": 2e-3}
    elif dmi == "random":
        D = {k: (1e-3 * (1 + rng.random(shape))) for k in ("xxy", "xyz", "xxz", "yxy", "yyz", "yxz",
                                                            "zxy", "zyz", "zxz")}
    m0 = random_unit(rng, shape)
    return make_pair(size, cs, params, m0, geometry=geom, openbc=openbc, world_bias=(0.0, 0.0, 0.02),
                     dmi=D, mastergrid=mg, pbc=pbc)


TERMS = {
    "exchange_field": of.exchange_field,
    "dmi_field": of.dmi_field,
    "anisotropy_field": of.anisotropy_field,
}


@pytest.mark.parametrize("openbc", [False, True])
@pytest.mark.parametrize("dmi", ["interfacial", "bulk", "random"])
@pytest.mark.parametrize("term", sorted(TERMS))
def test_local_terms(term, dmi, openbc):
    rng = np.random.default_rng(zlib.crc32(f"{term}-{dmi}-{openbc}".encode()))
    world, mag, ow, om = _setup(rng, openbc=openbc, dmi=dmi)
    got = getattr(mag, term).eval()
    want = TERMS[term](om, om.m)
    assert rel_max_err(got, want) < TOL


@pytest.mark.parametrize("geometry", [False, True])
@pytest.mark.parametrize("openbc", [False, True])
@pytest.mark.parametrize("dmi", ["interfacial", "bulk"])
def test_uniform_parameter_path(dmi, openbc, geometry):
    """All parameters uniform: the coefficient-folded stencil path."""
    rng = np.random.default_rng(zlib.crc32(f"u-{dmi}-{openbc}-{geometry}".encode()))
    world, mag, ow, om = _setup(rng, openbc=openbc, dmi=dmi, nonuniform=False, geometry=geometry)
    heff = of.effective_field(om, om.m, ow.bias)
    assert rel_max_err(mag.effective_field.eval(), heff) < TOL
    assert rel_max_err(mag.torque.eval(), of.llg_torque(om, om.m, heff)) < TOL
    for term in ("exchange_field", "dmi_field", "anisotropy_field"):
        assert rel_max_err(getattr(mag, term).eval(), TERMS[term](om, om.m)) < TOL


def test_uniform_cubic_and_periodic():
    rng = np.random.default_rng(17)
    size = (12, 10, 3)
    world, mag, ow, om = _setup(rng, geometry=False, nonuniform=False, size=size, mg=(12, 0, 3), pbc=(1, 0, 2))
    mag.ku1 = 0.0
    mag.ku2 = 0.0
    om.set("ku1", 0.0)
    om.set("ku2", 0.0)
    for k, v in {"kc1": 3e5, "kc2": -1e5, "kc3": 2e4, "anisC1": (1.0, 0.1, 0.0), "anisC2": (-0.1, 1.0, 0.2)}.items():
        setattr(mag, k, v)
        om.set(k, v)
    heff = of.effective_field(om, om.m, ow.bias)
    assert rel_max_err(mag.effective_field.eval(), heff) < TOL
    assert rel_max_err(mag.torque.eval(), of.llg_torque(om, om.m, heff)) < TOL


def test_periodic_exchange_dmi():
    rng = np.random.default_rng(5)
    size = (12, 10, 2)
    world, mag, ow, om = _setup(rng, geometry=False, size=size, mg=(12, 10, 0), pbc=(1, 1, 0))
    for term in ("exchange_field", "dmi_field"):
        assert rel_max_err(getattr(mag, term).eval(), TERMS[term](om, om.m)) < TOL


def test_cubic_anisotropy():
    rng = np.random.default_rng(9)
    world, mag, ow, om = _setup(rng, dmi="none")
    mag.ku1 = 0.0
    mag.ku2 = 0.0
    om.set("ku1", 0.0)
    om.set("ku2", 0.0)
    for k, v in {"kc1": 3e5, "kc2": -1e5, "kc3": 2e4, "anisC1": (1.0, 0.1, 0.0), "anisC2": (-0.1, 1.0, 0.2)}.items():
        setattr(mag, k, v)
        om.set(k, v)
    assert rel_max_err(mag.anisotropy_field.eval(), of.anisotropy_field(om, om.m)) < TOL


def test_external_effective_torque_energy():
    rng = np.random.default_rng(21)
    world, mag, ow, om = _setup(rng, dmi="interfacial")
    assert rel_max_err(mag.external_field.eval(), of.external_field(om, om.m, ow.bias)) < TOL
    heff = of.effective_field(om, om.m, ow.bias)
    assert rel_max_err(mag.effective_field.eval(), heff) < TOL
    assert rel_max_err(mag.torque.eval(), of.llg_torque(om, om.m, heff)) < TOL
    assert rel_max_err(mag.damping_torque.eval(), of.damping_torque(om, om.m, heff)) < TOL
    e_want = of.total_energy_density(om, om.m, ow.bias)
    assert rel_max_err(mag.total_energy_density.eval()[0], e_want) < 5 * TOL
    E = mag.total_energy.eval()
    assert abs(E - of.integrate(om, e_want)) <= 5 * TOL * abs(of.integrate(om, np.abs(e_want)))


def test_reference_exchange_formulation():
    """test/test_exchange.py:6-46 of the reference (relative error < 1e-3)."""
    import mumaxplus as mp
    world = mp.World((1e3, 2e3, 3e3))
    magnet = mp.Ferromagnet(world, mp.Grid((16, 16, 4)))
    magnet.aex = 3.2e7
    magnet.msat = 5.4
    result = magnet.exchange_field.eval()
    m = magnet.magnetization.get().astype(np.float64)
    cs = magnet.cellsize
    h = np.zeros(m.shape)
    for axis, c in ((1, cs[2]), (2, cs[1]), (3, cs[0])):
        m_ = np.roll(m, 1, axis=axis)
        sl = [slice(None)] * 4
        sl[axis] = slice(1, None)
        h[tuple(sl)] += (m_ - m)[tuple(sl)] / c ** 2
        m_ = np.roll(m, -1, axis=axis)
        sl[axis] = slice(0, -1)
        h[tuple(sl)] += (m_ - m)[tuple(sl)] / c ** 2
    wanted = 2 * 3.2e7 * h / 5.4
    assert np.max(np.abs(result - wanted) / np.abs(wanted)) < 1e-3


def test_reference_interfacial_dmi_formulation():
    """test/test_interfacialdmi.py:8-56 of the reference (open BC, < 1e-4)."""
    import mumaxplus as mp
    world = mp.World((1.1, 2.0, 3.2))
    magnet = mp.Ferromagnet(world, mp.Grid((4, 5, 3)))
    magnet.enable_openbc = True
    magnet.msat = 3.1
    magnet.dmi_tensor.xxz = 7.1
    magnet.dmi_tensor.yyz = -7.1
    result = magnet.dmi_field.eval()
    m = magnet.magnetization.get().astype(np.float64)
    cs = magnet.cellsize
    Dxxz, Dyyz = 7.1, -7.1
    h = np.zeros(m.shape)
    m_ = np.roll(m, 1, axis=3)
    h[0][:, :, 1:] -= Dxxz * m_[2][:, :, 1:] / cs[0]
    h[2][:, :, 1:] += Dxxz * m_[0][:, :, 1:] / cs[0]
    m_ = np.roll(m, -1, axis=3)
    h[0][:, :, 0:-1] += Dxxz * m_[2][:, :, 0:-1] / cs[0]
    h[2][:, :, 0:-1] -= Dxxz * m_[0][:, :, 0:-1] / cs[0]
    m_ = np.roll(m, 1, axis=2)
    h[1][:, 1:, :] -= Dyyz * m_[2][:, 1:, :] / cs[1]
    h[2][:, 1:, :] += Dyyz * m_[1][:, 1:, :] / cs[1]
    m_ = np.roll(m, -1, axis=2)
    h[1][:, 0:-1, :] += Dyyz * m_[2][:, 0:-1, :] / cs[1]
    h[2][:, 0:-1, :] -= Dyyz * m_[1][:, 0:-1, :] / cs[1]
    wanted = h / 3.1
    assert np.max(np.abs(result - wanted) / np.abs(wanted)) < 1e-4
