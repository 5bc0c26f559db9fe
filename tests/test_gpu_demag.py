"""Demag convolution on the GPU vs the FP64 oracle (Newell kernel + padded FFT).

Tolerance: max |h - h_oracle| / max |h_oracle| <= 2e-6 (FP32 FFT passes and an
FP32-rounded kernel spectrum against float64 throughout).
"""

import numpy as np
import pytest

from conftest import rel_max_err
from helpers import make_pair, random_unit
from oracle import demag as odemag
from oracle import newell as onewell

import mumaxplus as mp

pytestmark = pytest.mark.gpu

TOL = 2e-6

CASES = [
    # size, cellsize, geometry?, mastergrid, pbc
    ((16, 8, 4), (1e-9, 1.5e-9, 2e-9), False, (0, 0, 0), (0, 0, 0)),
    ((13, 7, 5), (2e-9, 2e-9, 2e-9), True, (0, 0, 0), (0, 0, 0)),
    ((32, 16, 1), (3.90625e-9, 3.90625e-9, 3e-9), False, (0, 0, 0), (0, 0, 0)),
    ((33, 9, 1), (2e-9, 3e-9, 1e-9), True, (0, 0, 0), (0, 0, 0)),
    ((1, 12, 6), (1e-9, 1e-9, 1e-9), False, (0, 0, 0), (0, 0, 0)),
    ((10, 1, 1), (1e-9, 1e-9, 1e-9), False, (0, 0, 0), (0, 0, 0)),
    ((1, 1, 1), (1e-9, 2e-9, 3e-9), False, (0, 0, 0), (0, 0, 0)),
    ((16, 8, 1), (5e-9, 5e-9, 5e-9), False, (16, 8, 0), (2, 2, 0)),
    ((8, 6, 4), (2e-9, 2e-9, 2e-9), True, (8, 0, 4), (1, 0, 2)),
]


@pytest.mark.parametrize("size,cs,use_geom,mg,pbc", CASES)
def test_demag_field_matches_oracle(size, cs, use_geom, mg, pbc):
    rng = np.random.default_rng(7)
    nx, ny, nz = size
    geom = rng.random((nz, ny, nx)) > 0.25 if use_geom else None
    msat = 4e5 + 8e5 * rng.random((nz, ny, nx))
    m0 = random_unit(rng, (nz, ny, nx))
    world, mag, ow, om = make_pair(size, cs, {"msat": msat}, m0, geometry=geom, mastergrid=mg, pbc=pbc)
    h = mag.demag_field.eval()
    wanted = odemag.DemagPlan(size, cs, mg, pbc).field(om.m, om.val("msat"), om.geom_mask())
    assert rel_max_err(h, wanted) < TOL


def test_demag_uniform_msat_3d_medium():
    rng = np.random.default_rng(3)
    size = (40, 24, 12)
    cs = (2e-9, 2e-9, 2e-9)
    m0 = random_unit(rng, (12, 24, 40))
    world, mag, ow, om = make_pair(size, cs, {"msat": 1.1e6}, m0)
    h = mag.demag_field.eval()
    wanted = odemag.DemagPlan(size, cs).field(om.m, 1.1e6)
    assert rel_max_err(h, wanted) < TOL


def test_demag_kernel_export_matches_reference_layout():
    world = mp.World((1e-9, 2e-9, 3e-9))
    mag = mp.Ferromagnet(world, mp.Grid((6, 4, 3)))
    k = mp._cpp._demag_kernel(mag._impl)
    want = onewell.reference_kernel((6, 4, 3), (1e-9, 2e-9, 3e-9))
    assert k.shape == want.shape
    assert np.abs(k - want).max() < 1e-6 * np.abs(want).max()


def test_reference_test_demag_formulation():
    """test/test_demag.py:6-45 of the reference, re-expressed: numpy FFT of the
    exported kernel vs the product's demag field (tolerance 2e-3 there)."""
    world = mp.World((1e-9, 1e-9, 1e-9))
    magnet = mp.Ferromagnet(world, mp.Grid((16, 4, 3)))
    kernel = mp._cpp._demag_kernel(magnet._impl).astype(np.float64)
    mag = magnet.msat.average()[0] * magnet.magnetization.get().astype(np.float64)
    pad = ((0, kernel.shape[1] - mag.shape[1]), (0, kernel.shape[2] - mag.shape[2]),
           (0, kernel.shape[3] - mag.shape[3]))
    m = np.fft.fftn(np.pad(mag, ((0, 0), *pad), "constant"), axes=(1, 2, 3))
    kxx, kyy, kzz, kxy, kxz, kyz = np.fft.fftn(kernel, axes=(1, 2, 3))
    hx = np.fft.ifftn(m[0] * kxx + m[1] * kxy + m[2] * kxz)
    hy = np.fft.ifftn(m[0] * kxy + m[1] * kyy + m[2] * kyz)
    hz = np.fft.ifftn(m[0] * kxz + m[1] * kyz + m[2] * kzz)
    h = -4 * np.pi * 1e-7 * np.array([hx, hy, hz]).real
    wanted = h[:, h.shape[1] - mag.shape[1]:, h.shape[2] - mag.shape[2]:, h.shape[3] - mag.shape[3]:]
    result = magnet.demag_field.eval()
    assert np.max(np.abs((wanted - result) / result)) < 2e-3


def test_uniform_prism_average_field_is_aharoni():
    """Volume average of -H/(mu0 Ms) of a uniformly magnetised prism equals
    Aharoni's closed-form demag factor (J. Appl. Phys. 83, 3432 (1998))."""
    from test_oracle_demag import aharoni_nzz
    nx, ny, nz = 12, 8, 5
    cs = (2e-9, 3e-9, 1.5e-9)
    world = mp.World(cs)
    mag = mp.Ferromagnet(world, mp.Grid((nx, ny, nz)))
    mag.msat = 1.0
    mag.magnetization = (0, 0, 1)
    hz = mag.demag_field.average()[2]
    want = aharoni_nzz(nx * cs[0], ny * cs[1], nz * cs[2])
    assert abs(-hz / odemag.MU0 - want) < 1e-6


def test_demag_operator_on_device_pointers():
    """mpa_demag_plan_exec on raw device buffers equals the World path."""
    from helpers import DeviceBuffer
    rng = np.random.default_rng(11)
    size, cs = (24, 10, 3), (2e-9, 2e-9, 2e-9)
    m0 = random_unit(rng, (3, 10, 24)).astype(np.float32)
    world, mag, ow, om = make_pair(size, cs, {"msat": 8e5}, m0)
    want = mag.demag_field.eval()
    plan = mp._cpp.DemagPlan(size, cs)
    md = DeviceBuffer.from_host(mag.magnetization.get())
    hd = DeviceBuffer(md.nbytes)
    plan.exec(md.ptr, hd.ptr, msat_uniform=8e5)
    got = hd.to_host(np.float32, want.shape)
    assert np.abs(got - want).max() <= 1e-6 * np.abs(want).max()
    ks, algo_bytes, dev_bytes = plan.info()
    assert ks == (64, 32, 8) and algo_bytes > 0 and dev_bytes > 0
