"""The uniform-parameter field kernel's LDS halo-tile variant, on the GPU.

A magnet with no geometry mask and no periodic wrap, at least 64 x 4 cells
per plane, takes its exchange and DMI neighbours from an LDS tile that a
workgroup marches through a chunk of z planes (world option "field_zchunk",
the planes per workgroup; 0 selects the one-cell-per-thread kernel).  Both
kernels evaluate the same cell function; the compiler may fuse its
multiply-adds differently in the two, so the values agree to a few FP32
ulps (any wrong neighbour or halo cell would differ at order one).  Cases:
chunks that divide the grid and chunks that do not, a single layer, ragged
tiles (x and y not multiples of the tile), open boundaries, cubic
anisotropy and a stray field from a second magnet, plus Runge-Kutta runs.
Parity with the FP64 oracle is covered by the field and time-stepping
tests, which run the tile kernel wherever it applies.
"""

import numpy as np
import pytest

from helpers import random_unit

import mumaxplus as mp

pytestmark = pytest.mark.gpu


def _magnet(zchunk, size, seed=0, cubic=False, openbc=False, second=False):
    world = mp.World((2e-9, 2.5e-9, 3e-9))
    nx, ny, nz = size
    mag = mp.Ferromagnet(world, mp.Grid(size))
    mag.msat, mag.aex, mag.alpha = 8e5, 1.3e-11, 0.05
    if cubic:
        mag.kc1, mag.kc2, mag.anisC1, mag.anisC2 = 2e4, -1e4, (1, 1, 0), (-1, 1, 0.2)
    else:
        mag.ku1, mag.anisU = 3e5, (0.1, 0.2, 1.0)
    mag.dmi_tensor.set_interfacial_dmi(2e-3)
    mag.dmi_tensor.xxy = 1e-3  # an in-plane entry, so the z boundary terms carry DMI too
    mag.enable_openbc = openbc
    world.bias_magnetic_field = (0.01, 0.0, 0.02)
    mag.magnetization = random_unit(np.random.default_rng(seed), (nz, ny, nx)).astype(np.float32)
    if second:
        other = mp.Ferromagnet(world, mp.Grid((8, 8, 1), origin=(0, 0, nz + 2)), name="other")
        other.msat = 1e6
        other.magnetization = (0.3, 0.4, 0.8)
    world._impl.set_option("field_zchunk", zchunk)
    assert world._impl.get_option("field_zchunk") == zchunk
    return world, mag


CASES = [
    ((64, 4, 1), {}),
    ((128, 12, 5), {}),
    ((96, 10, 7), {"cubic": True}),
    ((70, 9, 11), {"openbc": True}),
    ((64, 8, 3), {"second": True}),
]


@pytest.mark.parametrize("size,kw", CASES)
@pytest.mark.parametrize("zchunk", [1, 3, 8, 64])
def test_field_and_torque_equal_per_cell_kernel(size, kw, zchunk):
    w0, m0 = _magnet(0, size, **kw)
    w1, m1 = _magnet(zchunk, size, **kw)
    for q in ("effective_field", "torque", "exchange_field", "dmi_field"):
        a = getattr(m0, q).eval()
        b = getattr(m1, q).eval()
        scale = np.abs(a).max()
        assert scale > 0, q
        # a few ulps of the largest term: 1e-6 relative to the field's scale
        assert np.abs(a - b).max() <= 1e-6 * scale, q


@pytest.mark.parametrize("method", ["DormandPrince", "Fehlberg"])
def test_runs_equal_per_cell_kernel(method):
    out = []
    for zc in (0, 4):
        world, mag = _magnet(zc, (80, 16, 6), seed=4)
        ts = world.timesolver
        ts.set_method(method)
        ts.adaptive_timestep = False
        ts.timestep = 4e-14
        ts.steps(12)
        ts.adaptive_timestep = True
        ts.run(3e-13)
        out.append((mag.magnetization.get(), ts.time, ts.timestep))
    (a, ta, da), (b, tb, db) = out
    assert np.abs(a - b).max() < 1e-5
    assert ta == tb
    assert da == pytest.approx(db, rel=1e-3)


def test_zchunk_option():
    world, _ = _magnet(8, (64, 8, 2))
    assert world._impl.get_option("field_zchunk") == 8
    world._impl.set_option("field_zchunk", -1)
    assert world._impl.get_option("field_zchunk") == -1
    with pytest.raises(ValueError):
        world._impl.set_option("field_zchunk", -2)


@pytest.mark.parametrize("size", [(96, 10, 7), (64, 8, 1)])
def test_autotuned_chunk_equals_every_fixed_chunk(size):
    """field_zchunk -1: the autotuner times chunks 1, 2, 4, 8 (those below nz)
    on the first launch for the grid and records its choice; every chunk runs
    the same cell code, so fields, torques and a run equal each fixed chunk's
    bit for bit."""
    mp._cpp._field_tile_clear()
    res = {}
    for zc in (-1, 1, 2, 4, 8):
        world, mag = _magnet(zc, size, seed=2)
        h = mag.effective_field.eval()
        t = mag.torque.eval()
        ts = world.timesolver
        ts.set_method("DormandPrince")
        ts.adaptive_timestep = False
        ts.timestep = 4e-14
        ts.steps(6)
        res[zc] = (h, t, mag.magnetization.get())
    for zc in (1, 2, 4, 8):
        for a, b in zip(res[-1], res[zc]):
            assert np.array_equal(a, b), zc
    rep = mp._cpp._field_tile_report()
    mine = {k: v for k, v in rep.items() if k[:3] == size}
    assert mine, rep
    for rec in mine.values():
        assert rec["choice"] in (1, 2, 4, 8)
        assert rec["choice"] == 1 or rec["choice"] < size[2] * 2
        timed = [zc for zc in (1, 2, 4, 8) if rec[zc] is not None]
        assert timed and timed[0] == 1
        assert all(rec[zc] > 0 for zc in timed)
        assert rec[rec["choice"]] == min(rec[zc] for zc in timed)
        # chunks at or beyond nz beyond the first such one are not timed
        for zc in (2, 4, 8):
            if zc // 2 >= size[2]:
                assert rec[zc] is None
    mp._cpp._field_tile_clear()
    assert not mp._cpp._field_tile_report()
