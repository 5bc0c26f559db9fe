"""Pins of the oracle's demag convolution (oracle/demag.py)."""

import numpy as np
import pytest

from oracle import demag
from test_oracle_newell import aharoni_dz


def aharoni_nzz(a, b, c):
    return aharoni_dz(a, b, c)


def _rand_m(rng, shape):
    m = rng.normal(size=(3,) + shape)
    return m / np.linalg.norm(m, axis=0)


@pytest.mark.parametrize("size,cs,geom,mg,pbc", [
    ((4, 3, 2), (1e-9, 1.5e-9, 2e-9), False, (0, 0, 0), (0, 0, 0)),
    ((5, 4, 1), (2e-9, 1e-9, 3e-9), True, (0, 0, 0), (0, 0, 0)),
    ((3, 3, 3), (1e-9, 1e-9, 1e-9), True, (3, 0, 3), (1, 0, 2)),
    ((1, 6, 2), (1e-9, 2e-9, 1e-9), False, (0, 0, 0), (0, 0, 0)),
])
def test_fft_equals_direct_sum(size, cs, geom, mg, pbc):
    rng = np.random.default_rng(0)
    nx, ny, nz = size
    g = rng.random((nz, ny, nx)) > 0.3 if geom else None
    m = _rand_m(rng, (nz, ny, nx))
    ms = 1e5 + 1e6 * rng.random((nz, ny, nx))
    h1 = demag.demag_fft(m, ms, cs, g, mg, pbc)
    h2 = demag.demag_brute(m, ms, cs, g, mg, pbc)
    assert np.abs(h1 - h2).max() <= 1e-12 * np.abs(h2).max()


def test_equals_reference_test_layout():
    """Circular placement + crop == the reference's pre-shifted kernel grid."""
    rng = np.random.default_rng(1)
    m = _rand_m(rng, (3, 4, 16))
    cs = (1e-9, 1e-9, 1e-9)
    h1 = demag.demag_fft(m, 8e5, cs)
    h2 = demag.demag_reference_layout(m, 8e5, cs)
    assert np.abs(h1 - h2).max() <= 1e-12 * np.abs(h2).max()


@pytest.mark.parametrize("n,cs", [((4, 4, 4), (1e-9, 1e-9, 1e-9)), ((10, 6, 3), (2e-9, 3e-9, 1.5e-9)),
                                  ((7, 5, 1), (1e-9, 2e-9, 0.5e-9))])
def test_uniform_prism_average_is_aharoni(n, cs):
    """Cell-averaged Newell fields of uniformly magnetised cells sum exactly to
    the prism's field, so <H_z> = -mu0 Ms Dz(prism) (Aharoni's closed form)."""
    nx, ny, nz = n
    m = np.zeros((3, nz, ny, nx))
    m[2] = 1.0
    h = demag.demag_fft(m, 1.0, cs)
    got = -h[2].mean() / demag.MU0
    assert got == pytest.approx(aharoni_nzz(nx * cs[0], ny * cs[1], nz * cs[2]), rel=1e-10)
    m[:] = 0
    m[0] = 1.0
    got = -demag.demag_fft(m, 1.0, cs)[0].mean() / demag.MU0
    assert got == pytest.approx(aharoni_nzz(ny * cs[1], nz * cs[2], nx * cs[0]), rel=1e-10)


def test_thin_film_limit():
    """Periodic thin film magnetised out of plane: H_z -> -mu0 Ms."""
    n = 8
    m = np.zeros((3, 1, n, n))
    m[2] = 1.0
    h = demag.demag_fft(m, 1e6, (5e-9, 5e-9, 1e-9), mastergrid=(n, n, 0), pbc=(12, 12, 0))
    assert -h[2].mean() / (demag.MU0 * 1e6) == pytest.approx(1.0, abs=5e-3)


def test_octant_kernel_equals_direct():
    a = demag.kernel_offsets((6, 5, 4), (1e-9, 2e-9, 3e-9), (6, 0, 4), (1, 0, 1))
    b = demag.kernel_offsets_octant((6, 5, 4), (1e-9, 2e-9, 3e-9), (6, 0, 4), (1, 0, 1))
    assert np.abs(a - b).max() <= 1e-11 * np.abs(a).max()


def test_spectrum_is_real():
    spec, _ = demag.kernel_spectrum((9, 7, 5), (1e-9, 1.3e-9, 2e-9))
    assert np.abs(spec.imag).max() <= 1e-12 * np.abs(spec.real).max()
