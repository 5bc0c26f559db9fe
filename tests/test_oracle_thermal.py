"""Pins of oracle/thermal.py (the thermal field of the time steps)."""

import numpy as np
import pytest

from oracle import model as omodel
from oracle import thermal as oth


@pytest.mark.parametrize("ctr,key,want", [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
])
def test_philox_known_answers(ctr, key, want):
    """Philox4x32-10 known-answer vectors of the Random123 distribution."""
    got = oth.philox4x32_10([np.array([v], dtype=np.uint64) for v in ctr], key)
    assert tuple(int(g[0]) for g in got) == want


def test_normals_have_standard_moments_and_no_correlation():
    z = oth.normals(12345, 400000, 0, 7)
    n = z.shape[1]
    for c in range(3):
        assert abs(z[c].mean()) < 5 / np.sqrt(n)
        assert abs(z[c].var() - 1) < 10 / np.sqrt(n)
        assert abs((z[c] ** 4).mean() - 3) < 60 / np.sqrt(n)  # Gaussian kurtosis
    for a, b in ((0, 1), (0, 2), (1, 2)):
        assert abs(np.mean(z[a] * z[b])) < 5 / np.sqrt(n)
    # other draws, purposes and seeds give other numbers
    assert not np.allclose(z[:, :10], oth.normals(12345, 10, 0, 8))
    assert not np.allclose(z[:, :10], oth.normals(12345, 10, 1, 7))
    assert not np.allclose(z[:, :10], oth.normals(12346, 10, 0, 7))


def test_field_variance_is_the_closed_form():
    """<h_i^2> = 2 kB gamma alpha T / ((1 + alpha^2) Ms V) (thermalnoise.cu:40-43, 66-67)."""
    ow = omodel.World((2e-9, 3e-9, 1e-9))
    mag = ow.add_ferromagnet((64, 64, 8), defer_timestep=True)
    mag.set("alpha", 0.1)
    mag.set("temperature", 300.0)
    mag.set("msat", 8e5)
    h = oth.thermal_field(mag, 0, 3)
    want = 2 * 1.38064852e-23 * 1.7595e11 * 0.1 * 300.0 / ((1 + 0.01) * 8e5 * 6e-27)
    n = h[0].size
    for c in range(3):
        assert abs(h[c].var() / want - 1) < 10 / np.sqrt(n)
    # zero outside the geometry and where msat is zero
    geom = np.ones((8, 64, 64), dtype=bool)
    geom[:, :, :5] = False
    mag2 = ow.add_ferromagnet((64, 64, 8), geometry=geom, defer_timestep=True)
    mag2.set("temperature", 300.0)
    mag2.set("alpha", 0.1)
    assert np.all(oth.thermal_field(mag2, 0, 0)[:, :, :, :5] == 0)
