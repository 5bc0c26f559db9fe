"""Host-side utilities of the package (mumaxplus/util/config.py and
formulary.py of the reference): the initial magnetisations and the
characteristic lengths, against their defining properties."""

import numpy as np
import pytest

from mumaxplus.util import (MU0, Circle, antivortex, blochskyrmion, exchange_length, helical_length, Km,
                            magnetic_hardness, neelskyrmion, twodomain, vortex, wall_energy, wall_width)


def test_vortex_and_antivortex():
    f = vortex((1e-9, 2e-9, 0), 5e-9, 1, -1)
    assert f(1e-9, 2e-9, 0) == (0, 0, -1)
    m = np.array(f(11e-9, 2e-9, 0))  # on +x of the centre: counter-clockwise curl along +y
    assert np.isclose(np.linalg.norm(m), 1) and m[1] > 0.99 and m[2] < 0
    a = np.array(antivortex((0, 0, 0), 5e-9, 1, 1)(0, 10e-9, 0))  # on +y: mx = -dx/r = 0, my = +1
    assert np.isclose(np.linalg.norm(a), 1) and a[1] > 0.99
    with pytest.raises(ValueError):
        vortex((0, 0, 0), 1e-9, 2, 1)


def test_skyrmions():
    for fn, radial in ((neelskyrmion, True), (blochskyrmion, False)):
        f = fn((0, 0, 0), 10e-9, 1, 1)
        assert f(0, 0, 0) == (0, 0, 1)
        mx, my, mz = f(10e-9, 0, 0)  # at r = radius: mz = 2 (e^-1 - 1/2)
        assert np.isclose(mz, 2 * (np.exp(-1) - 0.5))
        inplane = abs(mx) if radial else abs(my)
        assert np.isclose(inplane, 1 - abs(mz))
    with pytest.raises(ValueError):
        neelskyrmion((0, 0, 0), 1e-9, 1, 0)


def test_twodomain():
    f = twodomain((1, 0, 0), (0, 1, 0), (-1, 0, 0), 5.0, 1.0)
    assert np.allclose(f(5.0, 0, 0), (0, 1, 0))
    assert np.allclose(f(-100.0, 0, 0), (1, 0, 0)) and np.allclose(f(100.0, 0, 0), (-1, 0, 0))
    assert twodomain((1, 0, 0), (0, 1, 0), (0, 0, 1), 0.0)(1.0, 0, 0) == (0, 0, 1)


def test_formulary_and_shape_export():
    ms, a, k, d = 8e5, 1.3e-11, 5e5, 3e-3
    assert np.isclose(Km(ms), 0.5 * MU0 * ms * ms)
    assert np.isclose(exchange_length(a, ms) ** 2 * Km(ms), a)
    assert np.isclose(wall_width(a, k) * wall_energy(a, k), a)
    assert np.isclose(helical_length(a, d) * d, a)
    assert np.isclose(magnetic_hardness(-k, ms) ** 2 * MU0 * ms * ms, k)
    assert Circle(10e-9) is not None
