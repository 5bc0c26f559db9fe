"""Angle quantities and get_rgb on the GPU, against closed forms.

* Ferromagnet.max_angle (exchange.cu:271-321): a planar spiral with a
  phase step k c between neighbours has max_angle k c; cells without
  exchange (harmonic mean 0), outside the geometry or with msat = 0 do not
  count;
* Antiferromagnet.angle_field / max_intracell_angle (afmexchange.cu:260-305):
  acos(sign(afmex_cell) m1 . m2), i.e. pi - theta for m2 at angle theta from
  m1 and a negative afmex_cell; zero where afmex_cell = 0;
* FieldQuantity.get_rgb (fieldops.cu:316-385): +x red, -x cyan, +z white,
  -z black, +y (hue 90 degrees) yellow-green (0.5, 1, 0); grey outside the
  geometry.
"""

import numpy as np
import pytest

import mumaxplus as mp

pytestmark = pytest.mark.gpu


def test_max_angle_of_a_spiral():
    nx, c, step = 40, 1e-9, 0.1
    world = mp.World((c, c, c))
    mag = mp.Ferromagnet(world, mp.Grid((nx, 2, 1)))
    mag.msat, mag.aex = 8e5, 1e-11
    ph = step * np.arange(nx)
    m = np.stack([np.cos(ph), np.sin(ph), 0 * ph])[:, None, None, :].repeat(2, axis=2)
    mag.magnetization = m.astype(np.float32)
    assert mag.max_angle.eval() == pytest.approx(step, rel=1e-4)
    # a larger step where exchange is off does not count
    ph2 = ph.copy()
    ph2[20:] += 0.5
    m2 = np.stack([np.cos(ph2), np.sin(ph2), 0 * ph2])[:, None, None, :].repeat(2, axis=2)
    mag.magnetization = m2.astype(np.float32)
    assert mag.max_angle.eval() == pytest.approx(0.6, rel=1e-4)
    a = np.full((1, 2, nx), 1e-11)
    a[:, :, 19:21] = 0.0  # both cells of the 0.6 link without exchange
    mag.aex = a
    assert mag.max_angle.eval() == pytest.approx(step, rel=1e-4)
    mag.aex = 0.0
    assert mag.max_angle.eval() == 0.0


def test_afm_angle_field():
    nx = 16
    world = mp.World((1e-9, 1e-9, 1e-9))
    afm = mp.Antiferromagnet(world, mp.Grid((nx, 1, 1)))
    afm.msat = 5e5
    theta = np.linspace(0.2, 1.4, nx)
    afm.sub1.magnetization = (1.0, 0.0, 0.0)
    afm.sub2.magnetization = np.stack([np.cos(theta), np.sin(theta), 0 * theta])[:, None, None, :].astype(np.float32)
    afm.afmex_cell = -1e-12
    got = afm.angle_field.eval()[0, 0, 0]
    assert np.abs(got - (np.pi - theta)).max() < 1e-5
    assert afm.max_intracell_angle.eval() == pytest.approx(np.pi - 0.2, rel=1e-5)
    with pytest.warns(UserWarning):
        afm.afmex_cell = 1e-12
    assert np.abs(afm.angle_field.eval()[0, 0, 0] - theta).max() < 1e-5
    afm.afmex_cell = 0.0
    assert afm.max_intracell_angle.eval() == 0.0


def test_get_rgb_colours():
    world = mp.World((1e-9, 1e-9, 1e-9))
    geom = np.ones((1, 1, 6), dtype=bool)
    geom[0, 0, 5] = False
    mag = mp.Ferromagnet(world, mp.Grid((6, 1, 1)), geometry=geom)
    m = np.array([[1, 0, 0], [-1, 0, 0], [0, 0, 1], [0, 0, -1], [0, 1, 0], [1, 0, 0]], dtype=np.float32)
    mag.magnetization = m.T[:, None, None, :]
    rgb = mag.magnetization.get_rgb()[:, 0, 0, :].T
    want = [[1, 0, 0], [0, 1, 1], [1, 1, 1], [0, 0, 0], [0.5, 1, 0], [0.5, 0.5, 0.5]]
    assert np.abs(rgb - np.array(want)).max() < 1e-6
