"""Pins of the oracle's regions and inter-region exchange (oracle/model.py
region_link, oracle/fields.py exchange_field, oracle/afm.py; the reference's
exchange.cu:61-101, afmexchange.cu:96-143, inter_parameter.cu).

The reference has no tests of regions, so the pins come from the
mathematics:

* the exchange field is -(1/(Ms V)) dE/dm of the link energy
  E = V sum_links A_link w_axis |m_i - m_j|^2, written here as a loop over
  the links of in-geometry neighbours, with A_link the harmonic mean inside a
  region and (inter != 0 ? inter : harmonic mean) * scale across regions.
  Without DMI the Neumann boundary adds nothing, so the identity is exact;
* scale 0 between two halves leaves exactly the fields of the two halves as
  separate magnets;
* an inter value equal to the harmonic mean, or half of it with scale 2,
  changes nothing, for the exchange and for the sublattice neighbour
  (afmex_nn) exchange of an antiferromagnet;
* set_between keeps the table symmetric and refuses i == j.
"""

import numpy as np
import pytest

from oracle import afm as oafm
from oracle import fields as of
from oracle import model as om

CS = (2e-9, 3e-9, 2.5e-9)


def _hm(a, b):
    return 0.0 if a + b == 0 else 2 * a * b / (a + b)


def _link_energy_gradient(mag, m):
    """dE/dm of E = V sum_links A_link w |m_i - m_j|^2, by a loop over links."""
    nz, ny, nx = mag.shape
    a = np.broadcast_to(np.asarray(mag.params["aex"], dtype=float), mag.shape)
    geom = mag.geom_mask()
    reg = mag.regions
    V = CS[0] * CS[1] * CS[2]
    grad = np.zeros((3, nz, ny, nx))
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                if not geom[z, y, x]:
                    continue
                for axis, (dx, dy, dz) in enumerate(((1, 0, 0), (0, 1, 0), (0, 0, 1))):
                    x2, y2, z2 = x + dx, y + dy, z + dz
                    if x2 >= nx or y2 >= ny or z2 >= nz or not geom[z2, y2, x2]:
                        continue
                    A = _hm(a[z, y, x], a[z2, y2, x2])
                    r1, r2 = reg[z, y, x], reg[z2, y2, x2]
                    if r1 != r2:
                        inter = mag.inter["inter_exchange"]
                        scale = mag.inter["scale_exchange"]
                        iv = inter if np.ndim(inter) == 0 else inter[r1, r2]
                        sv = scale if np.ndim(scale) == 0 else scale[r1, r2]
                        A = (iv if iv != 0 else A) * sv
                    w = 1.0 / CS[axis] ** 2
                    d = m[:, z, y, x] - m[:, z2, y2, x2]
                    grad[:, z, y, x] += 2 * A * w * V * d
                    grad[:, z2, y2, x2] -= 2 * A * w * V * d
    return grad


def _magnet(rng, size=(6, 5, 2), geometry=True):
    nx, ny, nz = size
    geom = rng.random((nz, ny, nx)) > 0.15 if geometry else None
    mag = om.Ferromagnet(size, CS, geometry=geom)
    mag.enable_demag = False
    mag.set("msat", 5e5 + 5e5 * rng.random((nz, ny, nx)))
    mag.set("aex", 1e-11 * (0.5 + rng.random((nz, ny, nx))))
    m = rng.normal(size=(3, nz, ny, nx))
    mag.set_m(m / np.linalg.norm(m, axis=0))
    mag.regions = rng.integers(0, 3, size=(nz, ny, nx))
    return mag


def test_exchange_is_the_gradient_of_the_link_energy():
    rng = np.random.default_rng(1)
    mag = _magnet(rng)
    mag.set_between("inter_exchange", 0, 1, 2.2e-11)   # pair (0, 2) keeps the harmonic mean
    mag.set_between("inter_exchange", 1, 2, 0.4e-11)
    mag.set_between("scale_exchange", 0, 2, 0.3)
    mag.set_between("scale_exchange", 1, 2, 1.7)
    h = of.exchange_field(mag, mag.m)
    grad = _link_energy_gradient(mag, mag.m)
    V = CS[0] * CS[1] * CS[2]
    ms = mag.val("msat")
    want = np.where(mag.geom_mask()[None], -grad / (ms * V), 0.0)
    assert np.abs(h - want).max() <= 1e-12 * np.abs(want).max()
    # the same with uniform inter-region values
    mag.inter = {"inter_exchange": 0.9e-11, "scale_exchange": 0.6}
    h = of.exchange_field(mag, mag.m)
    want = np.where(mag.geom_mask()[None], -_link_energy_gradient(mag, mag.m) / (ms * V), 0.0)
    assert np.abs(h - want).max() <= 1e-12 * np.abs(want).max()


def test_scale_zero_decouples_two_halves():
    rng = np.random.default_rng(2)
    nx, ny, nz = 8, 4, 2
    mag = _magnet(rng, (nx, ny, nz), geometry=False)
    reg = np.zeros((nz, ny, nx), dtype=int)
    reg[:, :, 5:] = 1
    mag.regions = reg
    mag.inter["scale_exchange"] = 0.0  # assured zero switches exchange off entirely (exchange.cu:14-18)
    assert mag.exchange_assured_zero()
    mag.inter["scale_exchange"] = 1.0
    mag.set_between("scale_exchange", 0, 1, 0.0)
    h = of.exchange_field(mag, mag.m)
    for sl in (slice(0, 5), slice(5, nx)):
        part = om.Ferromagnet((sl.stop - sl.start, ny, nz), CS)
        part.enable_demag = False
        part.set("msat", mag.params["msat"][:, :, sl])
        part.set("aex", mag.params["aex"][:, :, sl])
        part.set_m(mag.m[:, :, :, sl])
        hp = of.exchange_field(part, part.m)
        assert np.abs(h[:, :, :, sl] - hp).max() <= 1e-12 * np.abs(hp).max()


def test_inter_value_equal_to_the_harmonic_mean_changes_nothing():
    rng = np.random.default_rng(3)
    nx, ny, nz = 7, 5, 1
    mag = _magnet(rng, (nx, ny, nz), geometry=False)
    reg = np.zeros((nz, ny, nx), dtype=int)
    reg[:, :, 4:] = 1
    mag.regions = reg
    a0, a1 = 1.1e-11, 1.9e-11
    mag.set("aex", np.where(reg == 0, a0, a1).astype(float))
    plain = of.exchange_field(mag, mag.m)
    mag.set_between("inter_exchange", 0, 1, _hm(a0, a1) / 2)
    mag.set_between("scale_exchange", 0, 1, 2.0)
    assert np.abs(of.exchange_field(mag, mag.m) - plain).max() <= 1e-12 * np.abs(plain).max()
    # ... and a different value does change the field at the interface only
    mag.set_between("inter_exchange", 0, 1, 3e-11)
    d = np.abs(of.exchange_field(mag, mag.m) - plain).max(axis=(0, 1, 2))
    assert d[3] > 0 and d[4] > 0 and d[:3].max() == 0 and d[5:].max() == 0


def test_afm_neighbour_exchange_inter_value_equal_to_the_harmonic_mean():
    rng = np.random.default_rng(4)
    nx, ny, nz = 7, 4, 2
    w = om.World(CS)
    a = w.add_antiferromagnet((nx, ny, nz), defer_timestep=True)
    for sub in a.sublattices:
        sub.enable_demag = False
        sub.set("msat", 6e5)
        sub.set("aex", 1.2e-11)
        m = rng.normal(size=(3, nz, ny, nx))
        sub.set_m(m / np.linalg.norm(m, axis=0))
    reg = np.zeros((nz, ny, nx), dtype=int)
    reg[:, :, 3:] = 2
    a.regions = reg
    n0, n1 = -0.6e-12, -1.4e-12
    a.set("afmex_nn", np.where(reg == 0, n0, n1).astype(float))
    s1, s2 = a.sub1, a.sub2
    plain = oafm.afm_nn_field(s1, s1.m, s2.m)
    a.set_between("inter_afmex_nn", 0, 2, _hm(n0, n1) / 4)
    a.set_between("scale_afmex_nn", 0, 2, 4.0)
    assert np.abs(oafm.afm_nn_field(s1, s1.m, s2.m) - plain).max() <= 1e-12 * np.abs(plain).max()
    a.set_between("scale_afmex_nn", 0, 2, 0.0)
    d = np.abs(oafm.afm_nn_field(s1, s1.m, s2.m) - plain).max(axis=(0, 1, 2))
    assert d[2] > 0 and d[3] > 0 and d[:2].max() == 0 and d[4:].max() == 0
    # the sublattice's own exchange follows its inter_exchange / scale_exchange
    ex = oafm.exchange_field_sublattice(s1, s1.m, s2.m)
    s1.set_between("scale_exchange", 0, 2, 0.0)
    d = np.abs(oafm.exchange_field_sublattice(s1, s1.m, s2.m) - ex).max(axis=(0, 1, 2))
    assert d[2] > 0 and d[3] > 0 and d[:2].max() == 0 and d[4:].max() == 0


def test_set_between_is_symmetric_and_refuses_equal_indices():
    mag = om.Ferromagnet((4, 1, 1), CS)
    mag.regions = np.array([[[0, 1, 3, 3]]])
    mag.set_between("inter_exchange", 3, 1, 5.0)
    t = mag.inter["inter_exchange"]
    assert t.shape == (4, 4) and t[1, 3] == t[3, 1] == 5.0 and t[0, 1] == 0.0
    with pytest.raises(ValueError):
        mag.set_between("inter_exchange", 1, 1, 1.0)
