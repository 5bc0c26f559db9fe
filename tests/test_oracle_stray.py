"""Pins of oracle/stray.py (stray fields between grids, FP64 direct sum)."""

import numpy as np
import pytest

from oracle import demag as od
from oracle import stray as ost

CS = (1e-9, 2e-9, 3.1e-9)
MU0 = 1.25663706212e-6


def _M(rng, shape):
    m = rng.normal(size=(3,) + shape)
    return 8e5 * m / np.linalg.norm(m, axis=0)


def test_same_grid_is_the_demag_field():
    rng = np.random.default_rng(1)
    M = _M(rng, (2, 3, 4))
    h = ost.stray_field(M, (3, -1, 2), (4, 3, 2), (3, -1, 2), CS)
    want = od.demag_brute(M / 8e5, 8e5, CS)
    assert np.abs(h - want).max() <= 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("case", [
    ((8, 10, 1), (2, 15, 0), (6, 7, 2), (-5, 0, 2)),
    ((8, 10, 2), (2, 15, -1), (12, 12, 1), (-5, 0, 1)),
    ((5, 4, 2), (0, 0, 0), (5, 4, 2), (5, 0, 0)),
])
@pytest.mark.parametrize("pbc", [None, (1, 2, 0)])
def test_box_identity(case, pbc):
    """test/test_strayfield.py:86-136: the field on the destination is the
    demag field of the enclosing box with M only inside the source."""
    sn, so, dn, do = case
    rng = np.random.default_rng(7)
    M = _M(rng, (sn[2], sn[1], sn[0]))
    bo = [min(so[a], do[a]) for a in range(3)]
    bn = [max(so[a] + sn[a], do[a] + dn[a]) - bo[a] for a in range(3)]
    mg, reps = (0, 0, 0), (0, 0, 0)
    if pbc is not None:
        mg = tuple(bn[a] if pbc[a] else 0 for a in range(3))
        reps = pbc
    h = ost.stray_field(M, so, dn, do, CS, mg, reps)
    Mb = np.zeros((3, bn[2], bn[1], bn[0]))
    s0 = [so[a] - bo[a] for a in range(3)]
    Mb[:, s0[2]:s0[2] + sn[2], s0[1]:s0[1] + sn[1], s0[0]:s0[0] + sn[0]] = M
    hb = od.DemagPlan(tuple(bn), CS, mg, reps).field(Mb, np.ones((bn[2], bn[1], bn[0])))
    d0 = [do[a] - bo[a] for a in range(3)]
    want = hb[:, d0[2]:d0[2] + dn[2], d0[1]:d0[1] + dn[1], d0[0]:d0[0] + dn[0]]
    assert np.abs(h - want).max() <= 1e-9 * np.abs(want).max()


def test_far_field_dipole():
    """One source cell seen from (80, 26, 10) cells away: the point-dipole
    field mu0 V / (4 pi r^3) (3 (M.r^) r^ - M), to 1e-3 (the quadrupole
    terms of the non-cubic cell fall off as 1/r^2: 1.8e-3 at half the
    distance, 4.4e-4 here)."""
    M = np.zeros((3, 1, 1, 1))
    M[:, 0, 0, 0] = (3e5, -2e5, 5e5)
    do = (80, 26, 10)
    h = ost.stray_field(M, (0, 0, 0), (1, 1, 1), do, CS)[:, 0, 0, 0]
    r = np.array(do, dtype=float) * np.array(CS)
    rn = np.linalg.norm(r)
    rh = r / rn
    V = CS[0] * CS[1] * CS[2]
    mv = M[:, 0, 0, 0]
    want = MU0 * V / (4 * np.pi * rn ** 3) * (3 * np.dot(mv, rh) * rh - mv)
    assert np.abs(h - want).max() <= 1e-3 * np.abs(want).max()


@pytest.mark.parametrize("pbc", [None, (1, 1, 0)])
def test_mutual_energy_is_reciprocal(pbc):
    """Two magnets of different sizes and origins: sum M1 . B21 over magnet 1
    equals sum M2 . B12 over magnet 2 (the mutual energy of two magnets does
    not depend on which one is the source); a misplaced destination origin or
    crop breaks the equality."""
    rng = np.random.default_rng(3)
    n1, o1 = (5, 3, 2), (0, 0, 0)
    n2, o2 = (3, 4, 1), (7, 1, 3)
    M1 = _M(rng, (n1[2], n1[1], n1[0]))
    M2 = _M(rng, (n2[2], n2[1], n2[0]))
    mg, reps = (0, 0, 0), (0, 0, 0)
    if pbc is not None:
        mg, reps = (12, 8, 0), pbc
    b21 = ost.stray_field(M2, o2, n1, o1, CS, mg, reps)
    b12 = ost.stray_field(M1, o1, n2, o2, CS, mg, reps)
    e1, e2 = np.sum(M1 * b21), np.sum(M2 * b12)
    assert abs(e1 - e2) <= 1e-10 * max(np.sum(np.abs(M1 * b21)), np.sum(np.abs(M2 * b12)))
    assert abs(e1) > 1e-3 * np.sum(np.abs(M1 * b21))
