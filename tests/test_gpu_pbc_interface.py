"""The periodic-boundary interface of World, as the reference's own tests
state it (test/test_PBC_interface.py:7-177): construction and set_pbc with
valid and invalid (repetitions, mastergrid) pairs, the automatic mastergrid
(the bounding grid of the magnets on the repeated axes), the bounding grid,
setting the mastergrid and the repetitions afterwards, magnets outside the
mastergrid, and unset_pbc.  Errors: ValueError / TypeError for malformed
arguments, IndexError for arguments the magnets contradict.
"""

import numpy as np
import pytest

import mumaxplus as mp
from mumaxplus import Grid, World

pytestmark = pytest.mark.gpu

CS = (1, 1, 1)

GOOD = [((0, 0, 0), ((0, 0, 0), (0, 0, 0))),
        ((2, 0, 0), ((128, 0, 0), (1, 2, 3))),
        ((0, 5, 0), ((0, 109, 0), (8, 0, 6))),
        ((0, 0, 3), ((0, 0, 32), (0, 6, 0))),
        ((2, 3, 4), ((128, 32, 64), (1, 0, 0)))]

BAD = [((-1, 0, 0), (20, 0, 0)), ((0, -1, 0), (0, 16, 0)), ((0, 0, -1), (0, 0, 8)),
       ((0.5, 0.6, 0.7), (1, 1, 1)), ((2, 1), (1, 1, 1)),
       ((0, 0, 0), (128, 0, 0)), ((0, 0, 0), (0, 21, 0)), ((0, 0, 0), (0, 0, 64)),
       ((1, 0, 0), (0, 0, 0)), ((0, 2, 0), (0, 0, 0)), ((0, 0, 3), (0, 0, 0)), ((2, 0, 3), (0, 128, 0))]

PAIRS = [(((4, 2, 1), (0, 0, 0)), ((1, 2, 3), (0, 15, 29))),
         (((2, 4, 3), (-5, 7, 12)), ((8, 3, 4), (0, 1, 0))),
         (((8, 1, 6), (1, 2, 3)), ((7, 2, 3), (-5, 6, -9)))]


def _bounds(grids):
    lo = np.min([g[1] for g in grids], axis=0)
    hi = np.max([np.add(g[0], g[1]) for g in grids], axis=0)
    return tuple(int(v) for v in hi - lo), tuple(int(v) for v in lo)


def test_default_world_has_no_periodicity():
    w = World(CS)
    assert w.pbc_repetitions == (0, 0, 0)
    assert w.mastergrid == Grid((0, 0, 0))


@pytest.mark.parametrize("reps,mg", GOOD)
def test_valid_pairs(reps, mg):
    grid = Grid(*mg)
    w = World(CS, reps, grid)
    assert w.pbc_repetitions == reps and w.mastergrid == grid
    w2 = World(CS)
    w2.set_pbc(reps, grid)
    assert w2.pbc_repetitions == reps and w2.mastergrid == grid


@pytest.mark.parametrize("reps,mg", BAD)
def test_invalid_pairs(reps, mg):
    with pytest.raises((ValueError, TypeError)):
        World(CS, reps, Grid(mg))
    w = World(CS)
    with pytest.raises((ValueError, TypeError)):
        w.set_pbc(reps, Grid(mg))


@pytest.mark.parametrize("g1,g2", PAIRS)
@pytest.mark.parametrize("reps", [(1, 2, 1), (0, 1, 2), (2, 0, 1), (1, 2, 0), (0, 0, 0)])
def test_mastergrid_from_the_magnets(g1, g2, reps):
    w = World(CS)
    mp.Ferromagnet(w, Grid(*g1))
    mp.Ferromagnet(w, Grid(*g2))
    w.set_pbc(reps)
    size, origin = _bounds([g1, g2])
    assert w.pbc_repetitions == reps
    assert w.mastergrid == Grid(tuple(s if r else 0 for s, r in zip(size, reps)), origin)


@pytest.mark.parametrize("g1,g2", PAIRS)
def test_bounding_grid(g1, g2):
    w = World(CS)
    mp.Ferromagnet(w, Grid(*g1))
    mp.Ferromagnet(w, Grid(*g2))
    size, origin = _bounds([g1, g2])
    assert w.bounding_grid == Grid(size, origin)


def test_needs_magnets_or_a_mastergrid():
    w = World(CS)
    with pytest.raises(IndexError):
        w.set_pbc((1, 4, 3))
    with pytest.raises(IndexError):
        w.bounding_grid  # noqa: B018


def test_set_pbc_contradicting_a_magnet():
    w = World(CS, pbc_repetitions=(1, 1, 1), mastergrid=Grid((2, 4, 6)))
    mp.Ferromagnet(w, Grid((2, 4, 6)))
    with pytest.raises(IndexError):
        w.set_pbc(pbc_repetitions=(1, 2, 3), mastergrid=Grid((1, 2, 3)))


@pytest.mark.parametrize("origin", [(-1, 0, 0), (0, -1, 0), (0, 0, -1), (10, 0, 0), (0, 10, 0), (0, 0, 10)])
def test_magnet_outside_the_mastergrid(origin):
    w = World(CS, pbc_repetitions=(1, 1, 1), mastergrid=Grid((2, 4, 6)))
    with pytest.raises(IndexError):
        mp.Ferromagnet(w, Grid((1, 2, 3), origin=origin))


def test_set_mastergrid_and_repetitions_afterwards():
    w = World(CS, pbc_repetitions=(0, 1, 2), mastergrid=Grid((0, 12, 16)))
    w.mastergrid = Grid((0, 1, 6))
    assert w.mastergrid == Grid((0, 1, 6))
    with pytest.raises(ValueError):
        w.mastergrid = Grid((1, 0, 6))  # zeros must match the repetitions
    w2 = World(CS, pbc_repetitions=(0, 1, 2), mastergrid=Grid((0, 6, 2)))
    mp.Ferromagnet(w2, Grid((4, 6, 2)))
    with pytest.raises(IndexError):
        w2.mastergrid = Grid((0, 3, 1))  # smaller than the magnet
    w3 = World(CS, pbc_repetitions=(1, 2, 3), mastergrid=Grid((5, 4, 3)))
    w3.pbc_repetitions = (3, 2, 1)
    assert w3.pbc_repetitions == (3, 2, 1)


@pytest.mark.parametrize("reps", [(0, 1, 2), (1, 0, 2), (1, 2, 0), (-1, 2, 3), (4, 0.5, 6)])
def test_invalid_repetitions_afterwards(reps):
    w = World(CS, pbc_repetitions=(1, 2, 3), mastergrid=Grid((5, 4, 3)))
    with pytest.raises((ValueError, TypeError)):
        w.pbc_repetitions = reps


def test_unset_pbc():
    w = World(CS, pbc_repetitions=(1, 2, 3), mastergrid=Grid((128, 32, 26), origin=(-12, 23, 8)))
    w.unset_pbc()
    assert w.pbc_repetitions == (0, 0, 0)
    assert w.mastergrid == Grid((0, 0, 0))


def test_overlapping_magnets_are_refused():
    """MumaxWorld::checkAddibility (mumaxworld.cpp:42-47): a magnet may not
    overlap another one."""
    w = World(CS)
    mp.Ferromagnet(w, Grid((4, 4, 2)))
    with pytest.raises(IndexError):
        mp.Ferromagnet(w, Grid((2, 2, 1), origin=(3, 3, 1)))
    with pytest.raises(IndexError):
        mp.Antiferromagnet(w, Grid((2, 2, 1), origin=(0, 0, 0)))
    mp.Ferromagnet(w, Grid((2, 2, 1), origin=(4, 0, 0)))  # touching is fine
