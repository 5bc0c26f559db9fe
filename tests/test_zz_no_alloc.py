"""The allocation-free demag exec (mpa_demag_plan_exec_noalloc), on the GPU.

Its FFT intermediates live in the library's static device scratch array,
shared by every call on the GPU and keeping the previous call's contents, so
two calls in a row (the second on a smaller grid) must each equal the plan's
own exec bit for bit: same kernels, same arithmetic.  A plan used only this
way holds no intermediates of its own, and a grid whose intermediates exceed
the scratch is rejected.
"""

import numpy as np
import pytest

from helpers import DeviceBuffer, random_unit

import mumaxplus as mp

pytestmark = pytest.mark.gpu

CS = (2e-9, 2.5e-9, 3e-9)


def _inputs(size, seed):
    nx, ny, nz = size
    rng = np.random.default_rng(seed)
    m = random_unit(rng, (nz, ny, nx)).astype(np.float32)
    msat = rng.uniform(4e5, 1.2e6, size=(nz, ny, nx)).astype(np.float32)
    geom = (rng.random((nz, ny, nx)) > 0.2).astype(np.uint8)
    return m, msat, geom


def _both(plan, size, seed, per_cell):
    """The field of the noalloc exec and of the plan's own exec."""
    m, msat, geom = _inputs(size, seed)
    md = DeviceBuffer.from_host(m)
    sd = DeviceBuffer.from_host(msat) if per_cell else None
    gd = DeviceBuffer.from_host(geom) if per_cell else None
    h1 = DeviceBuffer(m.nbytes)
    h2 = DeviceBuffer(m.nbytes)
    kw = dict(msat_uniform=8e5, msat_ptr=sd.ptr if sd else None, geom_ptr=gd.ptr if gd else None)
    plan.exec_noalloc(md.ptr, h1.ptr, **kw)
    a = h1.to_host(np.float32, m.shape)
    plan.exec(md.ptr, h2.ptr, **kw)
    b = h2.to_host(np.float32, m.shape)
    return a, b


def test_two_calls_in_a_row_equal_the_plan_exec():
    big, small = (40, 24, 6), (16, 8, 2)
    p1 = mp._cpp.DemagPlan(big, CS)
    p2 = mp._cpp.DemagPlan(small, CS)
    own1, own2 = p1.info()[2], p2.info()[2]
    # both noalloc calls first: the second runs on the scratch the first left
    m1, ms1, g1 = _inputs(big, 1)
    m2, ms2, g2 = _inputs(small, 2)
    md1, md2 = DeviceBuffer.from_host(m1), DeviceBuffer.from_host(m2)
    sd1, gd1 = DeviceBuffer.from_host(ms1), DeviceBuffer.from_host(g1)
    hn1, hn2 = DeviceBuffer(m1.nbytes), DeviceBuffer(m2.nbytes)
    p1.exec_noalloc(md1.ptr, hn1.ptr, msat_ptr=sd1.ptr, geom_ptr=gd1.ptr)
    p2.exec_noalloc(md2.ptr, hn2.ptr, msat_uniform=7e5)
    a1 = hn1.to_host(np.float32, m1.shape)
    a2 = hn2.to_host(np.float32, m2.shape)
    # the plans hold no intermediates of their own after the noalloc calls
    assert p1.info()[2] == own1 and p2.info()[2] == own2
    he1, he2 = DeviceBuffer(m1.nbytes), DeviceBuffer(m2.nbytes)
    p1.exec(md1.ptr, he1.ptr, msat_ptr=sd1.ptr, geom_ptr=gd1.ptr)
    p2.exec(md2.ptr, he2.ptr, msat_uniform=7e5)
    b1 = he1.to_host(np.float32, m1.shape)
    b2 = he2.to_host(np.float32, m2.shape)
    assert p1.info()[2] > own1  # the plan's own exec allocated them
    assert np.abs(b1).max() > 0 and np.abs(b2).max() > 0
    assert np.array_equal(a1, b1)
    assert np.array_equal(a2, b2)
    # outside the geometry the field is zero
    assert np.all(a1[:, g1 == 0] == 0)


@pytest.mark.parametrize("size,per_cell", [((33, 17, 1), True), ((64, 64, 1), False), ((20, 12, 9), True)])
def test_single_layer_and_odd_grids(size, per_cell):
    plan = mp._cpp.DemagPlan(size, CS)
    a, b = _both(plan, size, 5, per_cell)
    assert np.abs(b).max() > 0
    assert np.array_equal(a, b)


def test_periodic_plan():
    size = (24, 16, 2)
    plan = mp._cpp.DemagPlan(size, CS, mastergrid=(24, 16, 0), pbc=(2, 1, 0))
    a, b = _both(plan, size, 9, True)
    assert np.array_equal(a, b)


def test_rejects_grids_beyond_the_scratch():
    plan = mp._cpp.DemagPlan((256, 256, 16), (2e-9, 2e-9, 2e-9))
    m = DeviceBuffer(4 * 3 * 256 * 256 * 16)
    h = DeviceBuffer(4 * 3 * 256 * 256 * 16)
    with pytest.raises(ValueError, match="scratch"):
        plan.exec_noalloc(m.ptr, h.ptr)
