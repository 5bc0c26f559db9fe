"""Parameter assignment semantics and averages, as the reference's own tests
state them (test/test_parameter.py:19-319, test/test_average.py:7-14).

A parameter takes a uniform value, a per-cell array, a function of the cell
position, a time-dependent term t -> value (alone or with a mask given as an
array, with or without the component axis, or as a function of position),
several terms at once, and removal of the terms.  is_uniform / is_dynamic and
eval() follow from the assignment in closed form; a mask of the wrong shape
is rejected with ValueError.  The average of a field equals numpy's average
of the same field.
"""

import numpy as np
import pytest

import mumaxplus as mp

pytestmark = pytest.mark.gpu

SHAPE3 = (3, 7, 4)  # (nz, ny, nx) of a 4 x 7 x 3 grid
T_NOW = 0.5


def sinc_term(t):
    return 24 * np.sinc(t)


def vec_term(t):
    return (0.0, 1.0, 0.25 * np.sin(t))


@pytest.fixture
def magnet():
    world = mp.World(cellsize=(1e-9, 1e-9, 1e-9))
    world.timesolver.time = T_NOW
    return mp.Ferromagnet(world, mp.Grid((4, 7, 3)))


def _full(ncomp, values):
    out = np.empty((ncomp,) + SHAPE3)
    for c in range(ncomp):
        out[c] = values[c]
    return out


def test_uniform_scalar_and_vector(magnet):
    magnet.ku1 = 1e6
    assert not magnet.ku1.is_dynamic and magnet.ku1.is_uniform
    np.testing.assert_allclose(magnet.ku1.eval(), 1e6)
    magnet.bias_magnetic_field = (0, 0, 1.5)
    assert not magnet.bias_magnetic_field.is_dynamic and magnet.bias_magnetic_field.is_uniform
    np.testing.assert_allclose(magnet.bias_magnetic_field.eval(), _full(3, (0, 0, 1.5)))


def test_per_cell_arrays(magnet):
    rng = np.random.default_rng(2)
    a = rng.random((1,) + SHAPE3)
    b = rng.random((3,) + SHAPE3)
    magnet.alpha = a
    magnet.bias_magnetic_field = b
    assert not magnet.alpha.is_uniform and not magnet.bias_magnetic_field.is_uniform
    np.testing.assert_allclose(magnet.alpha.eval(), a, rtol=1e-6)
    np.testing.assert_allclose(magnet.bias_magnetic_field.eval(), b, rtol=1e-6)


def test_scalar_time_term_and_removal(magnet):
    magnet.ku1 = sinc_term
    assert magnet.ku1.is_dynamic and magnet.ku1.is_uniform
    np.testing.assert_allclose(magnet.ku1.eval(), sinc_term(T_NOW), rtol=1e-6)
    magnet.ku1.remove_time_terms()
    assert not magnet.ku1.is_dynamic
    np.testing.assert_allclose(magnet.ku1.eval(), 0.0)


@pytest.mark.parametrize("mask_kind", ["array", "array_no_comp", "function"])
def test_scalar_time_term_with_mask(magnet, mask_kind):
    for value in (0.2, 0.8):
        if mask_kind == "array":
            mask = value * np.ones((1,) + SHAPE3)
        elif mask_kind == "array_no_comp":
            mask = value * np.ones(SHAPE3)
        else:
            mask = lambda x, y, z, v=value: v  # noqa: E731
        magnet.ku1 = (sinc_term, mask)
        assert magnet.ku1.is_dynamic and not magnet.ku1.is_uniform
        np.testing.assert_allclose(magnet.ku1.eval(), value * sinc_term(T_NOW), rtol=1e-6)


def test_scalar_mask_of_wrong_shape_is_rejected(magnet):
    with pytest.raises(ValueError):
        magnet.ku2 = (sinc_term, np.ones((1, 4, 7, 3)))  # x fastest is the last axis
    with pytest.raises(ValueError):
        magnet.ku2 = (sinc_term, lambda x, y, z: (0.2, 0.8))


def test_several_scalar_terms(magnet):
    term2 = lambda t: -24 * np.sinc(t)  # noqa: E731
    magnet.alpha = 0.5
    magnet.alpha.add_time_term(sinc_term)
    magnet.alpha.add_time_term(term2, 0.001 * np.ones((1,) + SHAPE3))
    assert magnet.alpha.is_dynamic and not magnet.alpha.is_uniform
    want = 0.5 + sinc_term(T_NOW) + 0.001 * term2(T_NOW)
    np.testing.assert_allclose(magnet.alpha.eval(), want, rtol=1e-6)
    magnet.alpha.remove_time_terms()
    assert not magnet.alpha.is_dynamic
    np.testing.assert_allclose(magnet.alpha.eval(), 0.5, rtol=1e-7)


def test_vector_time_term_and_removal(magnet):
    magnet.bias_magnetic_field = vec_term
    assert magnet.bias_magnetic_field.is_dynamic and magnet.bias_magnetic_field.is_uniform
    np.testing.assert_allclose(magnet.bias_magnetic_field.eval(), _full(3, vec_term(T_NOW)), rtol=1e-6,
                               atol=1e-7)
    magnet.bias_magnetic_field.remove_time_terms()
    assert not magnet.bias_magnetic_field.is_dynamic
    np.testing.assert_allclose(magnet.bias_magnetic_field.eval(), 0.0)


@pytest.mark.parametrize("mask_kind", ["array", "array_no_comp", "array_per_comp", "function",
                                       "function_scalar", "function_per_comp"])
def test_vector_time_term_with_mask(magnet, mask_kind):
    scales = {"array": (0.2,) * 3, "array_no_comp": (0.8,) * 3, "array_per_comp": (0.1, 0.2, 0.3),
              "function": (0.2,) * 3, "function_scalar": (0.8,) * 3, "function_per_comp": (0.1, 0.2, 0.3)}[mask_kind]
    if mask_kind == "array":
        mask = 0.2 * np.ones((3,) + SHAPE3)
    elif mask_kind == "array_no_comp":
        mask = 0.8 * np.ones(SHAPE3)
    elif mask_kind == "array_per_comp":
        mask = _full(3, scales)
    elif mask_kind == "function":
        mask = lambda x, y, z: (0.2, 0.2, 0.2)  # noqa: E731
    elif mask_kind == "function_scalar":
        mask = lambda x, y, z: 0.8  # noqa: E731
    else:
        mask = lambda x, y, z: (0.1, 0.2, 0.3)  # noqa: E731
    magnet.bias_magnetic_field = (vec_term, mask)
    assert magnet.bias_magnetic_field.is_dynamic and not magnet.bias_magnetic_field.is_uniform
    want = _full(3, [s * v for s, v in zip(scales, vec_term(T_NOW))])
    np.testing.assert_allclose(magnet.bias_magnetic_field.eval(), want, rtol=1e-6, atol=1e-7)


def test_vector_mask_of_wrong_shape_is_rejected(magnet):
    with pytest.raises(ValueError):
        magnet.bias_magnetic_field = (vec_term, np.ones((1, 4, 7, 3)))
    with pytest.raises(ValueError):
        magnet.bias_magnetic_field = (vec_term, lambda x, y, z: (0.2, 0.8))


def test_several_vector_terms(magnet):
    term1 = lambda t: np.array((0.0, 1.0, 0.25 * np.sin(t)))  # noqa: E731
    term2 = lambda t: np.array((0.0, 0.0, -24 * np.sinc(t)))  # noqa: E731
    magnet.bias_magnetic_field = (1, 1, 0)
    magnet.bias_magnetic_field.add_time_term(term1)
    magnet.bias_magnetic_field.add_time_term(term2, 0.001 * np.ones((3,) + SHAPE3))
    assert magnet.bias_magnetic_field.is_dynamic and not magnet.bias_magnetic_field.is_uniform
    want = np.array((1, 1, 0)) + term1(T_NOW) + 0.001 * term2(T_NOW)
    np.testing.assert_allclose(magnet.bias_magnetic_field.eval(), _full(3, want), rtol=1e-6, atol=1e-7)
    magnet.bias_magnetic_field.remove_time_terms()
    assert not magnet.bias_magnetic_field.is_dynamic
    np.testing.assert_allclose(magnet.bias_magnetic_field.eval(), _full(3, (1, 1, 0)))


def test_average_equals_numpy_average():
    """test/test_average.py:7-14: a 32 x 32 x 5 grid at origin (-10, 3, 0)
    with the default random magnetisation."""
    world = mp.World((1, 1, 1))
    mag = mp.Ferromagnet(world, mp.Grid((32, 32, 5), (-10, 3, 0)))
    m = mag.magnetization.get().astype(np.float64)
    want = m.mean(axis=(1, 2, 3))
    got = np.array(mag.magnetization.average())
    assert np.all(np.abs((want - got) / got) < 1e-5)
