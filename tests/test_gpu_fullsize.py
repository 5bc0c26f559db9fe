"""The benchmark workload at its full size (256 x 256 x 64 cells of 2 nm,
exchange + interfacial DMI + uniaxial anisotropy + demag, Dormand-Prince),
checked through properties that do not depend on the size (the FP64 oracle
is too slow at 4.2 million cells):

* the average demag field of the uniformly magnetised prism is Aharoni's
  closed-form demag factor along each axis (J. Appl. Phys. 83, 3432 (1998)),
  and the field is linear in Ms;
* the torque is perpendicular to m in every cell;
* with damping, 20 steps keep |m| = 1 and lower the total energy at every
  step; without damping, the energy changes by far less than the damped
  decrease, and the step is reproducible bit for bit.
"""

import numpy as np
import pytest

from oracle import demag as odemag
from test_oracle_demag import aharoni_nzz

import mumaxplus as mp

pytestmark = pytest.mark.gpu

N3 = (256, 256, 64)
C = 2e-9


def _magnet(alpha):
    world = mp.World((C, C, C))
    mag = mp.Ferromagnet(world, mp.Grid(N3))
    mag.msat, mag.aex, mag.ku1, mag.anisU, mag.alpha = 1.1e6, 1.6e-11, 8e5, (0, 0, 1), alpha
    mag.dmi_tensor.set_interfacial_dmi(3e-3)
    return world, mag


def _smooth_m():
    nx, ny, nz = N3
    x = np.arange(nx)[None, None, :] / nx
    y = np.arange(ny)[None, :, None] / ny
    th = 0.4 + 0.3 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y) + 0 * np.arange(nz)[:, None, None]
    ph = 2 * np.pi * (x + 0.5 * y)
    return np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)]).astype(np.float32)


def test_uniform_prism_demag_is_aharoni_and_linear():
    world = mp.World((C, C, C))
    mag = mp.Ferromagnet(world, mp.Grid(N3))
    mag.msat = 1.0
    lx, ly, lz = (n * C for n in N3)
    # N_zz(a, b, c) of a prism a x b x c; the other axes by permuting the sides
    want = {2: aharoni_nzz(lx, ly, lz), 0: aharoni_nzz(ly, lz, lx), 1: aharoni_nzz(lz, lx, ly)}
    assert abs(sum(want.values()) - 1) < 1e-12
    for axis in range(3):
        m = [0.0, 0.0, 0.0]
        m[axis] = 1.0
        mag.magnetization = tuple(m)
        h = mag.demag_field.average()[axis]
        assert abs(-h / odemag.MU0 - want[axis]) < 2e-5
    a = mag.demag_field.eval()
    mag.msat = 2.0
    b = mag.demag_field.eval()
    assert np.abs(b - 2 * a).max() <= 1e-6 * np.abs(b).max()


def test_torque_is_perpendicular_to_m():
    world, mag = _magnet(0.02)
    mag.magnetization = _smooth_m()
    m = mag.magnetization.get().astype(np.float64)
    t = mag.torque.eval().astype(np.float64)
    dot = np.abs((m * t).sum(axis=0)).max()
    assert dot <= 1e-6 * np.linalg.norm(t, axis=0).max()


def test_damped_steps_lower_the_energy_and_keep_the_norm():
    world, mag = _magnet(0.1)
    mag.magnetization = _smooth_m()
    ts = world.timesolver
    ts.set_method("DormandPrince")
    ts.adaptive_timestep = False
    ts.timestep = 5e-14
    e = [mag.total_energy.eval()]
    for _ in range(20):
        ts.steps(1)
        e.append(mag.total_energy.eval())
    assert all(b < a for a, b in zip(e, e[1:]))
    n = np.linalg.norm(mag.magnetization.get().astype(np.float64), axis=0)
    assert np.abs(n - 1).max() < 1e-6
    drop = e[0] - e[-1]
    # without damping the energy is conserved to far below the damped drop
    world2, mag2 = _magnet(0.0)
    mag2.magnetization = _smooth_m()
    ts2 = world2.timesolver
    ts2.set_method("DormandPrince")
    ts2.adaptive_timestep = False
    ts2.timestep = 5e-14
    e0 = mag2.total_energy.eval()
    ts2.steps(20)
    assert abs(mag2.total_energy.eval() - e0) < 0.02 * drop
    # the same steps again give the same state, bit for bit
    first = mag2.magnetization.get()
    mag2.magnetization = _smooth_m()
    ts2.time = 0.0
    ts2.steps(20)
    assert np.array_equal(mag2.magnetization.get(), first)
