"""Antiferromagnet checks the reference's own tests hold, re-expressed.

* test/test_domain_wall_afm.py:22-143 (Sanchez-Tejerina et al., PRB 101,
  014433): a 1 x 100 x 200 antiferromagnet of 2 nm cells without demag, a
  two-domain state along z.  The relaxed wall fits the Walker profile
  cos(2 atan(exp(-(z - q)/w))) with w = sqrt((2A - A_nn)/(2K)), and under a
  Slonczewski current the wall moves at v = pi w H_sh / (2 alpha Ms/gamma),
  H_sh = hbar P J / (2 e d); both within 0.5 %.
* test/test_spinwave_dispersion_afm.py:14-99: a 512-cell chain excited by a
  sinc pulse in the middle; the peak frequency of every wave number within
  |k| < 1/dx follows the analytical two-sublattice dispersion within 3 % of
  the pulse bandwidth.
"""

import numpy as np
import pytest
from scipy.optimize import curve_fit

import mumaxplus as mp
from mumaxplus.util.constants import GAMMALL, HBAR, QE

pytestmark = pytest.mark.gpu

# ---------------------------------------------------------------- domain wall
C2 = 2e-9
NY_W, NZ_W = 100, 200
Z_FIT = np.linspace(-NY_W * C2, NY_W * C2, NZ_W)  # the positions the reference fits on


def _wall_magnet(afmex_cell):
    world = mp.World((C2, C2, C2))
    afm = mp.Antiferromagnet(world, mp.Grid((1, NY_W, NZ_W)))
    afm.enable_demag = False
    afm.msat = 0.4e6
    afm.alpha = 0.1
    afm.ku1 = 64e3
    afm.anisU = (1, 0, 0)
    afm.aex = 10e-12
    afm.afmex_cell = afmex_cell
    afm.afmex_nn = -5e-12
    afm.sub1.dmi_tensor.set_interfacial_dmi(0.11e-3)
    afm.sub2.dmi_tensor.set_interfacial_dmi(0.11e-3)
    # two domains along z with a 4-cell wall, relaxed by 10 ps of dynamics
    m = np.zeros((3, NZ_W, NY_W, 1))
    h, w2 = NZ_W // 2, 2
    m[0, :h - w2] = -1
    m[2, h - w2:h + w2] = 1
    m[0, h + w2:] = 1
    afm.sub1.magnetization = m
    afm.sub2.magnetization = -m
    world.timesolver.run(1e-11)
    return world, afm


def _walker(z, q, w):
    return np.cos(2 * np.arctan(np.exp(-(z - q) / w)))


def _fit_wall(afm):
    prof = afm.sub1.magnetization.get()[0][:, NY_W // 2, 0]
    popt, _ = curve_fit(_walker, Z_FIT, prof, p0=(1e-9, 5e-9))
    return popt


def _analytic_width(afm):
    a, ann, k = afm.sub1.aex.uniform_value, afm.afmex_nn.uniform_value, afm.sub1.ku1.uniform_value
    return np.sqrt((2 * a - ann) / (2 * k))


def test_domain_wall_width():
    _, afm = _wall_magnet(-25e-12)
    w = _fit_wall(afm)[1]
    want = _analytic_width(afm)
    assert abs(w - want) / want < 5e-3, (w, want)


def test_domain_wall_speed_under_slonczewski_current():
    world, afm = _wall_magnet(-15e-12)
    afm.pol = 0.044
    afm.fixed_layer = (0, 1, 0)
    afm.Lambda = 1
    afm.free_layer_thickness = 2e-9
    afm.jcur = (0, 0, 1e12)
    t = 2e-11
    world.timesolver.run(t / 2)
    q1 = _fit_wall(afm)[0]
    world.timesolver.run(t / 2)
    q2 = _fit_wall(afm)[0]
    v = 2 * abs(q1 - q2) / t
    sub = afm.sub1
    h_sh = HBAR * sub.pol.uniform_value * sub.jcur.uniform_value[2] / (2 * QE * sub.free_layer_thickness.uniform_value)
    want = np.pi * _analytic_width(afm) * h_sh / (sub.alpha.uniform_value * 2 * sub.msat.uniform_value / GAMMALL)
    assert abs(v - want) / want < 5e-3, (v, want)


# ---------------------------------------------------------------- spin waves
def test_spinwave_dispersion():
    fmax, T = 5e13, 2e-12
    dt, dx, nx = 1 / (2 * fmax), 1e-9, 512
    bz, A, A_nn, A_c, Ms, alpha, K, mu0 = 0.2, 10e-12, -5e-12, -400e-12, 400e3, 0.005, 1e3, 1.256637062e-6
    world = mp.World((dx, dx, dx))
    world.bias_magnetic_field = (0, 0, bz)
    afm = mp.Antiferromagnet(world, mp.Grid((nx, 1, 1)))
    for sub in afm.sublattices:
        sub.msat, sub.aex, sub.alpha, sub.ku1, sub.anisU = Ms, A, alpha, K, (0, 0, 1)
    afm.afmex_nn = A_nn
    afm.afmex_cell = A_c
    afm.latcon = dx
    mask = np.zeros((1, 1, nx))
    mask[:, :, nx // 2 - 1:nx // 2 + 1] = 1
    pulse = lambda t: (1e2 * np.sinc(2 * fmax * (t - T / 2)), 0, 0)  # noqa: E731
    for sub in afm.sublattices:
        sub.bias_magnetic_field.add_time_term(pulse, mask)
    afm.sub1.magnetization = (0, 0, 1)
    afm.sub2.magnetization = (0, 0, -1)
    times = np.linspace(0, T, 1 + int(T / dt))
    out = world.timesolver.solve(times, {"m": lambda: afm.sub1.magnetization.eval()})
    mx = np.array(out["m"], dtype=float)[:, 0, 0, 0, :]
    power = np.abs(np.fft.fftshift(np.fft.fft2(mx))) ** 2
    f_sim = np.argmax(power[power.shape[0] // 2:, :], axis=0) / T
    k = np.linspace(-np.pi / dx, np.pi / dx * (nx - 2) / nx, nx)
    he = -4 * A_c / (dx * dx * mu0 * Ms)
    ha = 2 * K / (mu0 * Ms)
    hi = (2 * A - A_nn) / (mu0 * Ms)
    f_th = mu0 * GAMMALL / (2 * np.pi) * np.sqrt((2 * he + ha + hi * k ** 2) * (ha + hi * k ** 2)) \
        + bz * GAMMALL / (2 * np.pi)
    i0, i1 = np.argmin(abs(k + 1 / dx)), np.argmin(abs(k - 1 / dx))
    err = np.max(np.abs(f_sim - f_th)[i0:i1] / fmax)
    assert err < 3e-2, err
