"""Thermal field in FP64 (test infrastructure only).

Restates evalThermalNoise / k_thermalNoise (src/physics/thermalnoise.cu:21-70):
per cell sqrt(2 kB gamma alpha T / ((1 + alpha^2) Ms V)) times three standard
normal deviates, zero outside the geometry; the time stepper adds it to
every stage derivative divided by sqrt(dt) and draws it once per step
(rungekutta.cpp:85-106).

The deviates: the reference draws them from a clock-seeded cuRAND generator,
which cannot be reproduced.  The product draws them from Philox4x32-10
(Salmon, Moraes, Dror and Shaw, SC11) keyed by (seed, 0x6d756d78) with the
counter (cell, purpose, draw low, draw high), mapped to normals by
Box-Muller on the top 24 bits of each word: u = (x >> 8 + 0.5) / 2^24 in
FP32, z0, z1 = sqrt(-2 ln u0) (cos, sin)(2 pi u1), z2 = sqrt(-2 ln u2)
cos(2 pi u3).  This module restates the same generator so that a noisy
trajectory can be compared cell by cell.

Pinned by tests/test_oracle_thermal.py: the Philox4x32-10 known-answer
vectors of the Random123 distribution, the moments of the normals, and the
variance of the field against the closed form above.
"""

import numpy as np

M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
W0, W1 = 0x9E3779B9, 0xBB67AE85
KEY1 = 0x6D756D78
KB = 1.38064852e-23  # src/physics/constants.hpp:11
GAMMALL = 1.7595e11
MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(ctr, key):
    """ctr: 4 arrays (or ints) of uint32 words, key: 2 ints; returns 4 uint32 arrays."""
    c = [np.asarray(v, dtype=np.uint64) for v in ctr]
    k0, k1 = int(key[0]) & 0xFFFFFFFF, int(key[1]) & 0xFFFFFFFF
    for r in range(10):
        if r > 0:
            k0 = (k0 + W0) & 0xFFFFFFFF
            k1 = (k1 + W1) & 0xFFFFFFFF
        p0 = M0 * c[0]
        p1 = M1 * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & MASK
        c = [hi1 ^ c[1] ^ np.uint64(k0), lo1, hi0 ^ c[3] ^ np.uint64(k1), lo0]
    return [v.astype(np.uint32) for v in c]


def normals(seed, ncell, purpose, draw):
    """(3, ncell) standard normals for cells 0..ncell-1 of one draw."""
    cells = np.arange(ncell, dtype=np.uint64)
    x = philox4x32_10([cells, np.full(ncell, purpose, np.uint64),
                       np.full(ncell, draw & 0xFFFFFFFF, np.uint64), np.full(ncell, draw >> 32, np.uint64)],
                      (seed, KEY1))
    u = [((xi >> np.uint32(8)).astype(np.float32) + np.float32(0.5)) * np.float32(1.0 / 16777216.0) for xi in x]
    u = [v.astype(np.float64) for v in u]
    r1 = np.sqrt(-2.0 * np.log(u[0]))
    r2 = np.sqrt(-2.0 * np.log(u[2]))
    return np.stack([r1 * np.cos(2 * np.pi * u[1]), r1 * np.sin(2 * np.pi * u[1]), r2 * np.cos(2 * np.pi * u[3])])


def noise_amplitude(mag):
    """sqrt(2 kB gamma alpha T / ((1 + alpha^2) Ms V)) per cell (0 where Ms = 0)."""
    cs = mag.cellsize
    V = cs[0] * cs[1] * cs[2]
    a, T, ms = mag.val("alpha"), mag.val("temperature"), mag.val("msat")
    with np.errstate(divide="ignore", invalid="ignore"):
        amp = np.sqrt(2 * KB * GAMMALL / V * a * T / ((1 + a * a) * ms))
    return np.where(ms != 0, amp, 0.0)


def thermal_field(mag, purpose, draw):
    """The thermal field (3, nz, ny, nx) of draw `draw` of stream `purpose`."""
    nz, ny, nx = mag.shape
    z = normals(mag.noise_seed, nx * ny * nz, purpose, draw).reshape((3,) + mag.shape)
    amp = np.broadcast_to(noise_amplitude(mag), mag.shape)
    return np.where(mag.geom_mask()[None], amp[None] * z, 0.0)
