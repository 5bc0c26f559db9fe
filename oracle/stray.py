"""Stray fields between grids in FP64 (test infrastructure only).

Restates StrayFieldBruteExecutor (src/physics/strayfieldbrute.cu:15-88) for
a source grid and a destination grid: h(d) = -mu0 sum_s N(d - s) . M(s),
with N the Newell tensor at the cell offset including the periodic images
(strayfieldkernel.cu:25-56, here oracle.newell.kernel_on_axes), and the
external-field rule that every other source-enabled magnet with msat not
assured zero acts on a destination-enabled magnet (zeeman.cu:13-34, 70-88).

Pinned by tests/test_oracle_stray.py: equal to the oracle's own demag field
when source and destination are the same grid (the demag brute sum and FFT
pinned in test_oracle_demag.py), and to the demag field of the enclosing box
with the magnetisation set only inside the source (the identity of
test/test_strayfield.py:86-136); far-field dipole limit.
"""

import numpy as np

from . import newell

MU0 = 1.25663706212e-6


def stray_field(M, src_origin, dst_size, dst_origin, cs, mastergrid=(0, 0, 0), pbc=(0, 0, 0)):
    """M: (3, nz, ny, nx) source magnetisation Ms m (zero outside its
    geometry); returns (3, dz, dy, dx) on the destination grid."""
    snz, sny, snx = M.shape[1:]
    sn = (snx, sny, snz)
    dn = tuple(int(v) for v in dst_size)
    ko = [dst_origin[a] - (src_origin[a] + sn[a] - 1) for a in range(3)]
    ks = [dn[a] + sn[a] - 1 for a in range(3)]
    axes = [np.arange(ko[a], ko[a] + ks[a]) for a in range(3)]
    K = newell.kernel_on_axes(axes[0], axes[1], axes[2], cs, mastergrid, pbc)  # (6, kz, ky, kx)
    H = np.zeros((3, dn[2], dn[1], dn[0]))
    for sz, sy, sx in np.argwhere(np.any(M != 0, axis=0)):
        # offsets d - s for all destination cells d: index (d - s) - ko
        iz = dst_origin[2] + np.arange(dn[2]) - (src_origin[2] + sz) - ko[2]
        iy = dst_origin[1] + np.arange(dn[1]) - (src_origin[1] + sy) - ko[1]
        ix = dst_origin[0] + np.arange(dn[0]) - (src_origin[0] + sx) - ko[0]
        n = K[:, iz[:, None, None], iy[None, :, None], ix[None, None, :]]
        mv = M[:, sz, sy, sx]
        H[0] -= n[0] * mv[0] + n[3] * mv[1] + n[4] * mv[2]
        H[1] -= n[3] * mv[0] + n[1] * mv[1] + n[5] * mv[2]
        H[2] -= n[4] * mv[0] + n[5] * mv[1] + n[2] * mv[2]
    return MU0 * H
