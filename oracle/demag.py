"""Demag (self stray field) convolution in float64 (test infrastructure only).

Restates StrayFieldFFTExecutor::exec (src/physics/strayfieldfft.cu:191-246):
M = Ms * m inside the geometry (k_pad :51-73), H = -mu0 * (N * M) as a linear
convolution over the grid, zero outside the geometry (k_unpad :75-100).  The
convolution is evaluated with numpy's float64 FFT on the reference's own
kernel grid sizes, with the kernel placed circularly (offset r at index
r mod K).  For the cropped output this equals the reference's pre-shifted
kernel grid followed by taking the last N cells (:93), because every offset
that reaches a kept output lies in -(N-1)..(N-1) and K >= 2N-1.

``demag_brute`` restates k_demagfield (src/physics/strayfieldbrute.cu:15-57),
an O(N^2) direct sum, used to pin the FFT path on tiny grids.

mu0 is the float64 constant of src/physics/constants.hpp:22.

Pinned by (tests/test_oracle_demag.py): FFT == direct sum, the reference
test's own numpy formulation (test/test_demag.py:6-45), Aharoni's closed form
for the volume-averaged field of a uniformly magnetised prism, and the thin
film limit H_z -> -mu0 Ms.
"""

import numpy as np
import scipy.fft as _sfft

from . import newell

MU0 = 1.25663706212e-6
_WORKERS = 1


def set_fft_workers(n):
    """Threads used by the FFTs (scipy.fft workers)."""
    global _WORKERS
    _WORKERS = max(1, int(n))


def padded_size(n):
    """Kernel grid size of the reference for a self-interaction (kernelGrid)."""
    s = 2 * n - 1
    if s > 5 and s % 2 == 1:
        s += 1
    return s


def kernel_offsets(size, cs, mastergrid=(0, 0, 0), pbc=(0, 0, 0)):
    """Kernel on offsets -(N-1)..(N-1) per axis: array (6, 2Nz-1, 2Ny-1, 2Nx-1)."""
    nx, ny, nz = size
    return newell.kernel_on_axes(np.arange(-(nx - 1), nx), np.arange(-(ny - 1), ny),
                                 np.arange(-(nz - 1), nz), cs, mastergrid, pbc)


# parity of each component (xx, yy, zz, xy, xz, yz) along x, y, z: True = odd
ODD = ((False, False, False), (False, False, False), (False, False, False),
       (True, True, False), (True, False, True), (False, True, True))


def kernel_offsets_octant(size, cs, mastergrid=(0, 0, 0), pbc=(0, 0, 0)):
    """Same array as kernel_offsets, built from the offsets 0..N-1 and the
    even / odd symmetry of each component (for large grids)."""
    nx, ny, nz = size
    oct_ = newell.kernel_on_axes(np.arange(nx), np.arange(ny), np.arange(nz), cs, mastergrid, pbc)
    out = np.empty((6, 2 * nz - 1, 2 * ny - 1, 2 * nx - 1))
    ax = [np.abs(np.arange(-(n - 1), n)) for n in (nx, ny, nz)]
    sg = [np.sign(np.arange(-(n - 1), n)).astype(float) for n in (nx, ny, nz)]
    for c in range(6):
        v = oct_[c][np.ix_(ax[2], ax[1], ax[0])]
        for a, odd in enumerate(ODD[c]):
            if odd:
                shape = [1, 1, 1]
                shape[2 - a] = -1
                v = v * sg[a].reshape(shape)
        out[c] = v
    return out


def kernel_spectrum(size, cs, mastergrid=(0, 0, 0), pbc=(0, 0, 0), K=None, octant=None):
    """Spectrum (6, Kz, Ky, Kx//2+1) of the circularly placed kernel."""
    nx, ny, nz = size
    if K is None:
        K = (padded_size(nx), padded_size(ny), padded_size(nz))
    kx, ky, kz = K
    if octant is None:
        octant = nx * ny * nz > 4096
    koff = (kernel_offsets_octant if octant else kernel_offsets)(size, cs, mastergrid, pbc)
    full = np.zeros((6, kz, ky, kx))
    iz = np.arange(-(nz - 1), nz) % kz
    iy = np.arange(-(ny - 1), ny) % ky
    ix = np.arange(-(nx - 1), nx) % kx
    full[np.ix_(range(6), iz, iy, ix)] = koff
    del koff
    spec = _sfft.rfftn(full, axes=(1, 2, 3), workers=_WORKERS)
    return spec, K


class DemagPlan:
    """Precomputed kernel spectrum for one magnet grid."""

    def __init__(self, size, cs, mastergrid=(0, 0, 0), pbc=(0, 0, 0)):
        self.size = tuple(size)
        self.spec, self.K = kernel_spectrum(size, cs, mastergrid, pbc)

    def field(self, m, msat, geom=None):
        """Demag field in tesla for magnetisation m (3, nz, ny, nx)."""
        nx, ny, nz = self.size
        kx, ky, kz = self.K
        M = np.asarray(msat) * m
        if geom is not None:
            M = np.where(geom[None], M, 0.0)
        Mp = np.zeros((3, kz, ky, kx))
        Mp[:, :nz, :ny, :nx] = M
        Mf = _sfft.rfftn(Mp, axes=(1, 2, 3), workers=_WORKERS)
        xx, yy, zz, xy, xz, yz = self.spec
        Hf = np.empty_like(Mf)
        Hf[0] = xx * Mf[0] + xy * Mf[1] + xz * Mf[2]
        Hf[1] = xy * Mf[0] + yy * Mf[1] + yz * Mf[2]
        Hf[2] = xz * Mf[0] + yz * Mf[1] + zz * Mf[2]
        H = _sfft.irfftn(Hf, s=(kz, ky, kx), axes=(1, 2, 3), workers=_WORKERS)[:, :nz, :ny, :nx]
        H = -MU0 * H
        if geom is not None:
            H = np.where(geom[None], H, 0.0)
        return H


def demag_fft(m, msat, cs, geom=None, mastergrid=(0, 0, 0), pbc=(0, 0, 0)):
    nz, ny, nx = m.shape[1:]
    return DemagPlan((nx, ny, nz), cs, mastergrid, pbc).field(m, msat, geom)


def demag_brute(m, msat, cs, geom=None, mastergrid=(0, 0, 0), pbc=(0, 0, 0)):
    """Direct sum over source cells (strayfieldbrute.cu:15-57), tiny grids only."""
    nz, ny, nx = m.shape[1:]
    koff = kernel_offsets((nx, ny, nz), cs, mastergrid, pbc)
    M = np.asarray(msat) * m
    if geom is None:
        geom = np.ones((nz, ny, nx), dtype=bool)
    H = np.zeros((3, nz, ny, nx))
    src = np.argwhere(geom)
    for (dz, dy, dx) in np.argwhere(geom):
        h = np.zeros(3)
        for (sz, sy, sx) in src:
            rz, ry, rx = dz - sz + nz - 1, dy - sy + ny - 1, dx - sx + nx - 1
            nxx, nyy, nzz, nxy, nxz, nyz = koff[:, rz, ry, rx]
            Mv = M[:, sz, sy, sx]
            h[0] -= nxx * Mv[0] + nxy * Mv[1] + nxz * Mv[2]
            h[1] -= nxy * Mv[0] + nyy * Mv[1] + nyz * Mv[2]
            h[2] -= nxz * Mv[0] + nyz * Mv[1] + nzz * Mv[2]
        H[:, dz, dy, dx] = MU0 * h
    return H


def demag_reference_layout(m, msat, cs, mastergrid=(0, 0, 0), pbc=(0, 0, 0)):
    """The reference test's own formulation (test/test_demag.py:6-45).

    Pads M at the front of the reference kernel grid, multiplies full complex
    spectra and keeps the last N cells.  Used to pin ``DemagPlan``.
    """
    nz, ny, nx = m.shape[1:]
    kernel = newell.reference_kernel((nx, ny, nz), cs, mastergrid, pbc)
    mag = np.asarray(msat) * m
    pad = ((0, kernel.shape[1] - nz), (0, kernel.shape[2] - ny), (0, kernel.shape[3] - nx))
    mp = np.pad(mag, ((0, 0), *pad), "constant")
    mf = np.fft.fftn(mp, axes=(1, 2, 3))
    kxx, kyy, kzz, kxy, kxz, kyz = np.fft.fftn(kernel, axes=(1, 2, 3))
    hx = np.fft.ifftn(mf[0] * kxx + mf[1] * kxy + mf[2] * kxz)
    hy = np.fft.ifftn(mf[0] * kxy + mf[1] * kyy + mf[2] * kyz)
    hz = np.fft.ifftn(mf[0] * kxz + mf[1] * kyz + mf[2] * kzz)
    h = -MU0 * np.array([hx, hy, hz]).real
    return h[:, h.shape[1] - nz:, h.shape[2] - ny:, h.shape[3] - nx:]
