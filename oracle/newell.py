"""Newell demagnetisation tensor in float64 (test infrastructure only).

Restates src/physics/newell.cu:7-139 (Newell, Williams & Dunlop, JGR 98
(1993) 9551, eqs. 27 and 32) and the kernel grid / PBC image sum of
src/physics/strayfieldkernel.cu:25-99.

The indefinite integrals are evaluated once per node of the integer lattice
and the 27-point weighted sum is taken over shifted views of that node array.
F(idx + dr) is a pure function of the node, so this gives the same float64
terms as evaluating them inside the 27-point loop.  The weights and the
summation order (dx outer, dz inner) follow newell.cu:64-96.

One deliberate difference: the reference computes the lattice coordinates as
``int * float`` (FP32 build) before widening to double; here the cell size is
kept in float64 throughout.

Pinned by (tests/test_oracle_newell.py): the trace identity
Nxx+Nyy+Nzz = 1 of the self term, N = 1/3 for a cube, Aharoni's closed-form
prism demag factor (J. Appl. Phys. 83, 3432 (1998)), the point-dipole far
field, the parity symmetries and mpmath 50-digit spot checks.
"""

import numpy as np

COMPONENTS = ("xx", "yy", "zz", "xy", "xz", "yz")


def f_xx(ix, iy, iz, cs):
    """Indefinite integral for Nxx (newell.cu:7-28), vectorised over int arrays."""
    ix, iy, iz = np.broadcast_arrays(*(np.asarray(a, dtype=np.int64) for a in (ix, iy, iz)))
    x = ix * float(cs[0])
    y = iy * float(cs[1])
    z = iz * float(cs[2])
    R = np.sqrt(x * x + y * y + z * z)
    result = np.zeros(x.shape)
    with np.errstate(divide="ignore", invalid="ignore"):
        m1 = (ix != 0) | (iz != 0)
        t1 = (y / 2) * (z * z - x * x) * np.arcsinh(y / np.sqrt(x * x + z * z))
        result = result + np.where(m1, t1, 0.0)
        m2 = (ix != 0) | (iy != 0)
        t2 = (z / 2) * (y * y - x * x) * np.arcsinh(z / np.sqrt(x * x + y * y))
        result = result + np.where(m2, t2, 0.0)
        m3 = ix != 0
        t3 = x * y * z * np.arctan((y * z) / (x * R))
        result = result - np.where(m3, t3, 0.0)
    result = result + (1.0 / 6.0) * (2 * x * x - y * y - z * z) * R
    return result


def f_xy(ix, iy, iz, cs):
    """Indefinite integral for Nxy (newell.cu:34-58), vectorised over int arrays."""
    ix, iy, iz = np.broadcast_arrays(*(np.asarray(a, dtype=np.int64) for a in (ix, iy, iz)))
    x = ix * float(cs[0])
    y = iy * float(cs[1])
    z = iz * float(cs[2])
    R = np.sqrt(x * x + y * y + z * z)
    zero = (ix == 0) | (iy == 0)
    with np.errstate(divide="ignore", invalid="ignore"):
        result = (x * y * z) * np.arcsinh(z / np.sqrt(x * x + y * y))
        result = result + (y / 6) * (3 * z * z - y * y) * np.arcsinh(x / np.sqrt(y * y + z * z))
        result = result + (x / 6) * (3 * z * z - x * x) * np.arcsinh(y / np.sqrt(x * x + z * z))
        result = result - (z * y * y / 2) * np.arctan((x * z) / (y * R))
        result = result - (z * x * x / 2) * np.arctan((y * z) / (x * R))
        result = result - x * y * R / 3
        t = (z * z * z / 6) * np.arctan((x * y) / (z * R))
        result = result - np.where(iz != 0, t, 0.0)
    return np.where(zero, 0.0, result)


def _weight(dx, dy, dz):
    """newellWeight (newell.cu:64-77)."""
    return {0: 8, 1: -4, 2: 2, 3: -1}[dx * dx + dy * dy + dz * dz]


def _stencil_sum(F, ix, iy, iz, cs):
    """27-point weighted sum of an indefinite integral on the lattice.

    ``ix, iy, iz`` are 1-D integer coordinate axes; the result has shape
    (len(iz), len(iy), len(ix)).  F is evaluated once on the node lattice that
    covers every idx + dr.
    """
    ix = np.asarray(ix, dtype=np.int64)
    iy = np.asarray(iy, dtype=np.int64)
    iz = np.asarray(iz, dtype=np.int64)
    # node lattice per axis: the union of {i-1, i, i+1}
    nx = np.unique(np.concatenate([ix - 1, ix, ix + 1]))
    ny = np.unique(np.concatenate([iy - 1, iy, iy + 1]))
    nz = np.unique(np.concatenate([iz - 1, iz, iz + 1]))
    Z, Y, X = np.meshgrid(nz, ny, nx, indexing="ij")
    Fn = F(X, Y, Z, cs)
    px = {v: k for k, v in enumerate(nx)}
    py = {v: k for k, v in enumerate(ny)}
    pz = {v: k for k, v in enumerate(nz)}
    result = np.zeros((len(iz), len(iy), len(ix)))
    for dx in (-1, 0, 1):
        sx = np.array([px[v + dx] for v in ix])
        for dy in (-1, 0, 1):
            sy = np.array([py[v + dy] for v in iy])
            for dz in (-1, 0, 1):
                sz = np.array([pz[v + dz] for v in iz])
                w = _weight(dx, dy, dz)
                result = result + w * Fn[np.ix_(sz, sy, sx)]
    result = result / (4 * np.pi * float(cs[0]) * float(cs[1]) * float(cs[2]))
    return result


def _permute(comp, ix, iy, iz, cs):
    """Argument permutations that derive the other components (newell.cu:117-139).

    Returns (F, axes-in-F-order, cellsize-in-F-order, transpose) such that
    component(z, y, x) = stencil(F, a0, a1, a2, c).transpose(...).
    """
    cx, cy, cz = cs
    if comp == "xx":
        return f_xx, (ix, iy, iz), (cx, cy, cz)
    if comp == "yy":  # Nxx({y, z, x}, {cy, cz, cx})
        return f_xx, (iy, iz, ix), (cy, cz, cx)
    if comp == "zz":  # Nxx({z, x, y}, {cz, cx, cy})
        return f_xx, (iz, ix, iy), (cz, cx, cy)
    if comp == "xy":
        return f_xy, (ix, iy, iz), (cx, cy, cz)
    if comp == "xz":  # Nxy({x, z, y}, {cx, cz, cy})
        return f_xy, (ix, iz, iy), (cx, cz, cy)
    if comp == "yz":  # Nxy({y, z, x}, {cy, cz, cx})
        return f_xy, (iy, iz, ix), (cy, cz, cx)
    raise ValueError(comp)


def component(comp, ix, iy, iz, cs):
    """One Newell component on the lattice ix (x) iy (x) iz, shape (nz, ny, nx)."""
    F, (a0, a1, a2), c = _permute(comp, ix, iy, iz, cs)
    # stencil result is indexed [a2][a1][a0]; map back to [z][y][x]
    r = _stencil_sum(F, a0, a1, a2, c)
    if F is f_xy:
        # calcNewellNxy returns exactly 0 when the first or second (permuted)
        # offset is 0 (newell.cu:99-100)
        zero = (np.asarray(a1)[None, :, None] == 0) | (np.asarray(a0)[None, None, :] == 0)
        r = np.where(zero, 0.0, r)
    order = {"xx": "zyx", "yy": "xzy", "zz": "yxz", "xy": "zyx", "xz": "yzx", "yz": "xzy"}[comp]
    # r's axes are named by `order` (outermost first); bring them to z, y, x
    perm = [order.index(a) for a in "zyx"]
    return np.transpose(r, perm)


def component_point(comp, i, j, k, cs):
    """Scalar helper: component at one lattice offset."""
    return float(component(comp, [i], [j], [k], cs)[0, 0, 0])


def kernel_on_axes(ix, iy, iz, cs, mastergrid=(0, 0, 0), pbc=(0, 0, 0)):
    """All six components on the lattice axes, with the PBC image sum.

    Restates k_strayFieldKernel (strayfieldkernel.cu:25-56): the value at
    offset r is the sum over images r + (i, j, k) * mastergrid for
    |i| <= pbc.x etc.  Returns an array (6, nz, ny, nx) in COMPONENTS order.
    """
    ix = np.asarray(ix, dtype=np.int64)
    iy = np.asarray(iy, dtype=np.int64)
    iz = np.asarray(iz, dtype=np.int64)
    out = np.zeros((6, len(iz), len(iy), len(ix)))
    for i in range(-pbc[0], pbc[0] + 1):
        for j in range(-pbc[1], pbc[1] + 1):
            for k in range(-pbc[2], pbc[2] + 1):
                sx = ix + i * mastergrid[0]
                sy = iy + j * mastergrid[1]
                sz = iz + k * mastergrid[2]
                for c, name in enumerate(COMPONENTS):
                    out[c] += component(name, sx, sy, sz, cs)
    return out


def reference_kernel_grid(dst_size, dst_origin, src_size, src_origin):
    """kernelGrid (strayfieldkernel.cu:80-99): (size, origin) of the kernel grid."""
    size = [s + d - 1 for s, d in zip(src_size, dst_size)]
    origin = [do - (so + ss - 1) for do, so, ss in zip(dst_origin, src_origin, src_size)]
    for a in range(3):
        if size[a] > 5 and size[a] % 2 == 1:
            size[a] += 1
            origin[a] -= 1
    return tuple(size), tuple(origin)


def reference_kernel(size, cs, mastergrid=(0, 0, 0), pbc=(0, 0, 0), origin=(0, 0, 0)):
    """The kernel exactly as the reference lays it out for a magnet of `size`.

    Shape (6, Kz, Ky, Kx), value at grid index j is N(origin_k + j); this is
    what ``_demag_kernel(magnet)`` returns (wrap_magnet.cpp:56-59).
    """
    ksize, korigin = reference_kernel_grid(size, origin, size, origin)
    axes = [np.arange(korigin[a], korigin[a] + ksize[a]) for a in range(3)]
    return kernel_on_axes(axes[0], axes[1], axes[2], cs, mastergrid, pbc)
