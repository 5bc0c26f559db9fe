"""Electrical potential of a ferromagnet (test infrastructure only): FP64
restatement of the row construction of PoissonSystem
(src/physics/poissonsystem.cu:134-258), the AMR conductivity tensor
(src/physics/conductivitytensor.cu:11-32) and the four steppers of LinSolver
(src/linsolver/linsolver.cpp: Jacobi, conjugate gradient, minimal residual,
steepest descent), plus the direct sparse solution every converging method
approaches.
"""

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def conductivity_tensor(mag):
    """(6, nz, ny, nx): xx yy zz = sigma0 (1 - f (m_i^2 - 1/3)), xy xz yz = sigma0 f m_i m_j, f = 6 r / (6 + r)."""
    s0 = mag.val("conductivity")
    r = mag.val("amr_ratio")
    f = 6.0 * r / (6.0 + r)
    m = mag.m
    out = [s0 * (1.0 - f * (m[c] * m[c] - 1.0 / 3.0)) for c in range(3)]
    out += [s0 * f * m[a] * m[b] for a, b in ((0, 1), (0, 2), (1, 2))]
    return np.where(mag.geom_mask()[None], np.array(out), 0.0)


def rows(mag, normalize=True):
    """(A, b) of poissonsystem.cu:134-258 as a sparse matrix; rows divided by the diagonal when normalize."""
    nz, ny, nx = mag.shape
    N = nx * ny * nz
    geom = mag.geom_mask()
    aniso = not mag.assured_zero("amr_ratio")
    sig = conductivity_tensor(mag) if aniso else mag.val("conductivity")[None] * np.ones((1,) + mag.shape)
    pot = mag.val("applied_potential")
    cs = mag.cellsize

    def inside(x, y, z):
        return 0 <= x < nx and 0 <= y < ny and 0 <= z < nz and geom[z, y, x]

    def lin(x, y, z):
        return (z * ny + y) * nx + x

    A = sp.lil_matrix((N, N))
    b = np.zeros(N)
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                i = lin(x, y, z)
                if not inside(x, y, z) or sig[0, z, y, x] == 0 or not np.isnan(pot[z, y, x]):
                    A[i, i] = 1.0
                    b[i] = 0.0 if np.isnan(pot[z, y, x]) else pot[z, y, x]
                    continue
                row = {}

                def diff(p, q, w):
                    if inside(x + p[0], y + p[1], z + p[2]) and inside(x + q[0], y + q[1], z + q[2]):
                        row[p] = row.get(p, 0.0) + w
                        row[q] = row.get(q, 0.0) - w

                for ax in range(3):
                    for s in (-1, 1):
                        e = [0, 0, 0]
                        e[ax] = s
                        e = tuple(e)
                        if not inside(x + e[0], y + e[1], z + e[2]):
                            continue
                        area = cs[(ax + 1) % 3] * cs[(ax + 2) % 3]
                        n = (z + e[2], y + e[1], x + e[0])
                        c = ax if aniso else 0
                        diff((0, 0, 0), e, area * np.sqrt(sig[c, z, y, x] * sig[(c,) + n]) / cs[ax])
                        if not aniso:
                            continue
                        for ot in range(3):
                            if ot == ax:
                                continue
                            comp = {1: 3, 2: 4, 3: 5}[ax + ot]
                            w = -s * area * np.sqrt(sig[comp, z, y, x] * sig[(comp,) + n]) / cs[ot] / 4
                            o = [0, 0, 0]
                            o[ot] = 1
                            o = tuple(o)
                            mo = tuple(-v for v in o)
                            diff((0, 0, 0), mo, w)
                            diff(e, tuple(e[k] + mo[k] for k in range(3)), w)
                            diff(o, (0, 0, 0), w)
                            diff(tuple(e[k] + o[k] for k in range(3)), e, w)
                d = row[(0, 0, 0)] if normalize else 1.0
                for (dx, dy, dz), v in row.items():
                    if v != 0:
                        A[i, lin(x + dx, y + dy, z + dz)] = v / d
    return A.tocsr(), b


def potential(mag):
    """The exact solution of the (normalised) system, 0 outside the geometry."""
    A, b = rows(mag)
    x = spla.spsolve(A.tocsc(), b)
    return np.where(mag.geom_mask(), x.reshape(mag.shape), 0.0)


class Stepper:
    """LinSolver with x = 0 and one of its four steppers (linsolver.cpp)."""

    def __init__(self, A, b, method="jacobi"):
        self.A, self.b, self.method = A, b, method
        self.x = np.zeros(len(b))
        self.restart()

    def residual(self):
        return self.b - self.A @ self.x

    def restart(self):
        if self.method == "jacobi":
            return
        self.r = self.residual()
        self.p = self.r.copy() if self.method == "conjugategradient" else self.A @ self.r
        self.rr = self.r @ self.r

    def step(self):
        if self.method == "jacobi":
            self.x = self.x + self.residual()
        elif self.method == "conjugategradient":
            Ap = self.A @ self.p
            alpha = self.rr / (self.p @ Ap)
            self.x = self.x + alpha * self.p
            self.r = self.r - alpha * Ap
            rr_old, self.rr = self.rr, self.r @ self.r
            self.p = self.r + (self.rr / rr_old) * self.p
        else:
            den = self.p @ self.p if self.method == "minimalresidual" else self.p @ self.r
            alpha = self.rr / den
            self.x = self.x + alpha * self.r
            self.r = self.r - alpha * self.p
            self.p = self.A @ self.r
            self.rr = self.r @ self.r
