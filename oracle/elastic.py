"""Elastodynamics and magnetoelastic coupling in FP64.

Test infrastructure only: imported by the tests, the smoke check and the
benchmark's CPU baseline, never by the product.

Restates, for a magnet `mag` of oracle.model (its `neighbor(d)` gives the
in-geometry mask and index of the cell at offset d after the mastergrid
wrap):

* boundary-adapted first derivatives: src/physics/straintensor.cu:45-79
  (the same stencil chain is used in magnetoelasticforce.cu:42-76);
* strain tensor [xx, yy, zz, xy, xz, yz]: straintensor.cu:81-97;
* stress tensor: stresstensor.cu:9-40 (twice C44 on the real shear strain);
* elastic force, divergence of stress with traction-free boundaries:
  elasticforce.cu:25-101;
* magnetoelastic field from the dynamic strain (magnetoelasticfield.cu:44-70)
  or from a rigid applied strain (:72-103);
* magnetoelastic force: magnetoelasticforce.cu:10-94;
* elastic damping -eta v: elasticdamping.cu:13-28;
* effective body force and acceleration: elastodynamics.cu:23-84.

Fields are arrays (ncomp, nz, ny, nx); parameters are scalars or (nz, ny,
nx) arrays (vector parameters (3, nz, ny, nx)).

Pinned by tests/test_oracle_elastic.py against the closed forms of the
reference's tests (sine displacements and magnetisations on periodic grids:
test/test_strain.py, test_elasticforce_uniform_c.py,
test_magnetoelasticfield.py, test_magnetoelasticforce.py), the fourth-order
convergence of the interior stencil, the traction-free boundary values of a
uniformly strained bar, and rigid-equals-dynamic strain; the time
integration of (u, v) by a standing elastic wave (tests/test_oracle_afm.py).
"""

import numpy as np


def _shift(mag, axis, k):
    d = [0, 0, 0]
    d[axis] = k
    return mag.neighbor(d)


def derivative(mag, f, axis):
    """d f / d axis with the stencil chain of straintensor.cu:45-79."""
    im2, jm2 = _shift(mag, axis, -2)
    im1, jm1 = _shift(mag, axis, -1)
    ip1, jp1 = _shift(mag, axis, 1)
    ip2, jp2 = _shift(mag, axis, 2)

    def at(j):
        return f[(slice(None),) + j]

    f0 = f
    choices = [
        np.zeros_like(f0),                                     # --1--
        0.5 * (at(jp1) - at(jm1)),                             # -111-, 1111-, -1111
        f0 - at(jm1),                                          # -11--
        -f0 + at(jp1),                                         # --11-
        0.5 * at(jm2) - 2.0 * at(jm1) + 1.5 * f0,              # 111--
        -0.5 * at(jp2) + 2.0 * at(jp1) - 1.5 * f0,             # --111
    ]
    conds = [
        ~im1 & ~ip1,
        (~im2 | ~ip2) & im1 & ip1,
        ~im2 & ~ip1,
        ~im1 & ~ip2,
        im2 & ~ip1,
        ~im1 & ip1,
    ]
    central4 = (2.0 / 3.0) * (at(jp1) - at(jm1)) + (1.0 / 12.0) * (at(jm2) - at(jp2))
    out = np.empty_like(f0)
    for c in range(f0.shape[0]):
        out[c] = np.select(conds, [ch[c] for ch in choices], default=central4[c])
    return out / mag.cellsize[axis]


def _masked(mag, f):
    return np.where(mag.geom_mask()[None], f, 0.0)


def strain_tensor(mag, u):
    """straintensor.cu:13-97: components xx, yy, zz, xy, xz, yz."""
    der = [derivative(mag, u, i) for i in range(3)]  # der[i][j] = d_i u_j
    e = np.zeros((6,) + mag.shape)
    for i in range(3):
        for j in range(i, 3):
            if i == j:
                e[i] = der[i][j]
            else:
                e[i + j + 2] = 0.5 * (der[i][j] + der[j][i])
    return _masked(mag, e)


def stress_tensor(mag, strain):
    """stresstensor.cu:9-40."""
    C11, C12, C44 = mag.val("C11"), mag.val("C12"), mag.val("C44")
    s = np.zeros((6,) + mag.shape)
    for i in range(3):
        ip1, ip2 = (i + 1) % 3, (i + 2) % 3
        s[i] = C11 * strain[i] + C12 * strain[ip1] + C12 * strain[ip2]
        s[i + 3] = 2.0 * C44 * strain[i + 3]
    return _masked(mag, s)


def _tensor_comp(row, col):
    return row if row == col else row + col + 2


def elastic_force(mag, stress):
    """elasticforce.cu:25-101: f_j = d_i sigma_ij, traction-free boundaries."""
    f = np.zeros((3,) + mag.shape)
    for i in range(3):
        row = stress[[_tensor_comp(i, 0), _tensor_comp(i, 1), _tensor_comp(i, 2)]]
        im2, jm2 = _shift(mag, i, -2)
        im1, jm1 = _shift(mag, i, -1)
        ip1, jp1 = _shift(mag, i, 1)
        ip2, jp2 = _shift(mag, i, 2)

        def at(j):
            return row[(slice(None),) + j]

        w = 1.0 / mag.cellsize[i]
        choices = [
            np.zeros_like(row),
            w * (row + at(jp1) / 3.0),
            w * (-at(jm1) / 3.0 - row),
            0.5 * w * (at(jp1) - at(jm1)),
        ]
        conds = [~im1 & ~ip1, ~im1, ~ip1, ~im2 | ~ip2]
        default = w * ((4.0 / 6.0) * (at(jp1) - at(jm1)) + (1.0 / 12.0) * (at(jm2) - at(jp2)))
        for c in range(3):
            f[c] += np.select(conds, [ch[c] for ch in choices], default=default[c])
    return _masked(mag, f)


def magnetoelastic_field_dynamic(mag, m, strain):
    """magnetoelasticfield.cu:44-70."""
    B1, B2, ms = mag.val("B1"), mag.val("B2"), mag.val("msat")
    h = np.zeros_like(m)
    for i in range(3):
        ip1, ip2 = (i + 1) % 3, (i + 2) % 3
        h[i] = -2.0 / ms * (B1 * strain[i] * m[i] + B2 * (strain[i + ip1 + 2] * m[ip1] + strain[i + ip2 + 2] * m[ip2]))
    return _masked(mag, h)


def magnetoelastic_field_rigid(mag, m, n, s):
    """magnetoelasticfield.cu:72-103: n = rigid normal strain (xx, yy, zz),
    s = rigid shear strain (xy, xz, yz), both (3, nz, ny, nx)."""
    B1, B2, ms = mag.val("B1"), mag.val("B2"), mag.val("msat")
    h = np.zeros_like(m)
    for i in range(3):
        ip1, ip2 = (i + 1) % 3, (i + 2) % 3
        h[i] = -2.0 / ms * (B1 * n[i] * m[i] + B2 * (s[i + ip1 - 1] * m[ip1] + s[i + ip2 - 1] * m[ip2]))
    return _masked(mag, h)


def magnetoelastic_force(mag, m):
    """magnetoelasticforce.cu:10-94."""
    B1, B2 = mag.val("B1"), mag.val("B2")
    der = [derivative(mag, m, i) for i in range(3)]  # der[i][j] = d_i m_j
    f = np.zeros_like(m)
    for i in range(3):
        ip1, ip2 = (i + 1) % 3, (i + 2) % 3
        f[i] = 2.0 * B1 * m[i] * der[i][i]
        f[i] += B2 * m[i] * (der[ip1][ip1] + der[ip2][ip2])
        f[i] += B2 * (m[ip1] * der[ip1][i] + m[ip2] * der[ip2][i])
    return _masked(mag, f)


def elastic_damping(mag, v):
    """elasticdamping.cu:13-28."""
    return _masked(mag, -mag.val("eta") * v)
