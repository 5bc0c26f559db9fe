"""Local effective-field terms, torque and energies in float64 (test infrastructure only).

Every function takes an ``oracle.model.Ferromagnet`` (``mag``) and a
magnetisation array ``m`` of shape (3, nz, ny, nx) and returns arrays of the
same layout.  Cells outside the geometry get zero, as the reference kernels do.

Restated functions (reference file:line):

* ``exchange_field``    k_exchangeField FM variant, src/physics/exchange.cu:20-104
* ``dmi_field``         k_dmiFieldFM, src/physics/dmi.cu:87-178; getGamma dmi.hpp:53-69
* ``anisotropy_field``  k_unianisotropyField / k_cubicanisotropyField and the
                        uniaxial-else-cubic selection, src/physics/anisotropy.cu:25-140
* ``external_field``    evalExternalField without other magnets' stray fields,
                        src/physics/zeeman.cu:53-91
* ``demag_field``       evalDemagField, src/physics/demag.cpp:13-19
* ``effective_field``   evalEffectiveField, src/physics/effectivefield.cpp:15-28
* ``llg_torque``        k_llgtorque, src/physics/torque.cu:21-45
* ``damping_torque``    k_dampingtorque, src/physics/torque.cu:58-77
* energy densities      k_energyDensity (src/physics/energy.cu:18-50) and
                        k_unianisotropyEnergyDensity (anisotropy.cu:142-180)
* ``zhang_li_torque``, ``slonczewski_torque``, ``spin_transfer_torque``
                        k_ZhangLi, k_Slonczewski, evalSpinTransferTorque,
                        src/physics/stt.cu:9-188 (pinned by tests/test_oracle_stt.py:
                        the central-difference closed form on a spiral, zero torque
                        for uniform m, torque perpendicular to m, the empty-layer
                        test of test/test_free_layer_thickness.py, the sign flip of
                        the fixed-layer position; the GPU test of the domain-wall
                        drift -u t checks the product's signs physically)

Constants: GAMMALL = 1.7595e11 rad/(T s) (src/physics/constants.hpp:20).

Pinned by (tests/test_oracle_fields.py): the reference tests' own numpy
formulations (test/test_exchange.py:6-46, test/test_interfacialdmi.py:8-56),
closed-form anisotropy fields, the exchange spiral energy A k^2
(test/test_exchange.py:48-85), the 1-D DMI edge canting golden values of
test/test_boundaries.py:17-20 (Neumann, open, periodic) reached by damped
relaxation, single-spin precession (test/test_timesolver.py:19-29) and the
identity m . tau_LLG = 0.
"""

import numpy as np

from . import demag as _demag

GAMMALL = 1.7595e11
MU0 = _demag.MU0

DIRS = ((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1))


def harmonic_mean(a, b):
    """harmonicMean (src/core/datatypes.hpp:199-205), elementwise."""
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    s = a + b
    with np.errstate(divide="ignore", invalid="ignore"):
        hm = 2 * a * b / s
    return np.where(s == 0, 0.0, np.where(a == b, a, hm))


def cross(a, b):
    return np.stack([a[1] * b[2] - a[2] * b[1],
                     a[2] * b[0] - a[0] * b[2],
                     a[0] * b[1] - a[1] * b[0]])


def dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def get_gamma(D, n, m):
    """getGamma (dmi.hpp:53-69): boundary DMI term for absolute normal n."""
    nx, ny, nz = n
    return np.stack([
        -D["xxy"] * nx * m[1] - D["xxz"] * nx * m[2] - D["yxz"] * ny * m[2]
        - D["zxy"] * nz * m[1] - D["yxy"] * ny * m[1] - D["zxz"] * nz * m[2],
        D["xxy"] * nx * m[0] - D["zyz"] * nz * m[2] + D["yxy"] * ny * m[0]
        - D["xyz"] * nx * m[2] + D["zxy"] * nz * m[0] - D["yyz"] * ny * m[2],
        D["xxz"] * nx * m[0] + D["yyz"] * ny * m[1] + D["xyz"] * nx * m[1]
        + D["yxz"] * ny * m[0] + D["zxz"] * nz * m[0] + D["zyz"] * nz * m[1]])


def _active(mag):
    """Cells in the geometry with msat != 0 (where the stencil terms are non-zero)."""
    return mag.geom_mask() & (mag.val("msat") != 0)


def exchange_field(mag, m):
    nz, ny, nx = mag.shape
    h = np.zeros((3, nz, ny, nx))
    if mag.exchange_assured_zero():
        return h
    ms = mag.val("msat")
    a = mag.val("aex")
    D = mag.dmi_arrays()
    w = [1.0 / (c * c) for c in mag.cellsize]
    for d in DIRS:
        axis = int(np.flatnonzero(d)[0])
        nb_in, nb = mag.neighbor(d)
        m_nb = m[:, nb[0], nb[1], nb[2]]
        a_nb = a[nb]
        ms_nb = ms[nb]
        delta = d[axis] * mag.cellsize[axis]
        n = tuple(abs(v) for v in d)
        gamma = get_gamma(D, n, m)
        with np.errstate(divide="ignore", invalid="ignore"):
            m_neu = m + (gamma / (2 * a)) * delta
        if mag.openbc:
            include = nb_in & (ms_nb != 0)
        else:
            include = nb_in | (a != 0)
        m_ = np.where(nb_in[None], m_nb, m_neu)
        a_ = np.where(nb_in, a_nb, a)
        A = mag.region_link("inter_exchange", "scale_exchange", nb_in, nb, harmonic_mean(a, a_))
        contrib = 2 * A * w[axis] * (m_ - m)
        h += np.where(include[None], contrib, 0.0)
    act = _active(mag)
    with np.errstate(divide="ignore", invalid="ignore"):
        h = np.where(act[None], h / ms, 0.0)
    return h


_DMI_AXIS = {0: ("xxz", "xxy", "xyz"), 1: ("yxz", "yxy", "yyz"), 2: ("zxz", "zxy", "zyz")}


def dmi_field(mag, m):
    nz, ny, nx = mag.shape
    h = np.zeros((3, nz, ny, nx))
    if mag.dmi_assured_zero():
        return h
    ms = mag.val("msat")
    a = mag.val("aex")
    D = mag.dmi_arrays()
    open_eff = np.logical_or(mag.openbc, a == 0)
    self_idx = np.indices((nz, ny, nx))
    for d in DIRS:
        axis = int(np.flatnonzero(d)[0])
        nb_in, nb = mag.neighbor(d)
        nbi = tuple(np.where(nb_in, nb[i], self_idx[i]) for i in range(3))
        skip = open_eff & (~nb_in | (ms[nbi] == 0))
        kxz, kxy, kyz = _DMI_AXIS[axis]
        Dxz = harmonic_mean(D[kxz], D[kxz][nbi])
        Dxy = harmonic_mean(D[kxy], D[kxy][nbi])
        Dyz = harmonic_mean(D[kyz], D[kyz][nbi])
        Dzx, Dyx, Dzy = -Dxz, -Dxy, -Dyz
        delta = d[axis] * mag.cellsize[axis]
        n = tuple(abs(v) for v in d)
        gamma = get_gamma(D, n, m)
        neumann = ~nb_in & ~open_eff
        with np.errstate(divide="ignore", invalid="ignore"):
            m_neu = m + (gamma / (2 * a)) * delta
        m_ = np.where(neumann[None], m_neu, m[:, nbi[0], nbi[1], nbi[2]])
        c = np.stack([(Dxy * m_[1] + Dxz * m_[2]) / delta,
                      (Dyx * m_[0] + Dyz * m_[2]) / delta,
                      (Dzx * m_[0] + Dzy * m_[1]) / delta])
        h += np.where(skip[None], 0.0, c)
    act = _active(mag)
    with np.errstate(divide="ignore", invalid="ignore"):
        h = np.where(act[None], h / ms, 0.0)
    return h


def _normalized(v):
    n = np.sqrt(dot(v, v))
    with np.errstate(divide="ignore", invalid="ignore"):
        return v / n


def anisotropy_field(mag, m):
    nz, ny, nx = mag.shape
    h = np.zeros((3, nz, ny, nx))
    act = _active(mag)
    ms = mag.val("msat")
    if not mag.uniaxial_assured_zero():
        u = _normalized(mag.vec("anisU"))
        k1 = mag.val("ku1")
        k2 = mag.val("ku2")
        mu = dot(m, u)
        with np.errstate(divide="ignore", invalid="ignore"):
            h = (2 * k1 * mu + 4 * k2 * mu * mu * mu) * u / ms
    elif not mag.cubic_assured_zero():
        c1 = _normalized(mag.vec("anisC1"))
        c2 = _normalized(mag.vec("anisC2"))
        c3 = cross(c1, c2)
        kc1, kc2, kc3 = mag.val("kc1"), mag.val("kc2"), mag.val("kc3")
        c1m, c2m, c3m = dot(c1, m), dot(c2, m), dot(c3, m)
        with np.errstate(divide="ignore", invalid="ignore"):
            h = (-2 * kc1 * ((c2m**2 + c3m**2) * (c1m * c1) + (c1m**2 + c3m**2) * (c2m * c2)
                             + (c1m**2 + c2m**2) * (c3m * c3)) / ms
                 - 2 * kc2 * ((c2m**2 * c3m**2) * (c1m * c1) + (c1m**2 * c3m**2) * (c2m * c2)
                              + (c1m**2 * c2m**2) * (c3m * c3)) / ms
                 - 4 * kc3 * ((c2m**4 + c3m**4) * (c1m**3 * c1) + (c1m**4 + c3m**4) * (c2m**3 * c2)
                              + (c1m**4 + c2m**4) * (c3m**3 * c3)) / ms)
    else:
        return h
    return np.where(act[None], h, 0.0)


def external_field(mag, m, world_bias=(0.0, 0.0, 0.0)):
    nz, ny, nx = mag.shape
    geom = mag.geom_mask()
    h = np.zeros((3, nz, ny, nx))
    for c in range(3):
        h[c] = world_bias[c]
    h = h + mag.vec("bias")
    return np.where(geom[None], h, 0.0)


def external_assured_zero(mag, world_bias):
    return tuple(world_bias) == (0.0, 0.0, 0.0) and mag.assured_zero("bias")


def demag_field(mag, m):
    nz, ny, nx = mag.shape
    if not mag.enable_demag or mag.assured_zero("msat"):
        return np.zeros((3, nz, ny, nx))
    return mag.demag_plan().field(m, mag.val("msat"), mag.geom_mask())


def effective_field(mag, m, world_bias=(0.0, 0.0, 0.0)):
    """Sum of the terms in the order of effectivefield.cpp:15-28."""
    h = exchange_field(mag, m)
    h = h + anisotropy_field(mag, m)
    if not external_assured_zero(mag, world_bias):
        h = h + external_field(mag, m, world_bias)
    h = h + dmi_field(mag, m)
    h = h + demag_field(mag, m)
    return h


def llg_torque(mag, m, h):
    alpha = mag.val("alpha")
    mxh = cross(m, h)
    mxmxh = cross(m, mxh)
    t = -GAMMALL / (1 + alpha * alpha) * (mxh + alpha * mxmxh)
    ok = mag.geom_mask() & (mag.val("frozen_spins") == 0)
    return np.where(ok[None], t, 0.0)


def damping_torque(mag, m, h):
    t = -GAMMALL * cross(m, cross(m, h))
    ok = mag.geom_mask() & (mag.val("frozen_spins") == 0)
    return np.where(ok[None], t, 0.0)


def energy_density(mag, m, h, prefactor):
    """k_energyDensity: -prefactor * Ms * m . h inside the geometry."""
    e = -prefactor * mag.val("msat") * dot(m, h)
    return np.where(mag.geom_mask(), e, 0.0)


def anisotropy_energy_density(mag, m):
    """k_unianisotropyEnergyDensity, else k_cubanisotropyEnergyDensity
    (anisotropy.cu:142-221, chosen as in :223-253)."""
    if not mag.uniaxial_assured_zero():
        u = _normalized(mag.vec("anisU"))
        k1, k2 = mag.val("ku1"), mag.val("ku2")
        mu = dot(m, u)
        e = -k1 * mu * mu - k2 * mu ** 4
    elif not mag.cubic_assured_zero():
        c1 = _normalized(mag.vec("anisC1"))
        c2 = _normalized(mag.vec("anisC2"))
        c3 = cross(c1, c2)
        kc1, kc2, kc3 = mag.val("kc1"), mag.val("kc2"), mag.val("kc3")
        a1, a2, a3 = dot(c1, m) ** 2, dot(c2, m) ** 2, dot(c3, m) ** 2
        e = (kc1 * (a1 * a2 + a1 * a3 + a2 * a3) + kc2 * a1 * a2 * a3
             + kc3 * (a1 * a1 * a2 * a2 + a1 * a1 * a3 * a3 + a2 * a2 * a3 * a3))
    else:
        return np.zeros(mag.shape)
    return np.where(_active(mag), e, 0.0)


def total_energy_density(mag, m, world_bias=(0.0, 0.0, 0.0)):
    e = energy_density(mag, m, exchange_field(mag, m), 0.5)
    e = e + anisotropy_energy_density(mag, m)
    if not external_assured_zero(mag, world_bias):
        e = e + energy_density(mag, m, external_field(mag, m, world_bias), 1.0)
    e = e + energy_density(mag, m, dmi_field(mag, m), 0.5)
    e = e + energy_density(mag, m, demag_field(mag, m), 0.5)
    return e


def integrate(mag, density):
    """E = ncells * average * cellvolume, average over cells in the geometry
    (exchange.cu:247-256, reduce.cu:101-127)."""
    geom = mag.geom_mask()
    ncells = geom.size
    avg = density[geom].sum() / max(int(geom.sum()), 1)
    V = mag.cellsize[0] * mag.cellsize[1] * mag.cellsize[2]
    return ncells * avg * V


# ------------------------------------------------------------ spin transfer
QE = 1.60217646e-19  # src/physics/constants.hpp:13
MUB = 9.2740091523e-24
HBAR = 1.054571817e-34


def zhang_li_assured_zero(mag):
    """ZhangLiSTTAssuredZero (src/physics/stt.cu:13-16)."""
    return (not mag.enable_zhangli or mag.assured_zero("msat") or mag.assured_zero("jcur")
            or mag.assured_zero("pol"))


def slonczewski_assured_zero(mag):
    """SlonczewskiSTTAssuredZero (src/physics/stt.cu:18-26)."""
    return (not mag.enable_slonc or mag.assured_zero("msat") or mag.assured_zero("jcur")
            or mag.assured_zero("free_layer_thickness") or mag.assured_zero("fixed_layer")
            or (mag.assured_zero("epsilon_prime") and (mag.assured_zero("Lambda") or mag.assured_zero("pol"))))


def zhang_li_torque(mag, m):
    """k_ZhangLi (src/physics/stt.cu:27-102): u = muB P j / (e Ms (1 + xi^2)),
    h = sum over axes of u_a times the central difference of m over in-grid
    neighbours with non-zero msat (after the mastergrid wrap), then
    t = -1/(1+alpha^2) ((1 + xi alpha) m x (m x h) + (xi - alpha) m x h)."""
    ms, pol, xi, al = mag.val("msat"), mag.val("pol"), mag.val("xi"), mag.val("alpha")
    j = mag.vec("jcur")
    with np.errstate(divide="ignore", invalid="ignore"):
        u = MUB * pol / (QE * ms * (1 + xi * xi)) * j
    u = np.where((ms != 0)[None], u, 0.0)
    h = np.zeros((3,) + mag.shape)
    for d in DIRS:
        axis = 0 if d[0] else (1 if d[1] else 2)
        sign = d[0] + d[1] + d[2]
        in_grid, idx = _in_grid_neighbor(mag, d)
        ok = in_grid & (mag.val("msat")[idx] != 0)
        mn = m[(slice(None),) + idx]
        h += np.where(ok[None], sign * u[axis][None] * mn / (2 * mag.cellsize[axis]), 0.0)
    mxh = cross(m, h)
    mxmxh = cross(m, mxh)
    t = (-1 / (1 + al * al)) * ((1 + xi * al) * mxmxh + (xi - al) * mxh)
    zero = (ms == 0) | (pol == 0) | np.all(j == 0, axis=0)
    ok = mag.geom_mask() & (mag.val("frozen_spins") == 0) & ~zero
    return np.where(ok[None], t, 0.0)


def slonczewski_torque(mag, m):
    """k_Slonczewski (src/physics/stt.cu:104-150): B = (hbar/e) jz / (Ms d),
    eps = P L^2 / ((L^2 + 1) + (L^2 - 1) m.p),
    t = gamma B / (1+alpha^2) ((eps + eps' alpha) m x (p x m) + (eps' - eps alpha) p x m);
    d changes sign when the fixed layer is at the bottom."""
    ms, pol, al = mag.val("msat"), mag.val("pol"), mag.val("alpha")
    lam, epsp = mag.val("Lambda"), mag.val("epsilon_prime")
    th = mag.val("free_layer_thickness") * (1.0 if mag.fixed_on_top else -1.0)
    jz = mag.vec("jcur")[2]
    p = mag.vec("fixed_layer")
    with np.errstate(divide="ignore", invalid="ignore"):
        B = (HBAR / QE) * jz / (ms * th)
        l2 = lam * lam
        eps = pol * l2 / ((l2 + 1) + (l2 - 1) * dot(m, p))
    pxm = cross(p, m)
    mxpxm = cross(m, pxm)
    t = (B / (1 + al * al) * GAMMALL) * ((eps + epsp * al) * mxpxm + (epsp - eps * al) * pxm)
    zero = ((ms == 0) | (jz == 0) | (th == 0) | np.all(p == 0, axis=0)
            | ((epsp == 0) & ((lam == 0) | (pol == 0))))
    ok = mag.geom_mask() & (mag.val("frozen_spins") == 0) & ~zero
    return np.where(ok[None], np.nan_to_num(t), 0.0)


def spin_transfer_torque(mag, m):
    """evalSpinTransferTorque (src/physics/stt.cu:152-188): Slonczewski when it
    is not assured zero, else Zhang-Li, else zero."""
    if not slonczewski_assured_zero(mag):
        return slonczewski_torque(mag, m)
    if not zhang_li_assured_zero(mag):
        return zhang_li_torque(mag, m)
    return np.zeros((3,) + mag.shape)


def _in_grid_neighbor(mag, d):
    """(in-grid mask after the wrap, clipped index) of the neighbour in direction d."""
    nz, ny, nx = mag.shape
    Z, Y, X = np.indices(mag.shape)
    loc = [X, Y, Z]
    n = [nx, ny, nz]
    glob = []
    for a in range(3):
        g = loc[a] + mag.origin[a] + d[a]
        if mag.mastergrid[a] > 0:
            mo = mag.mastergrid_origin[a]
            g = mo + np.mod(g - mo, mag.mastergrid[a])
        glob.append(g - mag.origin[a])
    in_grid = np.ones(mag.shape, dtype=bool)
    clipped = []
    for a in range(3):
        in_grid &= (glob[a] >= 0) & (glob[a] < n[a])
        clipped.append(np.clip(glob[a], 0, n[a] - 1))
    return in_grid, (clipped[2], clipped[1], clipped[0])
