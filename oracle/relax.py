"""Relaxation and energy minimisation in FP64.

Test infrastructure only: imported by the tests, never by the product.

* ``total_energy_density(world, fm)`` -- evalTotalEnergyDensity of a
  ferromagnet or sublattice (src/physics/energy.cu:52-68): every active term's
  -prefactor Ms m.h (0.5 for exchange, DMI, demag, magnetoelastic and both AFM
  exchanges, 1 for Zeeman) plus the anisotropy energy; an independent magnet
  with elastodynamics adds its kinetic and elastic energy densities
  (elasticenergies.cu:7-101).  ``afm_total_energy_density`` adds both
  sublattices and the magnet's kinetic and elastic terms (energy.cu:70-76).
* ``relax(world, target, tol)`` -- Relaxer::exec (src/physics/relaxer.cpp:
  87-174), target a ferromagnet, an antiferromagnet or None (the world).
* ``minimize(world, target, tol, nsamples)`` -- Minimizer (minimizer.cu:8-160).

Pinned by tests/test_oracle_relax.py: the standard problem 2 coercivity in
the small-particle limit (test/test_standardproblem2_coercivity.py,
0.057069478 within 1e-5, minimize), the Neel-vector surface canting of
test/test_neel_canting.py (relax, 2 %), the Stoner-Wohlfarth single-spin
minimum (closed form), and the restore rules of the solver state.
"""

import math
from collections import deque

import numpy as np

from . import afm as _afm
from . import elastic as _elastic
from . import fields as _fields
from .model import Antiferromagnet, Ferromagnet


def _sister(fm):
    h = fm.host
    return h.sub2 if fm is h.sub1 else h.sub1


def total_energy_density(world, fm):
    m = fm.m
    own = fm.host if fm.host is not None else fm
    m2 = _sister(fm).m if fm.host is not None else None
    e = np.zeros(fm.shape)

    def add(h, pref):
        nonlocal e
        e = e + _fields.energy_density(fm, m, h, pref)

    if fm.host is not None and not fm.openbc:
        add(_afm.exchange_field_sublattice(fm, m, m2), 0.5)
    else:
        add(_fields.exchange_field(fm, m), 0.5)
    e = e + _fields.anisotropy_energy_density(fm, m)
    hs = world.stray_field_on(own)
    if not _fields.external_assured_zero(fm, world.bias) or hs is not None:
        h_ext = _fields.external_field(fm, m, world.bias)
        if hs is not None:
            h_ext = h_ext + np.where(fm.geom_mask()[None], hs, 0.0)
        add(h_ext, 1.0)
    add(_fields.dmi_field(fm, m) if fm.host is None else _afm.dmi_field_afm(fm, m, m2), 0.5)
    if fm.enable_demag and not fm.assured_zero("msat"):
        if fm.host is None:
            add(_fields.demag_field(fm, m), 0.5)
        else:
            h = fm.host
            M = h.sub1.val("msat") * h.sub1.m + h.sub2.val("msat") * h.sub2.m
            add(h.demag_plan().field(M, np.ones(h.shape), h.geom_mask()), 0.5)
    if not world.magnetoelastic_assured_zero(fm):
        add(world.magnetoelastic_field(fm, m, own.u), 0.5)
    if fm.host is not None:
        if not _afm.inhomo_assured_zero(fm):
            add(_afm.afm_nn_field(fm, m, m2), 0.5)
        if not _afm.homo_assured_zero(fm):
            add(_afm.afm_site_field(fm, m2), 0.5)
    if fm.host is None:
        e = e + _elastic_energy_densities(fm)
    return e


def _elastic_energy_densities(own):
    e = np.zeros(own.shape)
    if not own.enable_elastodynamics:
        return e
    geom = own.geom_mask()
    if not own.assured_zero("rho"):
        e = e + np.where(geom, 0.5 * own.val("rho") * np.sum(own.v * own.v, axis=0), 0.0)
    if not own.elasticity_assured_zero():
        strain = _elastic.strain_tensor(own, own.u)
        stress = _elastic.stress_tensor(own, strain)
        e = e + np.where(geom, 0.5 * np.sum(stress[:3] * strain[:3], axis=0)
                         + np.sum(stress[3:] * strain[3:], axis=0), 0.0)
    return e


def afm_total_energy_density(world, afm):
    return (total_energy_density(world, afm.sub1) + total_energy_density(world, afm.sub2)
            + _elastic_energy_densities(afm))


def total_energy(world, magnet):
    """ncells * average over the geometry * cell volume (energy.cu:78-92)."""
    if isinstance(magnet, Antiferromagnet):
        return _fields.integrate(magnet.sub1, afm_total_energy_density(world, magnet))
    return _fields.integrate(magnet, total_energy_density(world, magnet))


def _targets(world, target):
    if target is not None:
        return [target]
    return list(world.magnets) + list(world.afms)


def _ferromagnets(magnets):
    out = []
    for mag in magnets:
        out += [mag.sub1, mag.sub2] if isinstance(mag, Antiferromagnet) else [mag]
    return out


def damping_torque(world, fm):
    m2 = _sister(fm).m if fm.host is not None else None
    own = fm.host if fm.host is not None else fm
    return _fields.damping_torque(fm, fm.m, world.effective_field(fm, fm.m, m2, own.u))


def relax(world, target=None, tol=1e-9, thresholds=None):
    """Relaxer::exec.  thresholds: one per ferromagnet or sublattice of the
    targets (default -1: relax until the torque stops decreasing)."""
    ts = world.solver
    saved = (ts.time, ts.dt, ts.adaptive, ts.max_error, ts.method)
    targets = _targets(world, target)
    fms = _ferromagnets(targets)
    thr = list(thresholds) if thresholds is not None else [-1.0] * len(fms)
    try:
        ts.adaptive = True
        ts.method = "Fehlberg"
        world.relax_fms = fms
        world.relax_elastic = target is None
        ts.equations_changed()
        NS = 3

        def energy():
            return sum(total_energy(world, t) for t in targets)

        def torque_sum():
            return sum(float(np.sum(damping_torque(world, fm) ** 2)) for fm in fms)

        def run(n):
            for _ in range(n):
                ts.step()

        E0 = energy()
        run(NS)
        E1 = energy()
        while E1 < E0:
            run(NS)
            E0, E1 = E1, energy()
        if all(t < 0 for t in thr):
            t1 = torque_sum()
            err = ts.max_error
            while err > tol:
                err /= math.sqrt(2)
                ts.max_error = err
                run(NS)
                t0, t1 = t1, torque_sum()
                while t1 < t0:
                    run(NS)
                    t0, t1 = t1, torque_sum()
        elif any(t == 0 for t in thr):
            raise ValueError("The relax threshold should not be zero.")
        else:
            err = ts.max_error
            while err > tol:
                ok = all(_max_norm(damping_torque(world, fm), fm.geom_mask()) <= t for fm, t in zip(fms, thr))
                if ok:
                    err /= math.sqrt(2)
                    ts.max_error = err
                run(NS)
    finally:
        world.relax_fms = None
        world.relax_elastic = False
        ts.time, ts.dt, ts.adaptive, ts.max_error, ts.method = saved
        if ts.adaptive:
            ts.dt = ts.sensible_timestep()


def _max_norm(v, geom):
    n = np.sqrt(np.sum(v * v, axis=0))[geom]
    return float(n.max()) if n.size else 0.0


def minimize(world, target=None, tol=1e-6, nsamples=10, max_steps=None):
    """Minimizer::exec: steepest descent along the damping torque with
    Barzilai-Borwein step sizes starting at 1e-14."""
    fms = _ferromagnets(_targets(world, target))
    if target is None:
        nsamples *= len(fms)
    h = [1e-14] * len(fms)
    last = deque()
    t1 = [None] * len(fms)
    nsteps = 0

    def converged():
        return len(last) >= nsamples and all(d <= tol for d in last)

    while not converged():
        if max_steps is not None and nsteps >= max_steps:
            raise RuntimeError("minimize did not converge")
        m0 = [fm.m.copy() for fm in fms]
        t0 = [damping_torque(world, fm) if nsteps == 0 else t1[i] for i, fm in enumerate(fms)]
        m1 = []
        for i in range(len(fms)):
            t2 = h[i] * h[i] * np.sum(t0[i] * t0[i], axis=0)
            m1.append(((4 - t2) * m0[i] + 4 * h[i] * t0[i]) / (4 + t2))
        for fm, m in zip(fms, m1):
            fm.set_m(m)
        t1 = [damping_torque(world, fm) for fm in fms]
        for i, fm in enumerate(fms):
            geom = fm.geom_mask()
            dm = np.where(geom[None], m1[i] - m0[i], 0.0)
            dt = np.where(geom[None], t0[i] - t1[i], 0.0)
            if nsteps % 2 == 0:
                nom, div = float(np.sum(dm * dm)), float(np.sum(dm * dt))
            else:
                nom, div = float(np.sum(dm * dt)), float(np.sum(dt * dt))
            h[i] = 1e-14 if div == 0.0 else nom / div
            last.append(_max_norm(dm, geom))
            if len(last) > nsamples:
                last.popleft()
        nsteps += 1
    return nsteps
