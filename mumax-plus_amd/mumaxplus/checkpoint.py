"""Snapshots and checkpoints of a World's dynamic state.

The reference writes tables of quantities while it runs (TimeSolverOutput,
timesolver.py:7-46) but cannot save and resume a simulation.  A snapshot
holds everything a later run depends on: the magnetisation of every
ferromagnet and of both sublattices of every antiferromagnet, the elastic
displacement and velocity of magnets with elastodynamics, the thermal-noise
stream of every ferromagnet and sublattice (its key and 64-bit draw counters,
so that a restored run draws the noise the uninterrupted run would have drawn),
and the time solver's method, time, step and controller settings.  Parameters
are the script's business and are not part of it.

A checkpoint file is a NumPy ``.npz`` of plain arrays plus the settings as
a JSON string; it is read back with ``allow_pickle=False``.
"""

import json

import numpy as np

_SOLVER_KEYS = ("time", "timestep", "max_error", "headroom", "lower_bound", "upper_bound", "sensible_factor",
                "sensible_timestep_default")


def _state_arrays(world):
    out = {}
    for name, fm in world.ferromagnets.items():
        out[f"ferromagnet/{name}/magnetization"] = fm.magnetization.get()
        if fm.enable_elastodynamics:
            out[f"ferromagnet/{name}/elastic_displacement"] = fm.elastic_displacement.get()
            out[f"ferromagnet/{name}/elastic_velocity"] = fm.elastic_velocity.get()
    for name, afm in world.antiferromagnets.items():
        out[f"antiferromagnet/{name}/sub1"] = afm.sub1.magnetization.get()
        out[f"antiferromagnet/{name}/sub2"] = afm.sub2.magnetization.get()
        if afm.enable_elastodynamics:
            out[f"antiferromagnet/{name}/elastic_displacement"] = afm.elastic_displacement.get()
            out[f"antiferromagnet/{name}/elastic_velocity"] = afm.elastic_velocity.get()
    return out


def _noise_owners(world):
    """{key: ferromagnet or sublattice} of every thermal-noise stream."""
    out = {f"ferromagnet/{name}": fm for name, fm in world.ferromagnets.items()}
    for name, afm in world.antiferromagnets.items():
        out[f"antiferromagnet/{name}/sub1"] = afm.sub1
        out[f"antiferromagnet/{name}/sub2"] = afm.sub2
    return out


def snapshot(world):
    """The world's dynamic state: {"arrays": {key: ndarray}, "noise": {key: [seed, step draws,
    quantity draws]}, "solver": {...}}."""
    ts = world.timesolver
    solver = {k: float(getattr(ts, k)) for k in _SOLVER_KEYS}
    solver["adaptive_timestep"] = bool(ts.adaptive_timestep)
    solver["method"] = ts._impl.method
    noise = {k: list(mag._impl.noise_state) for k, mag in _noise_owners(world).items()}
    return {"arrays": {k: np.array(v, copy=True) for k, v in _state_arrays(world).items()}, "noise": noise,
            "solver": solver}


def restore(world, snap):
    """Put a snapshot of this world (same magnets) back."""
    arrays = snap["arrays"]
    fms, afms = world.ferromagnets, world.antiferromagnets
    for key, value in arrays.items():
        kind, name, what = key.split("/")
        magnets = fms if kind == "ferromagnet" else afms
        if name not in magnets:
            raise KeyError(f"the snapshot has a {kind} '{name}' this world does not have")
        mag = magnets[name]
        if what == "magnetization":
            mag.magnetization = value
        elif what in ("sub1", "sub2"):
            getattr(mag, what).magnetization = value
        elif what == "elastic_displacement":
            mag.elastic_displacement = value
        elif what == "elastic_velocity":
            mag.elastic_velocity = value
        else:
            raise KeyError(f"unknown state '{key}'")
    owners = _noise_owners(world)
    for key, state in snap.get("noise", {}).items():
        if key not in owners:
            raise KeyError(f"the snapshot has a noise stream '{key}' this world does not have")
        owners[key]._impl.noise_state = state
    s = snap["solver"]
    ts = world.timesolver
    ts.set_method(s["method"])
    ts.adaptive_timestep = s["adaptive_timestep"]
    for k in _SOLVER_KEYS:
        setattr(ts, k, s[k])


def save(world, path):
    """Write a checkpoint of the world to `path` (.npz)."""
    snap = snapshot(world)
    np.savez(path, _solver=np.array(json.dumps(snap["solver"])), _noise=np.array(json.dumps(snap["noise"])),
             **snap["arrays"])


def load(world, path):
    """Restore the world from a checkpoint written by `save`."""
    with np.load(path, allow_pickle=False) as f:
        solver = json.loads(str(f["_solver"]))
        noise = json.loads(str(f["_noise"])) if "_noise" in f.files else {}
        arrays = {k: f[k] for k in f.files if k not in ("_solver", "_noise")}
    restore(world, {"arrays": arrays, "noise": noise, "solver": solver})
