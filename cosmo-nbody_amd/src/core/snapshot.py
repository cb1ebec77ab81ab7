"""Snapshots of an evolved state (no reference counterpart: the reference
keeps its state in device memory only, src/core/state.py:5-18).

``save(path, engine)`` writes positions, velocities and masses (float32,
bit for bit), the scale factor, the step count, the configuration as JSON
and its SHA-256 to one .npz file (no pickled objects).  ``load(path)``
reads it back; with ``config`` given it refuses a file written under a
different configuration.  ``restore(engine, path)`` puts a snapshot into an
existing engine of the same particle count.
"""
from __future__ import annotations

import dataclasses
import hashlib
import json

import numpy as np
import torch

FORMAT = 1


def config_json(config) -> str:
    """Canonical JSON of a Config (sorted keys)."""
    d = dataclasses.asdict(config) if dataclasses.is_dataclass(config) else dict(config)
    return json.dumps(d, sort_keys=True, default=list)


def config_hash(config) -> str:
    return hashlib.sha256(config_json(config).encode()).hexdigest()


def save(path, engine):
    st = engine.state
    arrays = {f"pos_{k}": t.detach().cpu().numpy() for k, t in zip("xyz", st.pos)}
    arrays.update({f"vel_{k}": t.detach().cpu().numpy() for k, t in zip("xyz", st.vel)})
    arrays["mass"] = st.mass.detach().cpu().numpy()
    np.savez(path, format=np.int64(FORMAT), a=np.float64(engine.a),
             steps_done=np.int64(engine.steps_done),
             config=np.array(config_json(engine.config)),
             config_sha256=np.array(config_hash(engine.config)), **arrays)


def load(path, config=None):
    """(arrays dict, a, steps_done, config dict); arrays hold float32 numpy."""
    d = np.load(path, allow_pickle=False)
    if int(d["format"]) != FORMAT:
        raise ValueError(f"{path}: snapshot format {int(d['format'])}, expected {FORMAT}")
    stored = str(d["config_sha256"])
    if hashlib.sha256(str(d["config"]).encode()).hexdigest() != stored:
        raise ValueError(f"{path}: configuration record does not match its hash")
    if config is not None and config_hash(config) != stored:
        raise ValueError(f"{path}: written under a different configuration")
    arrays = {k: d[k] for k in d.files if k.startswith(("pos_", "vel_")) or k == "mass"}
    return arrays, float(d["a"]), int(d["steps_done"]), json.loads(str(d["config"]))


def restore(engine, path, check_config=True):
    arrays, a, steps, _ = load(path, engine.config if check_config else None)
    st = engine.state
    if arrays["mass"].shape[0] != st.n:
        raise ValueError(f"{path}: {arrays['mass'].shape[0]} particles, engine has {st.n}")
    dev = st.pos[0].device
    for t, k in zip(st.pos, "xyz"):
        t.copy_(torch.from_numpy(arrays[f"pos_{k}"]).to(dev))
    for t, k in zip(st.vel, "xyz"):
        t.copy_(torch.from_numpy(arrays[f"vel_{k}"]).to(dev))
    st.mass.copy_(torch.from_numpy(arrays["mass"]).to(dev))
    engine.a = a
    engine.steps_done = steps
    return engine
