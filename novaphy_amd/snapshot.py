"""Snapshot / restore of a fluid rollout (versioned ``.npz``).

A PBF step reads only the committed positions and velocities (densities and
lambdas are recomputed; they are kept for inspection), so a scene restored
from a snapshot continues bit for bit.  Rigid bodies are captured by their
transforms and velocities (the host rigid solver keeps no other state
between steps); their boundary particles and psi volumes are rebuilt from
the scene definition, as at the start of the rollout.

Arrays (original particle order, float32):
  positions (n, 3), velocities (n, 3), densities (n,), lambdas (n,)
  body_positions (b, 3), body_rotations (b, 4) xyzw, body_linear_velocities
  (b, 3), body_angular_velocities (b, 3)        -- only with rigid bodies
Scalars / metadata:
  format = "novaphy_amd.fluid_snapshot", version = 1, steps_done (int),
  meta = JSON text (free-form: request id, horizon, scene name, ...).
"""
from __future__ import annotations

import json
from typing import Dict, Optional

import numpy as np

FORMAT = "novaphy_amd.fluid_snapshot"
VERSION = 1
_ARRAYS = ("positions", "velocities", "densities", "lambdas", "body_positions", "body_rotations",
           "body_linear_velocities", "body_angular_velocities")


def save(path, snap: Dict[str, np.ndarray], steps_done: int = 0, meta: Optional[dict] = None) -> None:
    arrays = {k: np.ascontiguousarray(snap[k], np.float32) for k in _ARRAYS if k in snap}
    np.savez(path, format=np.array(FORMAT), version=np.array(VERSION, np.int32),
             steps_done=np.array(int(steps_done), np.int64), meta=np.array(json.dumps(meta or {})), **arrays)


def load(path):
    """Returns (snapshot dict, steps_done, meta dict).  Loads without pickles."""
    with np.load(path, allow_pickle=False) as z:
        if str(z["format"]) != FORMAT:
            raise ValueError(f"{path}: not a {FORMAT} file")
        if int(z["version"]) != VERSION:
            raise ValueError(f"{path}: snapshot version {int(z['version'])}, this build reads {VERSION}")
        snap = {k: np.array(z[k], np.float32) for k in _ARRAYS if k in z.files}
        return snap, int(z["steps_done"]), json.loads(str(z["meta"]))
