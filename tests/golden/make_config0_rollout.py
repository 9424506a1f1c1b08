"""Writes config0_configured_200.json: the oracle's 200-step rollout of the
8000-particle dam break with the configured s_corr (k = 0.1, n = 4,
dq = 0.2h; 4 iterations, XSPH 0.01, dt = 1/120, seed 0), as mean density,
mean density error |rho / rho0 - 1| and kinetic energy every 10 steps, plus
a SHA-256 of the final state (positions, velocities, densities, lambdas in
original particle order, float32 little-endian).

Run from the repository root in the build container:
    python tests/golden/make_config0_rollout.py
The oracle's serial run takes about ten minutes (the configured s_corr
collapses the block into stacks of coincident particles)."""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from novaphy_amd import scenes  # noqa: E402
from oracle import oracle as O  # noqa: E402


def summary(sc, pos, vel, rho):
    m = float(np.float32(sc.mass))
    v = vel.astype(np.float64)
    return {"mean_density": float(rho.astype(np.float64).mean()),
            "mean_density_error": float(np.abs(rho.astype(np.float64) / sc.params.rest_density - 1.0).mean()),
            "kinetic_energy": float(0.5 * m * (v * v).sum())}


def state_hash(pos, vel, rho, lam):
    h = hashlib.sha256()
    for a in (pos, vel, rho, lam):
        h.update(np.ascontiguousarray(a, "<f4").tobytes())
    return h.hexdigest()


def main():
    sc = scenes.dam_break_8k()
    pos, vel = sc.positions.copy(), sc.velocities.copy()
    rows = []
    for k in range(20):
        pos, vel, rho, lam = O.pbf_run(sc.params, pos, vel, sc.dt, sc.gravity, sc.mass, 10)
        rows.append({"step": 10 * (k + 1), **summary(sc, pos, vel, rho)})
        print(rows[-1], flush=True)
    out = {"scene": "dam_break_8k", "s_corr": {"k": 0.1, "n": 4, "dq": 0.2}, "steps": 200,
           "every": 10, "rows": rows, "final_state_sha256": state_hash(pos, vel, rho, lam)}
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "config0_configured_200.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
