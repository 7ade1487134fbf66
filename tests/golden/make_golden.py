"""Write tests/golden/golden.json (run from the repo root: python tests/golden/make_golden.py).

Contents:
  geometry        counts + SHA-256 of the uint32 flag words of the V60 cup
                  at 32^3, 48^3, 128^3 (scaled) and 224^3
  trilinear_kat   the ten probe positions of the reference's own test
                  tests/test_trilinear_interpolation.py:143-159 (16^3 grid,
                  linear field v = (x, y, z)); the expected interpolant is the
                  position itself (:166 of that file)
  particles_seed  SHA-256 of the seeded 50 000-particle set (positions,
                  radii, masses) of the v60_particles workload
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import __graft_entry__  # noqa: E402

__graft_entry__.load_package()
from pourover_amd.geometry import V60Params, build_v60  # noqa: E402
from pourover_amd import workloads as WL  # noqa: E402

TRILINEAR_KAT = [
    [5.0, 5.0, 5.0], [5.5, 5.5, 5.5], [5.25, 6.75, 7.1], [0.1, 0.1, 0.1], [14.9, 14.9, 14.9],
    [2.3, 8.7, 11.2], [7.8, 3.4, 9.6], [12.1, 13.5, 4.8], [1.7, 5.9, 14.3], [8.4, 11.2, 6.7],
]


def main():
    out = {"geometry": {}}
    for n in (32, 48, 128, 224):
        g = build_v60(V60Params() if n == 224 else V60Params.scaled(n))
        out["geometry"][str(n)] = {"counts": g.counts(), "sha256": g.digest(), "cone_top_z": g.cone_top_z}
        if n == 224:
            seed = WL.seed_particles(g, 50_000)
            h = hashlib.sha256()
            for a in (seed.pos, seed.radius, seed.mass):
                h.update(np.ascontiguousarray(a, dtype=np.float64).tobytes())
            out["particles_seed"] = {"seed": WL.SEED, "count": 50_000, "sha256": h.hexdigest(),
                                     "d_median": float(np.median(2 * seed.radius)),
                                     "z_range": [float(seed.pos[2].min()), float(seed.pos[2].max())]}
    out["trilinear_kat"] = {"grid": 16, "positions": TRILINEAR_KAT, "expected": TRILINEAR_KAT,
                            "source": "tests/test_trilinear_interpolation.py:143-166 (reference)"}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)
    print(path)


if __name__ == "__main__":
    main()
