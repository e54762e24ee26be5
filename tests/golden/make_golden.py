"""Writes the golden vectors of the solver step: the CPU oracle's FP32
state after a few frames of three small scenes (the committed scene files
shrunk to 16^3-20^3). Inputs are the scene settings and the seeded particle
set; outputs are the front buffers T, U, V, W, density and pressure.

Run from the repository root:  python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

CASES = {
    # name: (scene, grid, particles, pscale in cells, frames, overrides)
    "plume16": ("plume64.blz", 16, 1500, 0.9, 4, {"max_iter_solve": 30}),
    "features20": ("plume128.blz", 20, 2500, 1.2, 4, {"max_iter_solve": 30}),
    "explosion16": ("explosion256.blz", 16, 4000, 1.0, 4, {"max_iter_solve": 60, "residual_tol": 1e-2}),
}
FIELDS = {"T": 0, "U": 1, "V": 2, "W": 3, "density": 4, "P": 5}


def case_inputs(name):
    import blaze_amd as B

    scene, n, count, pcells, frames, over = CASES[name]
    s, _, _, gen = B.parse_scene(os.path.join(ROOT, "scenes", scene))
    s.grid_res[0] = s.grid_res[1] = s.grid_res[2] = n
    s.dx = 1.0 / n
    for k, v in over.items():
        setattr(s, k, v)
    gen.count = count
    gen.pscale = pcells / n
    return s, B.generate_particles(gen), frames, s.source_range[1]


def run_case(solver_cls, name):
    s, parts, frames, src_end = case_inputs(name)
    o = solver_cls(s)
    t = np.float32(0.0)
    for it in range(1, frames + 1):
        if it <= src_end:
            o.add_source(parts if it == 1 else None)
        o.step(float(t))
        t = np.float32(t + np.float32(s.dt))
    return {f: o.field(w) for f, w in FIELDS.items() if not (w == 4 and s.smoke_density <= 0)}


def digest(arrs):
    h = hashlib.sha256()
    for k in sorted(arrs):
        h.update(k.encode())
        h.update(np.ascontiguousarray(arrs[k], dtype=np.float32).tobytes())
    return h.hexdigest()


def main():
    import oracle as O

    manifest = {}
    for name in CASES:
        out = run_case(O.Oracle, name)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
        manifest[name] = {"sha256": digest(out), "fields": sorted(out)}
        print(name, manifest[name]["sha256"])
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=2)
        f.write("\n")


if __name__ == "__main__":
    main()
