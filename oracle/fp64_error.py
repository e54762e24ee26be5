"""FP32 against FP64 error of the oracle (test infrastructure: run by hand,
never by the product). The same CPU restatement of the reference step runs
in float and in double from the same inputs; prints the relative L2 error
per field, ||x32 - x64|| / ||x64||, after the given frames of a scene.

usage: python -m oracle.fp64_error SCENE.blz FRAMES [GRID]
(GRID overrides the grid's cells per axis, scaling dx and pscale, and the
particle count with the cell count, so each cell sees the same particles.)
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import blaze_amd as B  # noqa: E402  (settings parser and particle generator only)
import oracle as O  # noqa: E402

NAMES = {0: "T", 1: "U", 2: "V", 3: "W", 4: "density", 5: "P"}


def run(scene, frames, grid=None):
    s, _, _, gen = B.parse_scene(scene)
    if grid:
        f = s.grid_res[0] / grid
        for a in range(3):
            s.grid_res[a] = int(round(s.grid_res[a] / f))
        s.dx *= f
        gen.pscale *= f
        gen.count = max(1, int(round(gen.count / f ** 3)))
    parts = B.generate_particles(gen)
    o32, o64 = O.Oracle(s), O.Oracle(s, fp64=True)
    t = np.float32(0.0)
    for it in range(1, frames + 1):
        frame = s.source_range[0] + it - 1
        if frame <= s.source_range[1]:
            if it == 1 or s.source_animated == 0:
                o32.add_source(parts)
                o64.add_source(parts)
        o32.step(float(t))
        o64.step(float(t))
        t = np.float32(t + np.float32(s.dt))
    out = {}
    for w, name in NAMES.items():
        if w == 4 and s.smoke_density <= 0:
            continue
        a = o32.field(w).astype(np.float64)
        b = o64.field(w)
        n = np.linalg.norm(b)
        out[name] = float(np.linalg.norm(a - b) / n) if n > 0 else 0.0
    return tuple(s.grid_res), out


if __name__ == "__main__":
    scene, frames = sys.argv[1], int(sys.argv[2])
    grid = int(sys.argv[3]) if len(sys.argv) > 3 else None
    res, err = run(scene, frames, grid)
    print(f"{os.path.basename(scene)} {res[0]}x{res[1]}x{res[2]}, {frames} frames: "
          + ", ".join(f"{k} {v:.2e}" for k, v in err.items()))
