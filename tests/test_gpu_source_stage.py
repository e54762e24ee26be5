"""GPU tests of staged (asynchronous) particle uploads for animated sources:
frames uploaded on the copy stream from another host thread while the
previous frame steps give the same fields bit for bit as synchronous
uploads; the C++ drop-in (which prefetches the next .part file on a worker
thread) writes the same frames as the synchronous path."""
import concurrent.futures as cf
import json
import os
import subprocess
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(__file__))
from helpers import SCENES, compare  # noqa: E402

pytestmark = pytest.mark.gpu

B = pytest.importorskip("blaze_amd")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def plume32():
    s, _, _, gen = B.parse_scene(os.path.join(SCENES, "plume64.blz"))
    for a in range(3):
        s.grid_res[a] = 32
    s.dx = s.dx * 2
    gen.pscale = gen.pscale * 2
    return s, gen


def frames(gen, n):
    out = []
    for f in range(n):
        gen.seed = 100 + f
        gen.count = 12000 + 3000 * f
        out.append(B.generate_particles(gen))
    return out


def test_staged_uploads_equal_synchronous():
    print("[stage] python", flush=True)
    s, gen = plume32()
    parts = frames(gen, 5)
    a, b = B.Solver(s), B.Solver(s)
    t = np.float32(0.0)
    pool = cf.ThreadPoolExecutor(1)
    b.stage_source(parts[0])
    for f in range(5):
        a.add_source(parts[f])
        a.step(float(t))
        b.add_source_staged()
        nxt = pool.submit(b.stage_source, parts[f + 1]) if f + 1 < 5 else None
        b.step(float(t))          # concurrent with the staging thread
        if nxt is not None:
            nxt.result()
        t = np.float32(t + np.float32(s.dt))
    pool.shutdown()
    for w in (B.FIELD_T, B.FIELD_U, B.FIELD_V, B.FIELD_W, B.FIELD_P):
        compare(f"staged field {w}", b.field(w), a.field(w))


def test_cli_animated_part_files(tmp_path):
    """`blaze` on an animated .part scene (prefetching worker) writes the
    same PNG frames as the synchronous Python path."""
    print("[stage] cli", flush=True)
    s, gen = plume32()
    parts = frames(gen, 4)
    sd = tmp_path / "scenes"
    sd.mkdir()
    for f, p in enumerate(parts, start=1):
        p.astype(np.float32).tofile(str(sd / f"anim_{f:04d}.part"))
    d = json.load(open(os.path.join(SCENES, "plume64.blz")))
    d.pop("particleGen", None)
    d.update({"gridRes": [32, 32, 32], "dx": d["dx"] * 2, "particleFile": "anim_0001.part", "sourceAnimated": 1,
              "sourceRange": [1, 4, 1], "renderRange": [1, 4, 1], "renderRes": [64, 40],
              "primarySampleRate": d["dx"] * 2, "shadowSamplingRate": d["dx"] * 4,
              "renderFile": "../render/a_0001.png", "sourceMaxParticleCount": 30000})
    (sd / "a.blz").write_text(json.dumps(d))
    r = subprocess.run([os.path.join(ROOT, "blaze_amd", "bin", "blaze"), str(sd / "a.blz")], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-800:] + r.stderr[-800:]
    # the synchronous path in Python, same settings, frames and render keys
    s2, _, _, _ = B.parse_scene(str(sd / "a.blz"))
    g = B.Solver(s2)
    rs = B.parse_render(str(sd / "a.blz"))
    t = np.float32(0.0)
    for f in range(4):
        g.add_source(parts[f])
        g.step(float(t))
        t = np.float32(t + np.float32(s2.dt))
    out = tmp_path / "sync.png"
    B.write_image(str(out), g.render(rs))
    assert (tmp_path / "render" / "a_0004.png").read_bytes() == out.read_bytes()
