"""GPU test of blz_render_noalloc, the render entry point that allocates
nothing (its scatter volumes and frame buffer live in the library's static
per-device scratch): through the C ABI, two calls in a row, the second on
a smaller grid and frame, each equal bit for bit to blz_render on the same
state (whose buffers come from the handle's own allocation); a frame beyond
the scratch limit is refused before any launch."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(__file__))
from helpers import compare, settings  # noqa: E402
from test_gpu_render import ball_field, render_settings  # noqa: E402

pytestmark = pytest.mark.gpu

B = pytest.importorskip("blaze_amd")


def test_render_noalloc_equals_render_twice_smaller_second():
    print("[noalloc] 64^3 grid, 320x184 frame, then 32x24x28, 96x64", flush=True)
    for res, (w, h), seed in (((64, 64, 64), (320, 184), 3), ((32, 24, 28), (96, 64), 4)):
        s = settings(B, res)
        g = B.Solver(s)
        g.set_field(B.FIELD_T, ball_field(res, seed))
        r = render_settings(w, h, float(s.dx))
        a = g.render_noalloc(r)
        want = g.render(r)
        print(f"[noalloc] {res} {w}x{h}: lit {(a.sum(axis=2) > 0).mean():.2f}", flush=True)
        assert (a.sum(axis=2) > 0).mean() > 0.2
        compare(f"noalloc render {res}", a, want)
        g.close()


def test_render_noalloc_without_scatter_volumes():
    """No multi-scatter term: the volume the march reads is zero-filled by a
    kernel, also after a call that left other values in the scratch."""
    print("[noalloc] no multi-scatter", flush=True)
    s = settings(B, (32, 32, 32))
    g = B.Solver(s)
    g.set_field(B.FIELD_T, ball_field((32, 32, 32), 5))
    r = render_settings(64, 48, float(s.dx))
    g.render_noalloc(r)   # leaves scatter values in the scratch
    r.ms_scale = 0.0
    compare("noalloc render, no scatter", g.render_noalloc(r), g.render(r))


def test_render_noalloc_refuses_frames_beyond_the_scratch():
    print("[noalloc] oversized frame", flush=True)
    s = settings(B, (16, 16, 16))
    g = B.Solver(s)
    r = render_settings(4096, 2048, float(s.dx))   # 25M floats > 8M
    with pytest.raises(B.BlazeError) as e:
        g.render_noalloc(r)
    assert e.value.code == B.E_INVALID and "scratch limit" in str(e.value)
