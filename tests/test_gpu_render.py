"""GPU tests of the renderer (blaze_amd/csrc/k_render.hip) against the CPU
oracle (oracle/render_oracle.hpp): the same image bit for bit, on a grid
with dx a power of two (grid scaling by multiplication) and one without
(IEEE division), with lights, shadows, emission and the blurred
multi-scatter term; and on a plume state the solver itself produced."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(__file__))
from helpers import SCENES, compare, settings  # noqa: E402

pytestmark = pytest.mark.gpu

B = pytest.importorskip("blaze_amd")
O = pytest.importorskip("oracle")


def ball_field(res, seed):
    W, H, D = res
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.arange(D), np.arange(H), np.arange(W), indexing="ij")
    c = np.array([0.5 * W, 0.45 * H, 0.55 * D])
    r2 = ((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) / (0.3 * min(res)) ** 2
    T = 3.0 * np.exp(-r2) + 0.2 * rng.random((D, H, W))
    return T.astype(np.float32)


def render_settings(w, h, dx):
    r = B.default_render_settings()
    r.width, r.height = w, h
    r.cam_focal, r.cam_aperture = 50.0, 41.4214
    for a, v in enumerate([0.98, 0.0, -0.2]):
        r.cam_u[a] = v
    for a, v in enumerate([0.02, 0.995, 0.1]):
        r.cam_v[a] = v
    for a, v in enumerate([0.2, -0.1, 0.975]):
        r.cam_w[a] = v
    for a, v in enumerate([0.8, 0.4, 2.6]):
        r.cam_trans[a] = v
    r.p_step, r.s_step, r.cutoff = dx, 2 * dx, 0.01
    r.light_count = 2
    for a, v in enumerate([0.15, -0.35, -0.92]):
        r.light_dir[0][a] = v
    for a, v in enumerate([-0.7, -0.2, 0.68]):
        r.light_dir[1][a] = v
    for l in range(2):
        for a in range(3):
            r.light_color[l][a] = 1.0
    r.light_intensity[0], r.light_intensity[1] = 0.6, 0.3
    r.light_exposure[0], r.light_exposure[1] = 0.0, 1.0
    r.density_scale, r.shadow_density_scale = 2.0, 1.0
    for a, (vc, sc) in enumerate(zip([0.8, 0.9, 1.0], [0.75, 1.155, 1.5])):
        r.volume_color[a], r.scattering[a] = vc, sc
    r.absorption, r.gain, r.emission_scale = 0.1, 0.5, 2.0
    r.ramp_size = 4
    keys = [0.0, 0.33, 0.66, 1.0]
    cols = [[0, 0, 0], [1, 0.055, 0], [1, 0.431, 0], [1, 1, 0.077]]
    for q in range(4):
        r.ramp_keys[q] = keys[q]
        for a in range(3):
            r.ramp_colors[q][a] = cols[q][a]
    r.temp_range[0], r.temp_range[1] = 0.2, 3.0
    r.ms_scale, r.ms_blur_iter = 0.625, 3
    r.ms_mask[0], r.ms_mask[1] = 0.0, 0.5
    for a, v in enumerate([1.0, 0.2, 0.0]):
        r.ms_color[a] = v
    return r


@pytest.mark.parametrize("res", [(32, 32, 32), (24, 20, 28)], ids=["32c-pow2dx", "24x20x28"])
def test_render_bitwise(res):
    print(f"[render] {res}", flush=True)
    s = settings(B, res)
    g = B.Solver(s)
    T = ball_field(res, 3)
    g.set_field(B.FIELD_T, T)
    r = render_settings(64, 48, float(s.dx))
    img = g.render(r)
    want = O.render(T, float(s.dx), list(s.domain_min), list(s.domain_max), r)
    print(f"[render] max {img.max():.4f} lit pixels {(img.sum(axis=2) > 0).mean():.2f}", flush=True)
    assert (img.sum(axis=2) > 0).mean() > 0.2
    compare(f"render {res}", img, want)


def test_render_of_a_solver_state():
    """The plume scene's own render settings on the state after 6 frames
    at 32^3 (image reduced to 64x40)."""
    print("[render] plume state", flush=True)
    path = os.path.join(SCENES, "plume64.blz")
    s, _, _, gen = B.parse_scene(path)
    f = s.grid_res[0] / 32
    for a in range(3):
        s.grid_res[a] = 32
    s.dx = s.dx * f
    gen.pscale = gen.pscale * f
    parts = B.generate_particles(gen)
    g = B.Solver(s)
    t = np.float32(0.0)
    for _ in range(6):
        g.add_source(parts)
        g.step(float(t))
        t = np.float32(t + np.float32(s.dt))
    r = B.parse_render(path)
    r.width, r.height = 64, 40
    r.p_step, r.s_step = float(s.dx), 2 * float(s.dx)
    img = g.render(r)
    T = g.field(B.FIELD_T)
    want = O.render(T, float(s.dx), list(s.domain_min), list(s.domain_max), r)
    print(f"[render] plume max {img.max():.4f}", flush=True)
    assert img.max() > 0
    compare("render plume", img, want)


def test_render_rejects_bad_sizes():
    s = settings(B, (16, 16, 16))
    g = B.Solver(s)
    r = render_settings(60, 48, float(s.dx))
    with pytest.raises(B.BlazeError, match="multiples of 8"):
        g.render(r)
    r = render_settings(64, 48, float(s.dx))
    r.p_step = 0.0
    with pytest.raises(B.BlazeError, match="steps > 0"):
        g.render(r)


def test_cli_renders_frames(tmp_path):
    """`blaze -l scene.blz` renders and writes each frame of the render range
    (src/main.cu:89-106, RenderEngine::writeToDisk naming): the plume at
    32^3, frames 1-3 as PNG under <sceneDir>/../render/, the per-frame stats
    table on stdout and, with the scene summary, in the timestamped log under
    <sceneDir>/../logs/ (src/Timer.h:80-132, src/Logger.h:52-63)."""
    import json
    import subprocess
    import zlib

    print("[render] cli", flush=True)
    d = json.load(open(os.path.join(SCENES, "plume64.blz")))
    d["gridRes"] = [32, 32, 32]
    d["dx"] = d["dx"] * 2
    d["particleGen"]["pscale"] = d["particleGen"]["pscale"] * 2
    d["renderRange"] = [1, 3, 1]
    d["renderRes"] = [64, 40]
    d["primarySampleRate"] = d["dx"]
    d["shadowSamplingRate"] = 2 * d["dx"]
    d["renderFile"] = "../render/t_0001.png"
    (tmp_path / "scenes").mkdir()
    scene = tmp_path / "scenes" / "t.blz"
    scene.write_text(json.dumps(d))
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "blaze_amd", "bin", "blaze")
    r = subprocess.run([exe, "-l", str(scene)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-400:], flush=True)
    assert r.returncode == 0, r.stderr
    assert "Frame: 003" in r.stdout and "Gauss-Seidel Solve:" in r.stdout and "Render:" in r.stdout
    logs = list((tmp_path / "logs").glob("log_*.txt"))
    assert len(logs) == 1
    text = logs[0].read_text()
    assert text.startswith("sceneName:") and text.count("Frame: ") == 3 and "Total Time:" in text
    for f in (1, 2, 3):
        data = (tmp_path / "render" / f"t_{f:04d}.png").read_bytes()
        assert data[:8] == b"\x89PNG\r\n\x1a\n"
        idat = data.index(b"IDAT")
        n = int.from_bytes(data[idat - 4:idat], "big")
        raw = np.frombuffer(zlib.decompress(data[idat + 4:idat + 4 + n]), np.uint8).reshape(40, 1 + 4 * 64)
        if f == 3:
            assert raw[:, 1:].reshape(40, 64, 4)[..., :3].max() > 0
