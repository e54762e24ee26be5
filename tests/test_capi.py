"""Host-side tests of the C ABI (include/blaze.h) that need no GPU: the
library loads and exports every declared entry point, the struct layouts
seen by Python match the C compiler's, scene/particle I/O follows the
reference Parser (src/Parser.h:41-211), and error paths return codes
instead of aborting."""
import ctypes as C
import json
import os
import re
import subprocess
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(__file__))
from helpers import ROOT, SCENES  # noqa: E402

B = pytest.importorskip("blaze_amd")
HEADER = os.path.join(ROOT, "include", "blaze.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(blz_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 20
    lib = C.CDLL(B.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), f"{n} declared in blaze.h but not exported"
        assert n in B.ABI_SYMBOLS, f"{n} not bound in the Python mirror"
    assert B.abi_version() == 1


def test_struct_layouts_match_c(tmp_path):
    """sizeof/offsetof of the C structs (gcc on the header) equal ctypes'."""
    src = tmp_path / "layout.c"
    fields = {"blz_settings": [f[0] for f in B.Settings._fields_],
              "blz_stats": [f[0] for f in B.Stats._fields_],
              "blz_particle_gen": [f[0] for f in B.ParticleGen._fields_],
              "blz_batch_stats": [f[0] for f in B.BatchStats._fields_]}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    for st, fs in fields.items():
        lines.append(f'printf("{st} %zu\\n", sizeof({st}));')
        for f in fs:
            lines.append(f'printf("{st}.{f} %zu\\n", offsetof({st}, {f}));')
    lines.append("return 0;}")
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    out = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True, check=True)
               .stdout.splitlines())
    for st, cls in (("blz_settings", B.Settings), ("blz_stats", B.Stats), ("blz_particle_gen", B.ParticleGen),
                    ("blz_batch_stats", B.BatchStats)):
        assert int(out[st]) == C.sizeof(cls), st
        for f in fields[st]:
            assert int(out[f"{st}.{f}"]) == getattr(cls, f).offset, f"{st}.{f}"


@pytest.mark.parametrize("name", ["plume64.blz", "plume128.blz", "explosion256.blz"])
def test_parse_committed_scenes(name):
    s, d, pf, gen = B.parse_scene(os.path.join(SCENES, name))
    raw = json.load(open(os.path.join(SCENES, name)))
    assert list(s.grid_res) == raw["gridRes"]
    assert s.dx == pytest.approx(raw["dx"], rel=1e-7)
    assert s.max_iter_solve == raw["maxIterSolve"]
    assert list(s.closed_bounds) == raw["closedBounds"]
    # ranges are read as floats and truncated (src/Parser.h:58-64)
    assert list(s.source_range) == [int(raw["sourceRange"][0]), int(raw["sourceRange"][1])]
    assert s.noise_seed == raw["noiseSeed"]
    assert os.path.realpath(d) == os.path.realpath(SCENES)
    assert pf == raw["particleFile"]
    assert gen is not None and gen.count == raw["particleGen"]["count"]


REF_SCENES = "/root/reference/scenes"
REF_SCENE_NAMES = ["explosion", "fireDebris", "groundFire", "rocketLaunch", "smokePlume", "tornado"]


@pytest.mark.parametrize("name", REF_SCENE_NAMES)
def test_parse_reference_scenes(name):
    """The reference's own six scene files (its only data) parse through
    Parser::sceneParse's schema (src/Parser.h:41-172): every solver, source
    and render key equals an independent JSON reader's value, rounded to
    float32 as the reference stores it. Skipped where the reference tree is
    absent."""
    path = os.path.join(REF_SCENES, name + ".blz")
    if not os.path.exists(path):
        pytest.skip("reference scenes not present")
    raw = json.load(open(path))
    f32 = lambda v: float(np.float32(v))  # noqa: E731
    s, d, pf, gen = B.parse_scene(path)
    assert gen is None and pf == raw["particleFile"]
    assert os.path.realpath(d) == os.path.realpath(REF_SCENES)
    assert list(s.grid_res) == raw["gridRes"]
    assert list(s.closed_bounds) == raw["closedBounds"]
    assert s.max_iter_solve == raw["maxIterSolve"]
    assert s.wind_dir == raw["windDir"]
    assert s.source_animated == raw["sourceAnimated"]
    assert s.source_max_particle_count == raw["sourceMaxParticleCount"]
    # ranges are read as floats and truncated (src/Parser.h:58-64)
    assert list(s.source_range) == [int(raw["sourceRange"][0]), int(raw["sourceRange"][1])]
    scalars = {"dx": "dx", "dt": "dt", "density": "density", "gravity": "gravity", "buoyancy": "buoyancy",
               "cooling_rate": "coolingRate", "vorticity_conf": "vorticityConf", "drag": "drag",
               "wind_amp": "windAmp", "wind_speed": "windSpeed", "wind_turb_amp": "windTurbAmp",
               "wind_turb_scale": "windTurbScale", "turbulence_amp": "turbulenceAmp",
               "turbulence_scale": "turbulenceScale"}
    for field, key in scalars.items():
        assert getattr(s, field) == f32(raw[key]), key
    for field, key in (("domain_min", "domainBboxMin"), ("domain_max", "domainBboxMax"),
                       ("turb_mask_temp_ramp", "turbMaskTempRamp"), ("turb_mask_vel_ramp", "turbMaskVelRamp")):
        assert list(getattr(s, field)) == [f32(v) for v in raw[key]], key
    # extension keys are absent: reference behaviour
    assert s.noise_seed == 123 and s.residual_tol == 0.0 and s.cfl_number == 0.0 and s.smoke_density == 0.0
    r = B.parse_render(path)
    assert [r.width, r.height] == raw["renderRes"]
    assert r.light_count == raw["lightCount"] == len(raw["lights"])
    for q, L in enumerate(raw["lights"]):
        assert list(r.light_dir[q]) == [f32(v) for v in L["light_dir"]]
        assert list(r.light_color[q]) == [f32(v) for v in L["light_color"]]
        assert r.light_intensity[q] == f32(L["light_intensity"])
    ramp = raw["shaderColorRamp"]
    assert r.ramp_size == ramp["count"]
    assert list(r.ramp_keys)[: ramp["count"]] == [f32(v) for v in ramp["keys"]]
    for field, key in (("p_step", "primarySampleRate"), ("s_step", "shadowSamplingRate"), ("cutoff", "cutoff"),
                       ("density_scale", "densityScale"), ("absorption", "absorption"), ("gain", "gain"),
                       ("emission_scale", "emissionScale")):
        assert getattr(r, field) == f32(raw[key]), key
    assert r.ms_blur_iter == raw["multiScatterBlurIter"]


def reference_style_scene(tmp_path, drop=None, extra=None):
    """A .blz with every key of the reference's smokePlume.blz schema
    (solver, source and render keys) but no extension keys."""
    d = {
        "turbulenceScale": 2.0, "cutoff": 0.01, "lightCount": 0, "lights": [],
        "domainBboxMin": [0.389, 0.0012, -0.617], "turbMaskTempRamp": [0.033, 0.145],
        "emissionScale": 0.5, "renderRes": [1280, 720], "windAmp": 0.2, "buoyancy": 0.075,
        "turbMaskVelRamp": [0.0, 0.0586], "density": 0.1, "coolingRate": -0.1,
        "sourceAnimated": 1, "gravity": -9.8, "turbulenceAmp": 2.0, "renderRange": [1, 1000, 1],
        "domainBboxMax": [7.02, 5.39, 1.18], "gridRes": [768, 624, 208],
        "sourceMaxParticleCount": 64720, "particleFile": "../geo/smokePlume_0001.part",
        "maxIterSolve": 200, "drag": 0.0, "dx": 0.008635298892234763, "dt": 0.041666666666666664,
        "windTurbAmp": 7.5, "vorticityConf": 0.0, "sceneName": "smokePlume", "windTurbScale": 0.125,
        "sourceRange": [1.0, 1000.0, 1.0], "windDir": 0, "windSpeed": 0.2,
        "closedBounds": [0, 0, 1, 1, 1, 1], "renderFile": "../render/smokePlume_0001.exr",
    }
    if drop:
        d.pop(drop)
    if extra:
        d.update(extra)
    p = tmp_path / "scene.blz"
    p.write_text(json.dumps(d))
    return p


def test_reference_schema_defaults_to_reference_behaviour(tmp_path):
    s, _, pf, gen = B.parse_scene(str(reference_style_scene(tmp_path)))
    assert gen is None
    assert list(s.grid_res) == [768, 624, 208]
    assert s.dx == np.float32(0.008635298892234763)
    assert s.cooling_rate == np.float32(-0.1)
    assert s.noise_seed == 123          # compute_turbulence_kernel, src/FluidSolver.cu:304
    assert s.residual_tol == 0.0        # fixed sweep count
    assert s.cfl_number == 0.0          # no substeps
    assert s.smoke_density == 0.0       # no smoke grid
    assert pf == "../geo/smokePlume_0001.part"


@pytest.mark.parametrize("key", ["dx", "gridRes", "closedBounds", "maxIterSolve", "windDir"])
def test_missing_required_key_is_an_error(tmp_path, key):
    p = reference_style_scene(tmp_path, drop=key)
    with pytest.raises(B.BlazeError, match=key):
        B.parse_scene(str(p))


def test_malformed_and_missing_scene_files(tmp_path):
    bad = tmp_path / "bad.blz"
    bad.write_text('{"dx": 0.1, ')
    with pytest.raises(B.BlazeError, match="malformed"):
        B.parse_scene(str(bad))
    with pytest.raises(B.BlazeError, match="does not exist"):
        B.parse_scene(str(tmp_path / "nope.blz"))


def test_part_file_roundtrip_and_partial_record(tmp_path):
    """.part = 8 float32 per particle, no header; a trailing partial record is
    ignored (src/Parser.h:196-210)."""
    rng = np.random.default_rng(0)
    a = rng.standard_normal((37, 8)).astype(np.float32)
    p = tmp_path / "x_0001.part"
    p.write_bytes(a.tobytes() + b"\x01\x02\x03")
    got = B.read_particles(str(p))
    assert np.array_equal(got, a)
    with pytest.raises(B.BlazeError, match="does not exist"):
        B.read_particles(str(tmp_path / "missing.part"))


def test_particle_generator_deterministic_ball():
    g = B.ParticleGen(count=5000, seed=3, radius=0.05, pscale=0.01, temp=5.0, radial_vel=2.0)
    for a, v in enumerate((0.5, 0.25, 0.5)):
        g.center[a] = v
    a = B.generate_particles(g)
    b = B.generate_particles(g)
    assert np.array_equal(a, b)
    r = np.linalg.norm(a[:, :3].astype(np.float64) - [0.5, 0.25, 0.5], axis=1)
    assert r.max() <= 0.05 * (1 + 1e-6)
    speed = np.linalg.norm(a[:, 3:6], axis=1)
    assert np.allclose(speed, 2.0, rtol=1e-5)
    assert np.all(a[:, 6] == np.float32(0.01)) and np.all(a[:, 7] == 5.0)
    # uniform in the ball: the fraction inside half the radius is 1/8
    assert abs(np.mean(r < 0.025) - 0.125) < 0.02
    g.seed = 4
    assert not np.array_equal(a, B.generate_particles(g))


def test_create_rejects_bad_settings():
    s = B.default_settings()
    s.grid_res[0] = 15   # W must be even
    with pytest.raises(B.BlazeError, match="even"):
        B.Solver(s)
    s = B.default_settings()
    s.dx = 0.0
    with pytest.raises(B.BlazeError, match="positive"):
        B.Solver(s)
    s = B.default_settings()
    s.grid_res[2] = 2
    with pytest.raises(B.BlazeError, match="grid_res"):
        B.Solver(s)


def test_cli_argument_handling(tmp_path):
    """`blaze` mirrors src/main.cu:41-70: missing file, wrong extension,
    help; all exit non-zero before touching the GPU."""
    exe = os.path.join(ROOT, "blaze_amd", "bin", "blaze")
    assert os.path.exists(exe)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 1 and "No scene file" in r.stdout
    r = subprocess.run([exe, "-h"], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage" in r.stdout
    r = subprocess.run([exe, str(tmp_path / "x.txt")], capture_output=True, text=True)
    assert r.returncode == 1 and ".blz" in r.stdout
    r = subprocess.run([exe, str(tmp_path / "x.blz")], capture_output=True, text=True)
    assert r.returncode == 1 and "does not exist" in r.stdout
    r = subprocess.run([exe, "a.blz", "b.blz"], capture_output=True, text=True)
    assert r.returncode == 1 and "single scene" in r.stdout


def test_oracle_is_not_linked_into_the_product():
    """The product library never references the oracle."""
    out = subprocess.run(["ldd", B.LIB_PATH], capture_output=True, text=True).stdout
    assert "oracle" not in out
    data = open(B.LIB_PATH, "rb").read()
    assert b"orc_" not in data
