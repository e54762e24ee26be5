"""GPU parity: every solver stage and whole frames of the HIP path (through
the C ABI) against the CPU oracle on the same seeded inputs.

The bar is bit-exact FP32 equality: both sides evaluate the same formulas in
the same order with IEEE division/sqrt and no FMA contraction, and every
reference race (float atomics, inter-block Gauss-Seidel reads, boundary
copies) is replaced by one fixed order on both sides.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(__file__))
from helpers import SCENES, compare, dims, random_state, settings  # noqa: E402

pytestmark = pytest.mark.gpu

B = pytest.importorskip("blaze_amd")
O = pytest.importorskip("oracle")


def pair(s, state):
    g = B.Solver(s)
    o = O.Oracle(s)
    for w, a in state.items():
        g.set_field(w, a)
        o.set_field(w, a)
    return g, o


def test_field_roundtrip():
    s = settings(B, (16, 12, 10), smoke_density=1.0)
    st = random_state(s, 1, smoke=True)
    g = B.Solver(s)
    for w, a in st.items():
        g.set_field(w, a)
        compare(f"roundtrip {w}", g.field(w), a)


STAGE_CASES = [
    ("cooldown", B.STAGE_COOLDOWN, dict(cooling_rate=0.9)),
    ("drag", B.STAGE_DRAG, dict(drag=0.5)),
    ("vorticity", B.STAGE_VORTICITY, dict(vorticity_conf=0.5)),
    ("buoyancy", B.STAGE_BUOYANCY, dict(buoyancy=0.7)),
    ("wind_x", B.STAGE_WIND, dict(wind_dir=0, wind_amp=0.2, wind_speed=0.2, wind_turb_amp=7.5,
                                  wind_turb_scale=0.125)),
    ("wind_xneg", B.STAGE_WIND, dict(wind_dir=1, wind_amp=0.3, wind_speed=0.4, wind_turb_amp=2.0,
                                     wind_turb_scale=0.5)),
    ("wind_z", B.STAGE_WIND, dict(wind_dir=2, wind_amp=0.2, wind_speed=0.2, wind_turb_amp=7.5,
                                  wind_turb_scale=0.125)),
    ("wind_zneg", B.STAGE_WIND, dict(wind_dir=3, wind_amp=0.6, wind_speed=0.4, wind_turb_amp=3.75,
                                     wind_turb_scale=0.625)),
    ("turbulence", B.STAGE_TURBULENCE, dict(turbulence_amp=0.5, turbulence_scale=0.25, noise_seed=7,
                                            turb_mask_temp_ramp=(0.0, 2.0), turb_mask_vel_ramp=(0.0, 3.0))),
    ("turbulence_coarse", B.STAGE_TURBULENCE, dict(turbulence_amp=2.0, turbulence_scale=2.0,
                                                   turb_mask_temp_ramp=(0.03, 0.14),
                                                   turb_mask_vel_ramp=(0.0, 0.06))),
    ("divergence_closed", B.STAGE_DIVERGENCE, dict(closed_bounds=(1, 1, 1, 1, 1, 1))),
    ("divergence_mixed", B.STAGE_DIVERGENCE, dict(closed_bounds=(0, 1, 1, 0, 0, 1))),
    ("pressure_fixed", B.STAGE_PRESSURE, dict(max_iter_solve=7)),
    ("pressure_tol", B.STAGE_PRESSURE, dict(max_iter_solve=60, residual_tol=5.0, residual_check_every=7)),
    ("update_closed", B.STAGE_PROJECT_UPDATE, dict(closed_bounds=(1, 1, 1, 1, 1, 1))),
    ("update_open", B.STAGE_PROJECT_UPDATE, dict(closed_bounds=(0, 0, 0, 0, 0, 0))),
    ("update_mixed", B.STAGE_PROJECT_UPDATE, dict(closed_bounds=(1, 0, 1, 0, 0, 1))),
    ("advect", B.STAGE_ADVECT, dict()),
    ("advect_smoke", B.STAGE_ADVECT, dict(smoke_density=1.0)),
]


@pytest.mark.parametrize("name,stage,kw", STAGE_CASES, ids=[c[0] for c in STAGE_CASES])
@pytest.mark.parametrize("res", [(32, 24, 20), (16, 16, 16), (24, 20, 12)], ids=["32x24x20", "16c", "24x20x12"])
def test_stage_bitwise(name, stage, kw, res):
    """dx = 1/W: a power of two for the first two shapes (grid scaling by
    multiplication), not for the third (IEEE division)."""
    print(f"[stage] {name} res={res}", flush=True)
    s = settings(B, res, **kw)
    st = random_state(s, seed=sum(name.encode()), vel_scale=0.7, smoke=s.smoke_density > 0)
    g, o = pair(s, st)
    time, dt = 1.375, 1.0 / 24.0
    g.run_stage(stage, time, dt)
    o.run_stage(stage, time, dt)
    which = [0, 1, 2, 3, 5, 6] + ([4] if s.smoke_density > 0 else [])
    for w in which:
        compare(f"{name}:{['T', 'U', 'V', 'W', 'density', 'P', 'rhs'][w]}", g.field(w), o.field(w))
    if stage == B.STAGE_PRESSURE:
        gs, os_ = g.stats(), o.stats()
        assert gs.rbgs_iters_last == os_["rbgs_iters_last"]
        if s.residual_tol > 0:
            assert gs.residual == os_["residual"]
            assert gs.residual_initial == os_["residual_initial"]


RBGS_SHAPES = [(32, 24, 20), (64, 8, 40), (16, 10, 21), (32, 16, 64), (130, 6, 9), (128, 5, 8), (66, 4, 5),
               (60, 7, 6)]


# every variant the autotuner (Solver::autotune) can pick
ALL_VARIANTS = [2, 4, 8, 102, 103, 202, 203, 402, 403, 1103, 1104, 1105, 1106, 1203, 1204, 1403, 1404]


@pytest.mark.parametrize("variant", ALL_VARIANTS)
@pytest.mark.parametrize("res", RBGS_SHAPES, ids=["x".join(map(str, r)) for r in RBGS_SHAPES])
@pytest.mark.parametrize("iters", [9, 8])
def test_rbgs_variants_bitwise(variant, res, iters):
    """Every RBGS kernel variant the autotuner can pick gives the oracle's
    sweeps bit for bit on every shape: the two-pass in-place kernels (2/4/8
    slab groups in flight), the fused ping-pong sweeps (1x2/1x3 with 1, 2 or
    4 waves per block; an odd count ends with a two-pass sweep) and the
    row-pair fused sweeps (1x03-1x06: 3 to 6 groups in flight, 1, 2 and 4
    waves per block). Widths 130, 128 and 66 put the
    fused kernel's row over several waves, with a partly filled last wave;
    odd heights (5, 7 rows) leave the row-pair kernel a pair with one row,
    and width 60 fills its 30-cell segments exactly."""
    print(f"[rbgs] v{variant} res={res} iters={iters}", flush=True)
    s = settings(B, res, max_iter_solve=iters)
    st = random_state(s, seed=variant + iters)
    g, o = pair(s, st)
    g.set_rbgs_variant(variant)
    g.run_stage(B.STAGE_PRESSURE, 0.0, 1.0 / 24.0)
    o.run_stage(B.STAGE_PRESSURE, 0.0, 1.0 / 24.0)
    compare(f"rbgs v{variant}", g.field(B.FIELD_P), o.field(B.FIELD_P))


@pytest.mark.parametrize("variant", [0, 2, 4, 8, 102, 103, 202, 203, 402, 403, 1103, 1104, 1105, 1106, 1203, 1204,
                                     1403, 1404])
def test_rbgs_large_bitwise(variant):
    """128^3 (32 slab groups per column, 64 MB of planes read from HBM
    rather than cache, so loads are in flight for long): a residual-driven
    solve of 50 sweeps with five checks equals the oracle bit for bit in
    pressure, sweep count and residuals, for every variant the autotuner can
    pick; variant 0: its choice for this grid."""
    print(f"[rbgs-large] v{variant} 128^3", flush=True)
    s = settings(B, (128, 128, 128), closed_bounds=(1, 1, 1, 0, 1, 1), max_iter_solve=50, residual_tol=1e-30,
                 residual_check_every=10)
    st = random_state(s, seed=variant + 5)
    g, o = pair(s, st)
    if variant:
        g.set_rbgs_variant(variant)
    g.run_stage(B.STAGE_PRESSURE, 0.0, 1.0 / 24.0)
    o.run_stage(B.STAGE_PRESSURE, 0.0, 1.0 / 24.0)
    gs, os_ = g.stats(), o.stats()
    assert gs.rbgs_iters_last == os_["rbgs_iters_last"] == 50
    assert gs.residual_initial == os_["residual_initial"]
    assert gs.residual == os_["residual"]
    compare(f"rbgs-large v{variant}", g.field(B.FIELD_P), o.field(B.FIELD_P))


@pytest.mark.parametrize("variant", [8, 103, 1104, 1204])
@pytest.mark.parametrize("kind", ["tiny", "negzero", "upper"])
def test_rbgs_exact_division_fallback(variant, kind):
    """Numerators outside the fast division's exact range (|num| < 2^-125,
    and -0 for the fused kernels) take the exact path: still bit for bit.
    "upper": small normal values below slab 20 and subnormal ones above, so
    the column's chain decays into the exact range part-way up and the
    waves switch to the exact run mid-column (loads in flight, carried k-1
    values and the restart group all exercised)."""
    print(f"[rbgs-exact] v{variant} {kind}", flush=True)
    s = settings(B, (64, 12, 40 if kind == "upper" else 10), max_iter_solve=6)
    st = random_state(s, seed=11)
    if kind == "tiny":
        st[B.FIELD_P] = (st[B.FIELD_P] * np.float32(1e-38)).astype(np.float32)
        st[B.FIELD_RHS] = (st[B.FIELD_RHS] * np.float32(3e-38)).astype(np.float32)
    elif kind == "upper":
        for f in (B.FIELD_P, B.FIELD_RHS):
            a = st[f].astype(np.float64)
            a[:20] *= 1e-30
            a[20:] *= 1e-38
            st[f] = a.astype(np.float32)
    else:
        st[B.FIELD_P] = np.full_like(st[B.FIELD_P], -0.0)
        st[B.FIELD_RHS] = np.full_like(st[B.FIELD_RHS], -0.0)
        st[B.FIELD_RHS][3, 5, 7] = np.float32(1e-3)
    g, o = pair(s, st)
    g.set_rbgs_variant(variant)
    g.run_stage(B.STAGE_PRESSURE, 0.0, 1.0 / 24.0)
    o.run_stage(B.STAGE_PRESSURE, 0.0, 1.0 / 24.0)
    compare(f"rbgs-exact v{variant} {kind}", g.field(B.FIELD_P), o.field(B.FIELD_P))


@pytest.mark.parametrize("d", [3, 4, 5, 6])
def test_division_exhaustive(d):
    """The pressure solve's device division by a neighbour count equals IEEE
    division for all 2^32 float inputs."""
    assert B.selftest_division(d) == 0


def test_residual_matches_oracle():
    """The FP64 residual (max over cells, open-face cells skipped) equals the
    oracle's value exactly, for several boundary combinations."""
    for cb in [(1, 1, 1, 1, 1, 1), (0, 1, 1, 0, 1, 0), (0, 0, 0, 0, 0, 0)]:
        s = settings(B, (32, 24, 20), closed_bounds=cb, max_iter_solve=10, residual_tol=1e-30,
                     residual_check_every=10)
        st = random_state(s, seed=sum(cb))
        g, o = pair(s, st)
        g.run_stage(B.STAGE_PRESSURE, 0.0, 1.0 / 24.0)
        o.run_stage(B.STAGE_PRESSURE, 0.0, 1.0 / 24.0)
        gs, os_ = g.stats(), o.stats()
        assert gs.residual_initial == os_["residual_initial"], cb
        assert gs.residual == os_["residual"], cb


@pytest.mark.parametrize("case", ["tolerance", "fixed-point"])
def test_adaptive_checks_bitwise(case):
    """residual_check_adaptive on the GPU (checks skipped on the device until
    the interval the contraction sets, the FP32 fixed-point stop) against
    the oracle's rule: same sweep count, same final residual, same pressure
    bit for bit. "tolerance": a 64x48x40 solve to 1e-4; "fixed-point": an
    unreachable tolerance on 8^3, which ends at the fixed point."""
    res, tol, cap = ((64, 48, 40), 1e-4, 6000) if case == "tolerance" else ((8, 8, 8), 1e-30, 200000)
    print(f"[adaptive] {case} {res}", flush=True)
    s = settings(B, res, closed_bounds=(1, 1, 1, 1, 1, 1), max_iter_solve=cap, residual_tol=tol,
                 residual_check_every=10, residual_check_adaptive=1)
    W, H, D = res
    rng = np.random.default_rng(21)
    rhs = rng.standard_normal((D, H, W)).astype(np.float32)
    rhs -= np.float32(rhs.mean())
    g, o = B.Solver(s), O.Oracle(s)
    for x in (g, o):
        x.set_field(B.FIELD_RHS, rhs)
        x.set_field(B.FIELD_P, np.zeros_like(rhs))
    g.set_residual_trace(4096)
    g.run_stage(B.STAGE_PRESSURE, 0.0, 1.0 / 24.0)
    o.run_stage(B.STAGE_PRESSURE, 0.0, 1.0 / 24.0)
    gs, os_ = g.stats(), o.stats()
    checks = len(g.residual_trace(0))
    print(f"[adaptive] sweeps {gs.rbgs_iters_last} (oracle {os_['rbgs_iters_last']}), {checks} checks, "
          f"residual {gs.residual:.3e}", flush=True)
    assert gs.rbgs_iters_last == os_["rbgs_iters_last"] < cap
    assert gs.residual == os_["residual"]
    assert checks < gs.rbgs_iters_last // 10
    compare(f"adaptive {case} P", g.field(B.FIELD_P), o.field(B.FIELD_P))


def random_particles(n, seed, lo=-0.1, hi=1.1):
    rng = np.random.default_rng(seed)
    p = np.zeros((n, 8), np.float32)
    p[:, 0:3] = rng.uniform(lo, hi, (n, 3))
    p[:, 3:6] = rng.standard_normal((n, 3))
    p[:, 6] = rng.uniform(0.0, 0.15, n)      # some below dx/2 are skipped
    p[:, 7] = rng.uniform(-0.5, 2.0, n)
    p[::7, 3] = 1e-6                         # below the 1e-5 channel threshold
    p[::11, 7] = 0.0
    return p


@pytest.mark.parametrize("smoke", [False, True])
def test_p2g_bitwise(smoke):
    s = settings(B, (32, 24, 20), smoke_density=0.8 if smoke else 0.0)
    st = random_state(s, seed=3, smoke=smoke)
    g, o = pair(s, st)
    parts = random_particles(3000, 11)
    g.add_source(parts)
    o.add_source(parts)
    for w in [0, 1, 2, 3] + ([4] if smoke else []):
        compare(f"p2g:{w}", g.field(w), o.field(w))
    # re-splat of the uploaded set on a changed state
    st2 = random_state(s, seed=4, smoke=smoke)
    for w, a in st2.items():
        g.set_field(w, a)
        o.set_field(w, a)
    g.add_source(None)
    o.add_source(None)
    for w in [0, 1, 2, 3] + ([4] if smoke else []):
        compare(f"p2g-resplat:{w}", g.field(w), o.field(w))


def test_p2g_dense_hotspot():
    """Many particles per cell (the contention case of the reference)."""
    s = settings(B, (32, 32, 32))
    g, o = pair(s, {})
    gen = B.ParticleGen(count=50000, seed=5, radius=0.1, pscale=2.0 / 32, temp=5.0, radial_vel=2.0)
    for a, v in enumerate((0.5, 0.3, 0.5)):
        gen.center[a] = v
    parts = B.generate_particles(gen)
    g.add_source(parts)
    o.add_source(parts)
    for w in (0, 1, 2, 3):
        compare(f"hotspot:{w}", g.field(w), o.field(w))


def scene_settings(name, res=None, **kw):
    s, _, _, gen = B.parse_scene(os.path.join(SCENES, name))
    if res is not None:
        W, H, D = res
        s.grid_res[0], s.grid_res[1], s.grid_res[2] = W, H, D
        s.dx = 1.0 / W
        s.domain_max[0], s.domain_max[1], s.domain_max[2] = 1.0, H / W, D / W
    for k, v in kw.items():
        setattr(s, k, v)
    return s, gen


def run_frames(s, gen, frames, source_frames, check_every=1, fields=(0, 1, 2, 3, 4, 5)):
    parts = B.generate_particles(gen)
    g = B.Solver(s)
    o = O.Oracle(s)
    t = np.float32(0.0)
    for it in range(1, frames + 1):
        if it <= source_frames:
            g.add_source(parts if it == 1 else None)
            o.add_source(parts if it == 1 else None)
        g.step(float(t))
        o.step(float(t))
        gs, os_ = g.stats(), o.stats()
        assert gs.substeps == os_["substeps"], f"frame {it}: substeps {gs.substeps} vs {os_['substeps']}"
        assert gs.rbgs_iters == os_["rbgs_iters"], f"frame {it}: iters {gs.rbgs_iters} vs {os_['rbgs_iters']}"
        if it % check_every == 0 or it == frames:
            for w in fields:
                if w == 4 and s.smoke_density <= 0:
                    continue
                compare(f"frame {it} field {w}", g.field(w), o.field(w))
        t = np.float32(t + np.float32(s.dt))
    return g, o


def test_frames_plume64_bitwise():
    """64^3 plume (solid walls, open top, buoyancy, smoke density, CFL
    substeps): 12 frames bit-exact."""
    s, gen = scene_settings("plume64.blz")
    print("[frames] plume64", flush=True)
    run_frames(s, gen, 12, 12, check_every=3)


def test_frames_plume64_all_50_frames_bitwise():
    """The 64^3 plume for its whole length: 50 frames (20k particles
    re-splatted every frame, 100 sweeps per solve, CFL substeps), sweep and
    substep counts equal every frame, every field bit-exact every 10th."""
    s, gen = scene_settings("plume64.blz")
    print("[frames] plume64, 50 frames", flush=True)
    run_frames(s, gen, 50, 50, check_every=10)


def test_frames_plume128_features_bitwise():
    """Every force of the 128^3 plume (vorticity, curl noise, wind, open x
    sides and top) on a 48^3 grid, 8 frames bit-exact."""
    s, gen = scene_settings("plume128.blz", res=(48, 48, 48))
    gen.pscale = 0.75 / 48
    gen.count = 20000
    print("[frames] plume128 features @48^3", flush=True)
    run_frames(s, gen, 8, 8, check_every=2)


def test_frames_plume128_full_size_bitwise():
    """The 128^3 plume at its full size with every force (vorticity 0.3,
    curl-noise turbulence 0.5 at frequency 4 with seed 7, wind 0.2 with
    noise on the x boundary, open x sides and top), 100k particles
    re-splatted every frame, smoke density, CFL substeps and 200 fixed
    sweeps: two frames bit-exact against the oracle."""
    s, gen = scene_settings("plume128.blz")
    assert tuple(s.grid_res) == (128, 128, 128) and gen.count == 100000 and s.max_iter_solve == 200
    print("[frames] plume128 at full size, 2 frames", flush=True)
    run_frames(s, gen, 2, 2, check_every=1)


def test_frames_explosion_tol_bitwise():
    """Explosion features (radial source once, cooling, vorticity, residual-
    driven RBGS, CFL substeps, closed box) on a 40^3 grid."""
    s, gen = scene_settings("explosion256.blz", res=(40, 40, 40), residual_tol=2e-3, max_iter_solve=120)
    gen.count = 20000
    gen.pscale = 0.6 / 40 * 2
    print("[frames] explosion @40^3", flush=True)
    run_frames(s, gen, 6, 1, check_every=2)


def test_frames_explosion256_full_size_bitwise():
    """The benchmark's scene at its full size: 256^3, 1M particles, cooling,
    vorticity, buoyancy, the autotuned sweep kernel with residual checks, two
    frames of one substep each (CFL substepping off and 20 sweeps per solve
    keep the oracle to seconds; the first solve starts from p = 0, so its
    early groups take the exact division). Every field bit-exact."""
    s, gen = scene_settings("explosion256.blz", max_iter_solve=20, residual_tol=1e-4, cfl_number=0.0)
    assert tuple(s.grid_res) == (256, 256, 256) and gen.count == 1000000
    print("[frames] explosion256 at full size, 2 frames", flush=True)
    run_frames(s, gen, 2, 1, check_every=1, fields=(0, 1, 2, 3, 5))


@pytest.mark.parametrize("res", [(4, 4, 4), (6, 4, 4), (4, 5, 4), (6, 6, 5), (8, 4, 7)],
                         ids=["4c", "6x4x4", "4x5x4", "6x6x5", "8x4x7"])
def test_small_grid_autotuned_solve_bitwise(res):
    """Grids whose padded pressure planes outgrow the cell-centred scratch
    (2 * Np > 7 * W*H*D): the autotuner, which runs at the first solve and
    saves the pressure planes while it times its candidates, restores them
    intact, and the residual-driven solve equals the oracle bit for bit."""
    print(f"[small-grid autotune] {res}", flush=True)
    s = settings(B, res, closed_bounds=(1, 1, 1, 0, 1, 1), max_iter_solve=24, residual_tol=1e-30,
                 residual_check_every=6)
    st = random_state(s, seed=sum(res))
    g, o = pair(s, st)
    g.run_stage(B.STAGE_PRESSURE, 0.0, 1.0 / 24.0)
    o.run_stage(B.STAGE_PRESSURE, 0.0, 1.0 / 24.0)
    assert g.stats().rbgs_iters_last == o.stats()["rbgs_iters_last"] == 24
    compare(f"small {res}", g.field(B.FIELD_P), o.field(B.FIELD_P))
    # a whole frame through the drop-in path as well
    g2, o2 = pair(s, st)
    g2.step(0.0)
    o2.step(0.0)
    for w in (0, 1, 2, 3, 5):
        compare(f"small {res} frame field {w}", g2.field(w), o2.field(w))


def test_determinism_repeat():
    """Two GPU runs of the same frames are bitwise identical."""
    s, gen = scene_settings("plume128.blz", res=(32, 32, 32))
    gen.count = 5000
    outs = []
    for _ in range(2):
        parts = B.generate_particles(gen)
        g = B.Solver(s)
        t = np.float32(0.0)
        for it in range(1, 5):
            g.add_source(parts if it == 1 else None)
            g.step(float(t))
            t = np.float32(t + np.float32(s.dt))
        outs.append([g.field(w) for w in (0, 1, 2, 3, 5)])
    for a, b in zip(*outs):
        compare("repeat", a, b)
