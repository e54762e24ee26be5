"""GPU tests of batched scenes (blaze_amd/csrc/batch.hip): every scene of a
batch, stepped by captured-graph replays with multi-grid sweep and residual
launches, equals the same scene stepped alone (and hence the oracle) bit for
bit, before and after the autotuner changes the tile configuration, over
enough steps to replay both captured graphs."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(__file__))
from helpers import compare, settings  # noqa: E402

pytestmark = pytest.mark.gpu

B = pytest.importorskip("blaze_amd")

SHAPES = [(32, 24, 16), (48, 40, 32), (64, 32, 24), (40, 16, 20)]


def scene(res, seed):
    rng = np.random.default_rng(seed)
    s = settings(B, res, max_iter_solve=40, residual_tol=float(rng.uniform(1e-4, 5e-4)), residual_check_every=10,
                 buoyancy=float(rng.uniform(0.5, 2.0)), vorticity_conf=0.3, cooling_rate=0.5,
                 closed_bounds=(1, 1, 1, 0, 1, 1), wind_amp=0.2, wind_speed=0.2, wind_turb_amp=2.0,
                 wind_turb_scale=0.25)
    n = 4000
    W, H, D = res
    c = np.array([0.5 * W, 0.3 * H, 0.5 * D]) * s.dx
    r = float(rng.uniform(0.03, 0.1)) * W * s.dx
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    pos = c + d * (r * rng.random((n, 1)) ** (1 / 3))
    part = np.zeros((n, 8), np.float32)
    part[:, :3] = pos
    part[:, 6] = 2.5 * s.dx
    part[:, 7] = 1.0
    return s, part


@pytest.mark.parametrize("tune", [False, True, 1104, 1203])
def test_batch_equals_single_scenes(tune):
    """tune: False = the default variant, True = the autotuner's choice, a
    number = that row-pair variant (multi-grid row-pair sweep)."""
    print(f"[batch] tune={tune}", flush=True)
    cases = [scene(r, i) for i, r in enumerate(SHAPES)]
    b = B.Batch([c[0] for c in cases])
    singles = [B.Solver(c[0]) for c in cases]
    for i, (s, part) in enumerate(cases):
        b.scenes[i].add_source(part)
        singles[i].add_source(part)
    if tune not in (False, True):
        b.set_variant(tune)
        assert b.stats().variant == tune
    elif tune:
        v = b.autotune(10)
        st = b.stats()
        print(f"[batch] autotuned variant {v}: " + ", ".join(
            f"{st.tune_variant[k]}={st.tune_ms[k]:.4f}ms" for k in range(B.BATCH_TUNE)), flush=True)
    t = np.float32(0.0)
    for step in range(3):
        b.step(float(t))
        for g in singles:
            g.step(float(t))
        t = np.float32(t + np.float32(cases[0][0].dt))
        for i in range(len(cases)):
            gs, ss = b.scenes[i].stats(), singles[i].stats()
            assert gs.rbgs_iters == ss.rbgs_iters, (step, i)
            assert gs.residual == ss.residual, (step, i)
    for i in range(len(cases)):
        for w in (B.FIELD_T, B.FIELD_U, B.FIELD_V, B.FIELD_W, B.FIELD_P):
            compare(f"batch scene {i} field {w}", b.scenes[i].field(w), singles[i].field(w))
    st = b.stats()
    print(f"[batch] nodes {st.graph_nodes} capture {st.capture_ms:.1f} ms step {st.step_ms:.2f} ms", flush=True)
    assert st.graph_nodes > 0


@pytest.mark.parametrize("variant", [None, 203, 1104, 1204])
def test_batch_large_scenes_equal_single_scenes(variant):
    """Grids of 128^3 and 96x80x112 (loads from HBM, 28-32 slab groups per
    column) through the multi-grid sweep and residual kernels: every scene
    equals the same scene stepped alone bit for bit. variant None: the
    default multi-grid sweep."""
    print(f"[batch-large] variant {variant}", flush=True)
    cases = [scene(r, 20 + i) for i, r in enumerate([(128, 128, 128), (96, 80, 112)])]
    b = B.Batch([c[0] for c in cases])
    singles = [B.Solver(c[0]) for c in cases]
    for i, (s, part) in enumerate(cases):
        b.scenes[i].add_source(part)
        singles[i].add_source(part)
    if variant is not None:
        b.set_variant(variant)
    t = np.float32(0.0)
    for step in range(3):
        b.step(float(t))
        for g in singles:
            g.step(float(t))
        t = np.float32(t + np.float32(cases[0][0].dt))
        for i in range(len(cases)):
            gs, ss = b.scenes[i].stats(), singles[i].stats()
            assert gs.rbgs_iters == ss.rbgs_iters, (step, i)
            assert gs.residual == ss.residual, (step, i)
    for i in range(len(cases)):
        for w in (B.FIELD_T, B.FIELD_U, B.FIELD_V, B.FIELD_W, B.FIELD_P):
            compare(f"batch-large scene {i} field {w}", b.scenes[i].field(w), singles[i].field(w))


def test_batch_mixed_stages_equal_single_scenes():
    """Scenes that run different stage sets in one batch: smoke density
    (temperature and density advected as a pair), no buoyancy, drag and
    curl-noise turbulence, a power-of-two cell size (the multiplying
    position scale), adaptive check spacing with the fixed-point stop. The
    stages with table launches (buoyancy, divergence, update, advection)
    then run over a subset of the scenes or split by kind, the others per
    scene; every scene still equals the same scene stepped alone, bit for
    bit, with the same sweep counts and residuals."""
    print("[batch-mixed] 4 scenes", flush=True)
    cases = [scene(r, 50 + i) for i, r in enumerate([(32, 24, 16), (48, 40, 32), (64, 32, 24), (40, 16, 20)])]
    cases[0][0].smoke_density = 1.0
    cases[1][0].buoyancy = 0.0
    cases[1][0].drag = 0.5
    cases[1][0].turbulence_amp = 0.5
    cases[1][0].turbulence_scale = 0.3
    cases[2][0].residual_check_adaptive = 1
    cases[2][0].smoke_density = 0.5
    cases[3][0].vorticity_conf = 0.0
    assert cases[0][0].dx == 1.0 / 32 and cases[2][0].dx == 1.0 / 64  # powers of two
    b = B.Batch([c[0] for c in cases])
    singles = [B.Solver(c[0]) for c in cases]
    for i, (s, part) in enumerate(cases):
        b.scenes[i].add_source(part)
        singles[i].add_source(part)
    t = np.float32(0.0)
    for step in range(3):
        b.step(float(t))
        for g in singles:
            g.step(float(t))
        t = np.float32(t + np.float32(cases[0][0].dt))
        for i in range(len(cases)):
            gs, bs = singles[i].stats(), b.scenes[i].stats()
            assert bs.rbgs_iters == gs.rbgs_iters and bs.residual == gs.residual, (step, i)
    for i in range(len(cases)):
        fields = [B.FIELD_T, B.FIELD_U, B.FIELD_V, B.FIELD_W, B.FIELD_P]
        if cases[i][0].smoke_density > 0:
            fields.append(B.FIELD_DENSITY)
        for w in fields:
            compare(f"batch-mixed scene {i} field {w}", b.scenes[i].field(w), singles[i].field(w))


def test_batch_rejects_unknown_variant():
    s, _ = scene((32, 24, 16), 0)
    b = B.Batch([s])
    with pytest.raises(B.BlazeError, match="not a batch sweep variant"):
        b.set_variant(8)


def test_batch_rejects_unsupported_settings():
    s, _ = scene((32, 24, 16), 0)
    s.cfl_number = 1.0
    with pytest.raises(B.BlazeError, match="cfl_number"):
        B.Batch([s])


def test_single_grid_autotune_keeps_results():
    """The autotuner leaves the pressure as it found it and picks a variant
    whose sweeps are still bit-exact."""
    O = pytest.importorskip("oracle")
    s = settings(B, (64, 40, 24), max_iter_solve=12)
    rng = np.random.default_rng(9)
    rhs = rng.standard_normal((24, 40, 64)).astype(np.float32)
    p = rng.standard_normal((24, 40, 64)).astype(np.float32)
    g, o = B.Solver(s), O.Oracle(s)
    for x in (g, o):
        x.set_field(B.FIELD_RHS, rhs)
        x.set_field(B.FIELD_P, p)
    v, times = g.autotune(10)
    print(f"[autotune] best {v}: {times}", flush=True)
    compare("autotune keeps p", g.field(B.FIELD_P), p)
    g.run_stage(B.STAGE_PRESSURE, 0.0, 1.0 / 24.0)
    o.run_stage(B.STAGE_PRESSURE, 0.0, 1.0 / 24.0)
    compare(f"autotuned v{v}", g.field(B.FIELD_P), o.field(B.FIELD_P))


@pytest.mark.parametrize("variant", [None, 1204])
def test_batch_with_warm_start_and_handle_steps(variant):
    """With the learned warm start inside the captured graph, each scene still
    equals the same scene stepped alone with the same network; afterwards a
    scene handle stepped on its own (outside the graph, wind noise from its
    time argument) stays equal to the single solver. variant: the batch's
    sweep variant (None: the default one-row fused sweep; 1204: row-pair,
    also used by each scene's warm-start judge)."""
    print(f"[batch] warm start variant {variant}", flush=True)
    net = B.Net(seed=4)
    cases = [scene(r, 10 + i) for i, r in enumerate(SHAPES[:3])]
    for s, _ in cases:
        s.warm_start = 20
    b = B.Batch([c[0] for c in cases])
    if variant is not None:
        b.set_variant(variant)
    b.set_warm_start(net)
    singles = [B.Solver(c[0]) for c in cases]
    for i, (s, part) in enumerate(cases):
        singles[i].set_warm_start(net)
        b.scenes[i].add_source(part)
        singles[i].add_source(part)
    t = np.float32(0.0)
    for step in range(3):
        b.step(float(t))
        for g in singles:
            g.step(float(t))
        t = np.float32(t + np.float32(cases[0][0].dt))
    kept = sum(b.scenes[i].stats().warm_start_used for i in range(len(cases)))
    rej = sum(b.scenes[i].stats().warm_start_fallbacks for i in range(len(cases)))
    print(f"[batch] warm start kept {kept} rejected {rej}", flush=True)
    for i in range(len(cases)):
        assert b.scenes[i].stats().warm_start_fallbacks == singles[i].stats().warm_start_fallbacks
        for w in (B.FIELD_T, B.FIELD_U, B.FIELD_V, B.FIELD_W, B.FIELD_P):
            compare(f"batch+ws scene {i} field {w}", b.scenes[i].field(w), singles[i].field(w))
    for i in range(len(cases)):
        b.scenes[i].step(float(t))
        singles[i].step(float(t))
        compare(f"handle step scene {i}", b.scenes[i].field(B.FIELD_U), singles[i].field(B.FIELD_U))


def graphs_launched(batch_end, graphs=5, per_graph=16):
    """Graphs a batch step launches when the last solve ends in batch
    `batch_end` of the solve loop (0: at the check before the sweeps): graph
    0 holds the batches up to per_graph, each later sweep graph per_graph
    more; the host launches graph 1 unchecked, then graph i only after
    graph i - 2 ended with some solve still running, and the last graph."""
    k = 0 if batch_end == 0 else (batch_end - 1) // per_graph
    return min(k + 2, graphs - 1) + 1


@pytest.mark.parametrize("warm", [False, True])
def test_batch_chunked_graphs_equal_single_scenes(warm):
    """A sweep cap of 600 (60 batches of 10 sweeps; with the learned
    correction before the sweeps and every 20, a cap of 1200 for as many
    batches) splits each captured step into graphs of 16 batches and a last
    one with the update and advection. With a tight tolerance every solve
    runs to the cap and the step launches all five graphs; with a loose one
    the solves end early and the step launches only the graphs up to one
    past the graph in which the last solve ended, then the last graph.
    Either way every scene equals the same scene stepped alone, bit for bit,
    over both captured parities (warm: re-applications and their judges on
    either side of the graph boundaries)."""
    net = B.Net(seed=4) if warm else None
    cap = 1200 if warm else 600
    runs = {}
    loose_tol = 2e-2 if warm else 0.3
    for tol in (1e-7, loose_tol):
        print(f"[batch-chunks] warm={warm} tol={tol:g}", flush=True)
        cases = [scene(r, 40 + i) for i, r in enumerate(SHAPES)]
        for s, _ in cases:
            s.max_iter_solve = cap
            s.residual_tol = tol
            s.closed_bounds[3] = 1
            s.warm_start = 20 if warm else 0
        b = B.Batch([c[0] for c in cases])
        singles = [B.Solver(c[0]) for c in cases]
        if warm:
            b.set_warm_start(net)
        for i, (s, part) in enumerate(cases):
            if warm:
                singles[i].set_warm_start(net)
            b.scenes[i].add_source(part)
            singles[i].add_source(part)
        t = np.float32(0.0)
        graphs, iters = [], []
        for step in range(3):
            b.step(float(t))
            for g in singles:
                g.step(float(t))
            t = np.float32(t + np.float32(cases[0][0].dt))
            graphs.append(b.stats().graphs_run)
            iters.append([b.scenes[i].stats().rbgs_iters for i in range(len(cases))])
            for i in range(len(cases)):
                gs, ss = b.scenes[i].stats(), singles[i].stats()
                assert gs.rbgs_iters == ss.rbgs_iters, (tol, step, i)
                assert gs.residual == ss.residual, (tol, step, i)
                assert gs.warm_start_fallbacks == ss.warm_start_fallbacks, (tol, step, i)
        for i in range(len(cases)):
            for w in (B.FIELD_T, B.FIELD_U, B.FIELD_V, B.FIELD_W, B.FIELD_P):
                compare(f"batch-chunks tol {tol:g} scene {i} field {w}", b.scenes[i].field(w), singles[i].field(w))
        print(f"[batch-chunks] graphs run per step {graphs}, sweeps per step {iters}", flush=True)
        runs[tol] = (graphs, iters)
    (g_tight, i_tight), (g_loose, i_loose) = runs[1e-7], runs[loose_tol]
    assert all(max(x) == cap for x in i_tight), i_tight
    assert g_tight == [5, 5, 5], g_tight
    for n, x in zip(g_loose, i_loose):
        if warm:
            # the judge's batch first, then 20 sweeps per batch of the loop
            end = -(-(max(x) - 10) // 20)
        else:
            end = max(x) // 10
        if not warm:
            assert n == graphs_launched(end), (n, x)
        assert n <= graphs_launched(end), (n, x)


def test_batch_scenes_equal_the_oracle():
    """The multi-grid sweep and residual kernels against the CPU oracle
    directly: three uneven scenes (wind noise, vorticity, open top,
    residual-driven RBGS) batched for two steps equal the oracle stepping
    each scene on its own, field by field, bit for bit, with the same sweep
    counts and final residuals."""
    O = pytest.importorskip("oracle")
    print("[batch-oracle] 3 scenes", flush=True)
    cases = [scene(r, 30 + i) for i, r in enumerate([(32, 24, 16), (40, 16, 20), (24, 24, 28)])]
    b = B.Batch([c[0] for c in cases])
    oracles = [O.Oracle(c[0]) for c in cases]
    for i, (s, part) in enumerate(cases):
        b.scenes[i].add_source(part)
        oracles[i].add_source(part)
    t = np.float32(0.0)
    for step in range(2):
        b.step(float(t))
        for o in oracles:
            o.step(float(t))
        t = np.float32(t + np.float32(cases[0][0].dt))
        for i in range(len(cases)):
            gs, os_ = b.scenes[i].stats(), oracles[i].stats()
            assert gs.rbgs_iters == os_["rbgs_iters"], (step, i)
            assert gs.residual == os_["residual"], (step, i)
    for i in range(len(cases)):
        for w in (B.FIELD_T, B.FIELD_U, B.FIELD_V, B.FIELD_W, B.FIELD_P):
            compare(f"batch-oracle scene {i} field {w}", b.scenes[i].field(w), oracles[i].field(w))


def test_batched_workload_64_scenes_equal_single_scenes():
    """The benchmark's batched workload itself (64 scenes of 32^3..128^3,
    seed 5, the shipped warm-start network inside the captured graph,
    autotuned tiles): after two steps every scene equals the same scene
    stepped alone with the same network, bit for bit."""
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench
    W = pytest.importorskip("blaze_amd.warm_start")
    print("[batch-64] building 64 scenes", flush=True)
    sets, parts = bench.batched_scenes(64, warm_start=True)
    net = W.load_net()
    b = B.Batch(sets)
    b.set_warm_start(net)
    for q in range(64):
        b.scenes[q].add_source(parts[q])
    v = b.autotune(10)
    print(f"[batch-64] {b.stats().voxels} voxels, variant {v}", flush=True)
    t = np.float32(0.0)
    for _ in range(2):
        b.step(float(t))
        t = np.float32(t + np.float32(1.0 / 24.0))
    for q in range(64):
        g = B.Solver(sets[q])
        g.set_warm_start(net)
        g.set_rbgs_variant(v)
        g.add_source(parts[q])
        t = np.float32(0.0)
        for _ in range(2):
            g.step(float(t))
            t = np.float32(t + np.float32(1.0 / 24.0))
        gs, bs = g.stats(), b.scenes[q].stats()
        assert bs.rbgs_iters == gs.rbgs_iters and bs.residual == gs.residual, q
        for w in (B.FIELD_T, B.FIELD_U, B.FIELD_V, B.FIELD_W, B.FIELD_P):
            compare(f"batch-64 scene {q} field {w}", b.scenes[q].field(w), g.field(w))
        g.close()
    print("[batch-64] all 64 scenes equal", flush=True)
