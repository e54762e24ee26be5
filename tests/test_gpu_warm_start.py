"""GPU tests of the learned warm start (blaze_amd/csrc/k_warm.hip): RBGS
stays the judge and the fallback, and a trained network lowers the sweeps
needed to reach the same residual tolerance (the ablation against the
solve without it)."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(__file__))
from helpers import SCENES, compare, random_state, settings  # noqa: E402

pytestmark = pytest.mark.gpu

B = pytest.importorskip("blaze_amd")
O = pytest.importorskip("oracle")
W = pytest.importorskip("blaze_amd.warm_start")


def test_rejected_prediction_gives_the_plain_result_bitwise():
    """A destructive prediction raises the FP64 residual, is rejected, the
    saved pressure is restored, and the sweeps then produce the oracle's
    plain result bit for bit."""
    print("[ws] rejected prediction", flush=True)
    s = settings(B, (32, 24, 20), max_iter_solve=40, residual_tol=1e-7, residual_check_every=10, warm_start=1)
    st = random_state(s, seed=3)
    g, o = B.Solver(s), O.Oracle(s)
    for w, a in st.items():
        g.set_field(w, a)
        o.set_field(w, a)
    net = B.Net(seed=1)
    net.set_params(net.params() * np.float32(40.0))
    g.set_warm_start(net)
    g.run_stage(B.STAGE_PRESSURE, 0.0, 1.0 / 24.0)
    o.run_stage(B.STAGE_PRESSURE, 0.0, 1.0 / 24.0)
    compare("ws-rejected P", g.field(B.FIELD_P), o.field(B.FIELD_P))
    gs, os_ = g.stats(), o.stats()
    assert gs.warm_start_used == 0 and gs.warm_start_fallbacks == 1
    assert gs.rbgs_iters_last == os_["rbgs_iters_last"]
    assert gs.residual == os_["residual"] and gs.residual_initial == os_["residual_initial"]


def test_detached_or_disabled_warm_start_is_the_plain_path():
    s = settings(B, (32, 24, 20), max_iter_solve=30, residual_tol=1e-7, residual_check_every=10, warm_start=0)
    st = random_state(s, seed=4)
    g, o = B.Solver(s), O.Oracle(s)
    for w, a in st.items():
        g.set_field(w, a)
        o.set_field(w, a)
    g.set_warm_start(B.Net(seed=2))   # attached but warm_start == 0
    g.run_stage(B.STAGE_PRESSURE, 0.0, 1.0 / 24.0)
    o.run_stage(B.STAGE_PRESSURE, 0.0, 1.0 / 24.0)
    compare("ws-disabled P", g.field(B.FIELD_P), o.field(B.FIELD_P))
    assert g.stats().residual_checks == o.stats()["residual_checks"]


def run_frames(scene, res, frames, tol, cap, net, start=0):
    """Frames `start`..`start+frames-1` of `scene` at grid `res`, stage by
    stage; returns the sweeps of each frame's solve and the kept/rejected
    counts."""
    s, _, _, gen = B.parse_scene(os.path.join(SCENES, scene))
    f = s.grid_res[0] / res[0]
    for a in range(3):
        s.grid_res[a] = res[a]
    s.dx = s.dx * f
    gen.pscale = gen.pscale * f
    s.cfl_number = 0.0
    s.residual_tol = tol
    s.max_iter_solve = cap
    s.residual_check_every = 10
    s.warm_start = 1 if net is not None else 0
    parts = B.generate_particles(gen)
    g = B.Solver(s)
    if net is not None:
        g.set_warm_start(net)
    dt = float(s.dt)
    t = np.float32(0.0)
    sweeps, kept = [], 0
    for fr in range(start + frames):
        g.add_source(parts)
        for stg in W._force_stages(s):
            g.run_stage(stg, float(t), dt)
        g.run_stage(B.STAGE_DIVERGENCE, float(t), dt)
        g.run_stage(B.STAGE_PRESSURE, float(t), dt)
        st = g.stats()
        if fr >= start:
            sweeps.append(st.rbgs_iters_last)
            kept += st.warm_start_used
        g.run_stage(B.STAGE_PROJECT_UPDATE, float(t), dt)
        g.run_stage(B.STAGE_ADVECT, float(t), dt)
        t = np.float32(t + np.float32(dt))
    return np.array(sweeps), kept, g.stats().warm_start_fallbacks


def test_trained_network_reduces_sweeps():
    """Ablation: a network trained on pairs from the first frames of a plume
    run (the solver's own data, 32^3) is used on later frames; the sweeps
    to reach the same FP64 residual tolerance are compared with the plain
    solve of the same frames."""
    print("[ws] ablation: data", flush=True)
    pairs = W.scene_pairs(os.path.join(SCENES, "plume64.blz"), 40, res=(32, 32, 32), sweeps=1500, every=1)
    net = B.Net(seed=0)
    print(f"[ws] {len(pairs)} pairs; untrained error {W.relative_error(net, pairs[-5:]):.3f}", flush=True)
    W.train(net, W.augment(pairs), steps=400, batch=4, lr=2e-3)
    err = W.relative_error(net, pairs[-5:])
    print(f"[ws] trained error {err:.3f}", flush=True)
    plain, _, _ = run_frames("plume64.blz", (32, 32, 32), 20, 2e-4, 600, None, start=40)
    warm, kept, rej = run_frames("plume64.blz", (32, 32, 32), 20, 2e-4, 600, net, start=40)
    print(f"[ws] plain sweeps {plain.sum()} {plain.tolist()}", flush=True)
    print(f"[ws] warm  sweeps {warm.sum()} {warm.tolist()} kept {kept} rejected {rej}", flush=True)
    assert err < 0.9
    assert kept > 0
    assert warm.sum() < plain.sum()
