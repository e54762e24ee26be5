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
    """A prediction that cannot lower the FP64 residual (here a network whose
    output overflows, so the correction is empty) is rejected, the saved
    pressure is restored, and the sweeps then produce the oracle's plain
    result bit for bit."""
    print("[ws] rejected prediction", flush=True)
    s = settings(B, (32, 24, 20), max_iter_solve=40, residual_tol=1e-7, residual_check_every=10, warm_start=1)
    st = random_state(s, seed=3)
    g, o = B.Solver(s), O.Oracle(s)
    for w, a in st.items():
        g.set_field(w, a)
        o.set_field(w, a)
    net = B.Net(seed=1)
    net.set_params(net.params() * np.float32(1e30))
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
    g, o, plain = B.Solver(s), O.Oracle(s), B.Solver(s)
    for w, a in st.items():
        g.set_field(w, a)
        o.set_field(w, a)
        plain.set_field(w, a)
    g.set_warm_start(B.Net(seed=2))   # attached but warm_start == 0
    for x in (g, o, plain):
        x.run_stage(B.STAGE_PRESSURE, 0.0, 1.0 / 24.0)
    compare("ws-disabled P", g.field(B.FIELD_P), o.field(B.FIELD_P))
    assert g.stats().residual_checks == plain.stats().residual_checks
    assert g.stats().warm_start_used == 0 and g.stats().warm_start_fallbacks == 0


def run_frames(scene, res, frames, tol, cap, net, start=0, every_ws=1):
    """Frames `start`..`start+frames-1` of `scene` at grid `res`, stage by
    stage; returns the sweeps of each frame's solve, the kept/rejected
    counts of learned corrections and each frame's final FP64 residual."""
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
    s.warm_start = every_ws if net is not None else 0
    parts = B.generate_particles(gen)
    g = B.Solver(s)
    if net is not None:
        g.set_warm_start(net)
    dt = float(s.dt)
    t = np.float32(0.0)
    sweeps, kept, res = [], 0, []
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
            res.append(st.residual)
        g.run_stage(B.STAGE_PROJECT_UPDATE, float(t), dt)
        g.run_stage(B.STAGE_ADVECT, float(t), dt)
        t = np.float32(t + np.float32(dt))
    return np.array(sweeps), kept, g.stats().warm_start_fallbacks, np.array(res)


def test_trained_network_ablation():
    """Ablation of the learned part on the solver's own data (plume, 32^3):
    the network trained on pairs from the first frames predicts held-out
    corrections far better than the same network untrained, and a solve
    that applies it (every 50 sweeps, judged by the FP64 residual) ends no
    worse than the plain solve of the same frames at equal sweeps. The
    measured gains are printed: the final-residual ratio warm/plain and the
    kept/rejected counts."""
    print("[ws] ablation: data", flush=True)
    pairs = W.scene_pairs(os.path.join(SCENES, "plume64.blz"), 48, res=(32, 32, 32), sweeps=1500, every=1)
    train_pairs, held = pairs[:40], pairs[40:]
    net = B.Net(seed=0)
    e0 = W.relative_error(net, held)
    W.train(net, W.augment(train_pairs), steps=1000, batch=4, lr=3e-4)
    e1 = W.relative_error(net, held)
    print(f"[ws] held-out shape error: untrained {e0:.3f} trained {e1:.3f}", flush=True)
    assert e1 < 0.85 * e0
    plain = run_frames("plume64.blz", (32, 32, 32), 12, 1e-12, 300, None, start=40)
    warm = run_frames("plume64.blz", (32, 32, 32), 12, 1e-12, 300, net, start=40, every_ws=50)
    ratio = float(np.mean(warm[3]) / np.mean(plain[3]))
    print(f"[ws] final residual plain {np.mean(plain[3]):.3e} warm {np.mean(warm[3]):.3e} ratio {ratio:.3f} "
          f"kept {warm[1]} rejected {warm[2]}", flush=True)
    assert warm[1] + warm[2] > 0
    assert ratio <= 1.25
