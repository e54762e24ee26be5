#!/usr/bin/env python3
"""bench.py — headline benchmark of the MI355X solver step.

Workload: the 256^3 FP32 explosion (scenes/explosion256.blz): 1M particles
sourced once, cooling, vorticity confinement, buoyancy, RBGS pressure solve
to ||div u||_inf <= 1e-4 (cap 500 sweeps), CFL substeps. A bench "step" is one
frame (FluidSolver::step); every CFL substep counts as a solver step in the
voxel-steps/s metric. Warmup frames (including the sourcing frame) are
untimed; the K timed frames are bracketed by stream synchronisations.

Prints one JSON line with the metric, the roofline of the dominant kernel
(the RBGS sweep kernel; algorithmic 12 B/voxel per sweep: read p, read rhs,
write p -- one launch of the fused sweep, or two colour passes of 6 B/voxel
each -- timed with HIP events on the solver stream in a second pass over the
same frames, since the per-stage events cost about 2% of the step; value is
the pass without them) and a CPU baseline (the oracle on the host cores,
bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
RBGS_BYTES_PER_VOXEL_SWEEP = 12.0  # read p, read rhs, write p
RBGS_KERNELS = {1: ("k_rbgs_pass", "RBGS colour pass, two per sweep", "rbgs_pass_traffic.json"),
                2: ("k_rbgs_fused", "RBGS fused sweep, one per sweep", "rbgs_fused_traffic.json"),
                3: ("k_rbgs_pair", "RBGS row-pair fused sweep, one per sweep", "rbgs_pair_traffic.json")}
# compulsory bytes per voxel of the non-solve part of a substep of this
# workload: cooling 8, vorticity 24, buoyancy 12, divergence 16, update 28,
# advection 40 (T, U, V, W, density)
FIXED_BYTES_PER_VOXEL = 128.0
RESIDUAL_BYTES_PER_VOXEL = 8.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scene", default=os.path.join(ROOT, "scenes", "explosion256.blz"))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sweeps", type=int, default=6)
    ap.add_argument("--no-cpu-1thread", action="store_true", help="skip the one-thread CPU baseline")
    ap.add_argument("--rbgs-variant", type=int, default=0)
    ap.add_argument("--calib-copy", action="store_true",
                    help="first run a device copy of 256 MiB (reference kernel for HBM counter scaling)")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="only run RBGS sweeps (for hardware counter runs)")
    ap.add_argument("--probe-sweeps", type=int, default=20)
    ap.add_argument("--stream-probe", action="store_true",
                    help="time plain streaming kernels (torch) over the sweep's working set: four pressure/rhs-sized "
                         "planes read, two written")
    ap.add_argument("--sweep-scaling", action="store_true",
                    help="time every sweep variant on batches of 1 and 2 grids of 256^3 and 8 and 16 of 128^3 "
                         "(the batch autotuner's timings; one JSON line per batch)")
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--no-warm-start", action="store_true", help="skip the learned warm-start leg")
    ap.add_argument("--warm-every", type=int, default=40,
                    help="warm_start setting of that leg (1: before the sweeps, N >= 2: also every N sweeps)")
    ap.add_argument("--warm-parts", type=int, default=3,
                    help="warm-start corrections of the headline-settings leg: 1 network, 2 global, 3 both")
    ap.add_argument("--tol-steps", type=int, default=3,
                    help="timed frames of the at-tolerance legs (plain, global, network + global, exact coarse "
                         "+ global, all to the scene's tolerance with the sweep cap lifted)")
    ap.add_argument("--tol-cap", type=int, default=20000, help="sweep cap of the at-tolerance legs")
    ap.add_argument("--no-render", action="store_true", help="skip the render of the plain leg's final state")
    ap.add_argument("--render-reps", type=int, default=3)
    ap.add_argument("--render-cpu-rows", type=int, default=8,
                    help="rows of the CPU oracle's render sample (spread evenly over the image)")
    ap.add_argument("--animated", action="store_true",
                    help="instead: animated .part sources (a new file per frame), synchronous vs staged uploads")
    ap.add_argument("--batched", action="store_true",
                    help="instead: 64 scenes of 32^3..128^3 in one captured-graph step (warm start on)")
    ap.add_argument("--scenes", type=int, default=64)
    ap.add_argument("--batched-tol-cap", type=int, default=0,
                    help="with --batched: also run the scenes with closed tops, warm start on and off, to each "
                         "tolerance of --batched-tol under this sweep cap (0: skip)")
    ap.add_argument("--batched-tol", default="0.05",
                    help="comma-separated tolerances of those runs (1e-4 lies below the FP32 floor of "
                         "||div u||_inf that these scenes' solves reach with closed tops, 1e-3 to 2e-2)")
    ap.add_argument("--batched-tol-steps", type=int, default=2, help="timed steps of those runs")
    return ap.parse_args()


def log(msg, quiet=False):
    if not quiet:
        print(msg, file=sys.stderr, flush=True)


def calib_copy(quiet, mib=256):
    """Device-to-device copy of a known size through the HIP runtime (the
    reference kernel for scaling the HBM byte counters)."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    hip.hipMemset(a, 1, C.c_size_t(n))
    for _ in range(3):
        assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0   # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"[calib] copied 3 x {mib} MiB device to device", quiet)


def traffic_probe(a):
    """RBGS sweeps alone on the scene's grid (random smooth state), for
    hardware counter runs: no particle sort, no host synchronisation inside
    the sweeps."""
    import numpy as np

    import blaze_amd as B

    s, _, _, _ = B.parse_scene(a.scene)
    s.residual_tol = 0.0
    s.max_iter_solve = a.probe_sweeps
    W, H, D = s.grid_res
    g = B.Solver(s)
    g.set_rbgs_variant(a.rbgs_variant)
    rng = np.random.default_rng(0)
    g.set_field(B.FIELD_RHS, rng.standard_normal((D, H, W)).astype(np.float32))
    g.run_stage(B.STAGE_PRESSURE, 0.0, float(s.dt))
    log(f"[probe] {a.probe_sweeps} sweeps on {W}x{H}x{D} done", a.quiet)


def stream_probe(a):
    """The bandwidth a plain streaming kernel reaches on the same working set
    as one RBGS sweep (four padded planes of the scene's grid read, two
    written; x = a + b and y = c + d, torch's elementwise kernels), as the
    reference point for the sweep kernel's rate. Prints one JSON line."""
    import torch

    import blaze_amd as B

    s, _, _, _ = B.parse_scene(a.scene)
    W, H, D = s.grid_res
    n = (W // 2 + 2) * (H + 2) * ((D + 3) // 4 + 2) * 4
    dev = torch.device("cuda:0")
    pl = [torch.rand(n, device=dev) for _ in range(6)]
    reps = 20

    def once():
        torch.add(pl[0], pl[1], out=pl[4])
        torch.add(pl[2], pl[3], out=pl[5])

    for _ in range(3):
        once()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        once()
    e1.record()
    e1.synchronize()
    us = e0.elapsed_time(e1) * 1e3 / reps
    gbs = 6 * n * 4 / (us * 1e3)
    print(json.dumps({"stream_probe": {"planes_floats": n, "bytes_per_pass": 6 * n * 4, "us_per_pass": us,
                                       "GB_per_s": gbs, "kernels": "torch.add x2 (2 reads, 1 write each)"}}),
          flush=True)


def sweep_scaling(a):
    """Sweep time per cell as the working set and the wave count grow: one
    and two 256^3 grids (212 and 424 MB of sweep working set, against the
    256 MB Infinity Cache), eight and sixteen 128^3 grids (the same cells
    as one and two 256^3 grids in 2.2 times the waves). Each batch's
    autotuner times every sweep variant over all its grids in one launch
    per sweep. Prints one JSON line per batch."""
    import blaze_amd as B

    s, _, _, _ = B.parse_scene(a.scene)
    s.residual_tol = 1e-4
    s.max_iter_solve = 20
    s.residual_check_every = 10
    s.cfl_number = 0.0
    for n, res in ((1, 256), (2, 256), (8, 128), (16, 128)):
        t = s.copy()
        for ax in range(3):
            t.grid_res[ax] = res
        t.dx = 1.0 / res
        b = B.Batch([t] * n)
        b.autotune(20)
        st = b.stats()
        cells = n * res ** 3
        us = {int(st.tune_variant[c]): st.tune_ms[c] * 1e3 for c in range(len(st.tune_variant))}
        print(json.dumps({"sweep_scaling": {"grids": n, "res": res, "cells": cells,
                                            "us_per_sweep": us,
                                            "us_per_sweep_per_million_cells": {v: u / cells * 1e6 for v, u in us.items()}}}),
              flush=True)
        del b


def cpu_model():
    """The host CPU's model name (from /proc/cpuinfo) and its logical CPUs."""
    name = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return name, os.cpu_count() or 0


def cpu_baseline(s, parts, substeps, iters_per_substep, checks_per_substep, n_sweeps, quiet, threads=0):
    """Time the oracle (threads = 0: all host threads, bitwise equal to
    serial; else that many) on one substep of the same 256^3 workload, stage
    by stage, plus n_sweeps RBGS sweeps, and scale to the GPU run's average
    sweeps and residual checks per substep. Returns the CPU voxel-steps/s
    and a description."""
    import oracle as O

    s2 = s.copy()
    s2.residual_tol = 0.0
    s2.max_iter_solve = n_sweeps
    all_threads = O.max_threads()
    if threads:
        O.set_threads(threads)
    o = O.Oracle(s2)
    cores = O.max_threads()
    t0 = time.perf_counter()
    if parts is not None:   # the stages' cost does not depend on the values
        o.add_source(parts)
    t_src = time.perf_counter() - t0
    dt = float(s.dt) / max(1, substeps)
    stages = []
    if abs(s.cooling_rate) > 0:
        stages.append(0)
    if s.drag > 0:
        stages.append(1)
    if abs(s.vorticity_conf) > 0:
        stages.append(2)
    if abs(s.buoyancy) > 0:
        stages.append(3)
    if s.wind_amp > 0:
        stages.append(4)
    if s.turbulence_amp > 0:
        stages.append(5)
    stages += [6]
    t_fixed = 0.0
    per = {}
    for st in stages:
        t0 = time.perf_counter()
        o.run_stage(st, 0.0, dt)
        per[st] = time.perf_counter() - t0
        t_fixed += per[st]
    t0 = time.perf_counter()
    o.run_stage(7, 0.0, dt)   # n_sweeps sweeps
    t_sweeps = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.residual(dt)
    t_res = time.perf_counter() - t0
    for st in (8, 9):
        t0 = time.perf_counter()
        o.run_stage(st, 0.0, dt)
        per[st] = time.perf_counter() - t0
        t_fixed += per[st]
    if threads:
        O.set_threads(all_threads)
    t_sweep = t_sweeps / n_sweeps
    t_substep = t_fixed + iters_per_substep * t_sweep + checks_per_substep * t_res
    W, H, D = s.grid_res
    value = W * H * D / t_substep
    log(f"[cpu] cores {cores} source {t_src:.2f}s fixed {t_fixed:.2f}s sweep {t_sweep * 1e3:.1f}ms "
        f"residual {t_res * 1e3:.1f}ms -> substep {t_substep:.2f}s", quiet)
    sample = (f"oracle FP32 on {W}x{H}x{D}: one substep's stages timed once (fixed part "
              f"{t_fixed:.2f}s), {n_sweeps} RBGS sweeps ({t_sweep * 1e3:.1f} ms/sweep) and one residual "
              f"check ({t_res * 1e3:.1f} ms), combined with the GPU run's averages of "
              f"{iters_per_substep:.1f} sweeps and {checks_per_substep:.1f} checks per substep")
    model, logical = cpu_model()
    return {"value": value, "unit": "voxel-steps/s", "cores": cores, "kind": "port", "sample": sample,
            "cpu_model": model, "host_logical_cpus": logical}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              "(the solver step is single-GPU; multi-GPU slab decomposition is design-only)",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np

    import blaze_amd as B

    if a.calib_copy:
        calib_copy(a.quiet)
    if a.stream_probe:
        stream_probe(a)
        return
    if a.sweep_scaling:
        sweep_scaling(a)
        return
    if a.traffic_probe:
        traffic_probe(a)
        return

    if a.batched:
        batched(a)
        return
    if a.animated:
        animated(a)
        return
    s, _, _, gen = B.parse_scene(a.scene)
    W, H, D = s.grid_res
    N3 = W * H * D
    parts = B.generate_particles(gen)
    # the headline frames without instrumentation (value, ms_per_step), then
    # the same frames with per-stage HIP events (roofline, stage times)
    plain = run_leg(a, s, parts, None, render=not a.no_render)
    timed = run_leg(a, s, parts, None, timing=True)
    warm = None
    if not a.no_warm_start:
        import blaze_amd.warm_start as WS

        if os.path.exists(WS.WEIGHTS):
            net = WS.load_net()
            warm = {"scene_tol": run_leg(a, s, parts, net),
                    "at_tol": {name: tolerance_leg(a, s, parts, net if bits else None, bits)
                               for name, bits in TOL_LEGS}}
    report(a, s, gen, parts, plain, timed, warm)


# the at-tolerance legs: name, warm-start parts (0: plain RBGS; bit 0
# network, bit 1 global correction, bit 2 exact coarse solve in place of the
# network -- the upper bound of what a perfect network would give)
TOL_LEGS = (("plain", 0), ("global", 2), ("network+global", 3), ("exact_coarse+global", 6))


def tolerance_leg(a, s, parts, net, bits):
    """The solve to the scene's tolerance ||div u||_inf <= tol with the sweep
    cap lifted (a.tol_cap): the same plain warmup frames as every other leg
    (scene settings, no warm start, so all legs start the timed frames from
    the same state bit for bit), then a.tol_steps timed frames with the
    limits lifted and, when bits != 0, the warm start with those parts.
    Returns voxel-steps/s, ms per frame, counted and executed sweeps (the
    latter includes the judge's batches on the path not kept), kept and
    rejected corrections, final residuals and the residual history of the
    first timed solve (sweeps, residual, kind)."""
    import numpy as np

    import blaze_amd as B

    W, H, D = s.grid_res
    s2 = s.copy()
    s2.warm_start = a.warm_every if bits else 0
    g = B.Solver(s2)
    g.set_residual_trace(4096)
    t = np.float32(0.0)
    frame = s.source_range[0]

    def one_frame():
        nonlocal t, frame
        if frame <= s.source_range[1]:
            g.add_source(parts if frame == s.source_range[0] else None)
        g.step(float(t))
        t = np.float32(t + np.float32(s.dt))
        frame += 1
        return g.stats()

    for _ in range(a.warmup):
        one_frame()
    g.set_solve_limits(float(s.residual_tol), a.tol_cap, int(s.residual_check_every))
    if bits:
        g.set_warm_start(net, parts=bits)
    g.sync()
    rej0 = g.stats().warm_start_fallbacks
    sub = iters = executed = kept = 0
    res, curve = [], None
    t0 = time.perf_counter()
    for k in range(a.tol_steps):
        st = one_frame()
        sub += st.substeps
        iters += st.rbgs_iters
        executed += st.rbgs_sweeps_executed
        kept += st.warm_start_used
        res.append(st.residual)
        if k == 0:
            curve = g.residual_trace(0)
        log(f"[tol {a.tol_cap} parts {bits}] frame {frame - 1}: substeps {st.substeps} sweeps {st.rbgs_iters} "
            f"executed {st.rbgs_sweeps_executed} residual {st.residual:.3e} step {st.step_ms:.1f} ms", a.quiet)
    g.sync()
    el = time.perf_counter() - t0
    rej = g.stats().warm_start_fallbacks - rej0
    g.close()
    step = max(1, len(curve) // 40)
    return {"value": W * H * D * sub / el, "ms_per_step": el * 1e3 / a.tol_steps, "substeps": sub,
            "rbgs_sweeps": iters, "rbgs_sweeps_executed": executed, "corrections_kept": kept,
            "corrections_rejected": rej, "max_final_residual": max(res), "final_residuals": res,
            "all_reached_tolerance": bool(max(res) <= float(s.residual_tol)),
            "first_solve_residual_curve": [[e[0], e[2], "ckr"[e[1]]] for e in curve[::step]] + (
                [[curve[-1][0], curve[-1][2], "ckr"[curve[-1][1]]]] if curve and (len(curve) - 1) % step else [])}


def run_leg(a, s, parts, net, tol=None, timing=False, render=False):
    """Warmup frames, then a.steps timed frames of the scene; net != None:
    with the learned warm start (same tolerance and cap); tol: another
    residual tolerance for the leg; timing: per-stage HIP events (stage and
    sweep-kernel times; about 2% slower); render: render the final state."""
    import numpy as np

    import blaze_amd as B

    W, H, D = s.grid_res
    N3 = W * H * D
    s = s.copy()
    if tol is not None:
        s.residual_tol = tol
    if net is not None:
        s.warm_start = a.warm_every
    g = B.Solver(s)
    g.set_timing(timing)
    g.set_rbgs_variant(a.rbgs_variant)
    if net is not None:
        g.set_warm_start(net, parts=a.warm_parts)
    t = np.float32(0.0)
    frame = s.source_range[0]
    sr_end = s.source_range[1]
    tag = ("warm" if net is not None else "plain") + ("" if tol is None else "@eq") + (" timed" if timing else "")

    def one_frame():
        nonlocal t, frame
        if frame <= sr_end:
            g.add_source(parts if frame == s.source_range[0] else None)
        g.step(float(t))
        st = g.stats()
        t = np.float32(t + np.float32(s.dt))
        frame += 1
        return st

    tw = time.perf_counter()
    for _ in range(a.warmup):
        st = one_frame()
        log(f"[{tag} warmup] frame {frame - 1}: substeps {st.substeps} sweeps {st.rbgs_iters} "
            f"residual {st.residual:.3e} step {st.step_ms:.1f} ms", a.quiet)
    g.sync()
    log(f"[{tag} warmup] {time.perf_counter() - tw:.1f}s", a.quiet)
    kept = rejected0 = 0
    rejected0 = g.stats().warm_start_fallbacks

    tot_sub = tot_iters = tot_checks = tot_launch = tot_timed_sweeps = tot_executed = 0
    kernel_kind = 0
    tot_rbgs_ms = 0.0
    stage_ms = [0.0] * B.NUM_STAGES
    residuals = []
    g.sync()
    t0 = time.perf_counter()
    for k in range(a.steps):
        st = one_frame()
        tot_sub += st.substeps
        tot_iters += st.rbgs_iters
        tot_executed += st.rbgs_sweeps_executed
        tot_checks += st.residual_checks
        tot_launch += st.rbgs_launches
        tot_timed_sweeps += st.rbgs_sweeps_timed
        tot_rbgs_ms += st.rbgs_kernel_ms
        kernel_kind = st.rbgs_kernel or kernel_kind
        for i in range(B.NUM_STAGES):
            stage_ms[i] += st.stage_ms[i]
        residuals.append(st.residual)
        kept += st.warm_start_used
        log(f"[{tag} frame {frame - 1}] substeps {st.substeps} sweeps {st.rbgs_iters} checks "
            f"{st.residual_checks} residual {st.residual:.3e} step {st.step_ms:.1f} ms", a.quiet)
    g.sync()
    elapsed = time.perf_counter() - t0
    rend = render_leg(a, g, s) if render else None
    return dict(elapsed=elapsed, tot_sub=tot_sub, tot_iters=tot_iters, tot_executed=tot_executed,
                tot_checks=tot_checks, tot_launch=tot_launch,
                tot_timed_sweeps=tot_timed_sweeps, tot_rbgs_ms=tot_rbgs_ms, kernel_kind=kernel_kind,
                stage_ms=stage_ms, residuals=residuals, kept=kept,
                rejected=g.stats().warm_start_fallbacks - rejected0, variant=g.stats().rbgs_kernel, render=rend)


def render_leg(a, g, s):
    """Render the final state with the scene's render settings (the frame
    loop's consumer): wall time per frame (scatter build, blur passes, ray
    march, copy of the image to the host), and the image for the CPU
    comparison."""
    import blaze_amd as B

    rs = B.parse_render(a.scene)
    img = g.render(rs)  # first call allocates the scatter volumes and the frame
    t0 = time.perf_counter()
    for _ in range(a.render_reps):
        img = g.render(rs)
    ms = (time.perf_counter() - t0) * 1e3 / a.render_reps
    log(f"[render] {rs.width}x{rs.height}: {ms:.2f} ms per frame, lit {(img.sum(axis=2) > 0).mean():.2f}, "
        f"max {img.max():.3f}", a.quiet)
    return dict(ms=ms, settings=rs, image=img, T=g.field(B.FIELD_T),
                dom=(list(s.domain_min), list(s.domain_max)), dx=float(s.dx))


def render_report(a, R):
    """GPU render time, the CPU oracle on an evenly spread sample of rows
    (scaled to the frame) and the sampled rows compared bit for bit."""
    import numpy as np

    rs, img = R["settings"], R["image"]
    out = {"workload": f"render of the final 256^3 state, {rs.width}x{rs.height}, {rs.light_count} lights, "
                       f"multi-scatter {rs.ms_blur_iter} blur passes (scene render keys)",
           "ms_per_frame": R["ms"], "pixels_per_s": rs.width * rs.height / (R["ms"] / 1e3),
           "lit_fraction": float((img.sum(axis=2) > 0).mean())}
    if not a.no_cpu_baseline:
        import oracle as O

        n = max(1, a.render_cpu_rows)
        stride = max(1, rs.height // n)
        row0 = stride // 2
        t0 = time.perf_counter()
        ref = O.render(R["T"], R["dx"], R["dom"][0], R["dom"][1], rs, rows=(row0, rs.height, stride))
        t = time.perf_counter() - t0
        rows = list(range(row0, rs.height, stride))
        same = bool(np.array_equal(img[rows].view(np.uint32), ref[rows].view(np.uint32)))
        cpu_ms = t * 1e3 * rs.height / len(rows)
        out["cpu_baseline"] = {"ms_per_frame": cpu_ms, "cores": O.max_threads(), "kind": "port",
                               "sample": f"oracle render of {len(rows)} rows (every {stride}th from row {row0}, "
                                         f"scatter volume included) scaled to {rs.height} rows"}
        out["sampled_rows_bitwise_equal_to_oracle"] = same
        log(f"[render] cpu oracle {cpu_ms:.0f} ms per frame (sample of {len(rows)} rows), rows equal {same}",
            a.quiet)
    return out


def report(a, s, gen, parts, L, LT, warm):
    """L: the headline frames without instrumentation; LT: the same frames
    with per-stage HIP events (sweep-kernel and stage times)."""
    import blaze_amd as B

    W, H, D = s.grid_res
    N3 = W * H * D
    elapsed, tot_sub, tot_iters, tot_checks = L["elapsed"], L["tot_sub"], L["tot_iters"], L["tot_checks"]
    tot_launch, tot_timed_sweeps, tot_rbgs_ms = LT["tot_launch"], LT["tot_timed_sweeps"], LT["tot_rbgs_ms"]
    kernel_kind, stage_ms, residuals = LT["kernel_kind"], LT["stage_ms"], L["residuals"]

    value = N3 * tot_sub / elapsed
    ms_per_step = elapsed * 1e3 / a.steps
    # algorithmic bytes of the timed sweeps over their device time; per
    # launch that is 12 B/voxel (fused sweep) or 6 B/voxel (colour pass)
    avg_launch_s = (tot_rbgs_ms / 1e3) / max(1, tot_launch)
    achieved = (RBGS_BYTES_PER_VOXEL_SWEEP * N3 * tot_timed_sweeps / (tot_rbgs_ms / 1e3) / 1e9
                if tot_rbgs_ms > 0 else 0.0)
    kname, kdesc, kfile = RBGS_KERNELS.get(kernel_kind, ("?", "?", ""))
    algo_bytes = N3 * (FIXED_BYTES_PER_VOXEL * tot_sub + 12.0 * tot_iters
                       + RESIDUAL_BYTES_PER_VOXEL * tot_checks)
    step_gbs = algo_bytes / elapsed / 1e9

    traffic = None
    tpath = os.path.join(ROOT, "profiles", kfile) if kfile else ""
    if tpath and os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        if tuple(tj.get("grid", ())) == (W, H, D):
            traffic = tj.get("bytes_per_launch")

    cpu = cpu1 = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline(s, parts, max(1, round(tot_sub / a.steps)), tot_iters / max(1, tot_sub),
                           tot_checks / max(1, tot_sub), a.cpu_sweeps, a.quiet)
        if not a.no_cpu_1thread:
            # the parity mode: one thread (the same bits as every thread count)
            cpu1 = cpu_baseline(s, None, max(1, round(tot_sub / a.steps)), tot_iters / max(1, tot_sub),
                                tot_checks / max(1, tot_sub), 2, a.quiet, threads=1)

    out = {
        "metric": "voxel-steps/s at 256\u00b3, plain RBGS vs warm-start at equal \u2016\u2207\u00b7u\u2016\u221e; HBM GB/s",
        "value": value,
        "unit": "voxel-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded particle ball, scenes/explosion256.blz)",
        "config": {
            "workload": f"explosion256: {W}x{H}x{D} FP32 MAC grid, {gen.count} particles r={gen.radius} "
                        f"radial vel {gen.radial_vel} temp {gen.temp}, cooling {s.cooling_rate}, "
                        f"vorticity {s.vorticity_conf}, RBGS to {s.residual_tol:g} (cap {s.max_iter_solve}), "
                        f"CFL {s.cfl_number:g} (max {s.max_substeps} substeps), frames "
                        f"{a.warmup + 1}..{a.warmup + a.steps}",
            "grid": [W, H, D],
            "substeps": tot_sub,
            "rbgs_sweeps": tot_iters,
            "residual_checks": tot_checks,
            "final_residual": residuals[-1] if residuals else None,
            "parallelism": "single GPU",
        },
        "hbm_gbs_algorithmic_step": step_gbs,
        "stage_ms": {B.STAGE_NAMES[i]: stage_ms[i] for i in range(B.NUM_STAGES) if stage_ms[i] > 0},
        "roofline": {
            "kernel": f"{kname} ({kdesc})",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "avg_launch_us": avg_launch_s * 1e6,
            "launches": tot_launch,
            "timing_pass_ms_per_step": LT["elapsed"] * 1e3 / a.steps,
        },
        "cpu_baseline": cpu,
        "cpu_baseline_1thread": cpu1,
    }
    if L.get("render") is not None:
        out["render"] = render_report(a, L["render"])
    if warm is not None:
        def leg(L):
            return {"value": N3 * L["tot_sub"] / L["elapsed"], "ms_per_step": L["elapsed"] * 1e3 / a.steps,
                    "rbgs_sweeps": L["tot_iters"], "rbgs_sweeps_executed": L["tot_executed"],
                    "substeps": L["tot_sub"],
                    "mean_final_residual": sum(L["residuals"]) / max(1, len(L["residuals"])),
                    "corrections_kept": L["kept"], "corrections_rejected": L["rejected"]}
        ws = leg(warm["scene_tol"])
        T = warm["at_tol"]
        pl = T["plain"]
        out["warm_start"] = {
            "setting": f"warm_start={a.warm_every} (corrections before the sweeps and every {a.warm_every} sweeps), "
                       "weights blaze_amd/warm_start_weights.npy",
            "at_tolerance": {
                "tolerance": float(s.residual_tol),
                "cap": a.tol_cap,
                "frames": f"{a.warmup + 1}..{a.warmup + a.tol_steps} (after {a.warmup} identical plain warmup frames)",
                "legs": T,
                "speedup_vs_plain": {k: v["value"] / pl["value"] for k, v in T.items() if k != "plain"},
                "time_ratio_vs_plain": {k: pl["ms_per_step"] / v["ms_per_step"] for k, v in T.items() if k != "plain"},
                "network_gain_over_global": T["global"]["ms_per_step"] / T["network+global"]["ms_per_step"],
            },
            "at_headline_settings": dict(ws, parts="network + global", ratio_vs_plain=ws["value"] / value,
                                         plain_mean_final_residual=sum(residuals) / max(1, len(residuals))),
        }
    print(json.dumps(out), flush=True)


def animated(a):
    """The plume at 128^3 with a new .part file per frame (100k particles,
    files written to a temporary directory first): frames with the
    reference's order (read the file, upload, splat, step) against frames
    whose next file is read and uploaded on a worker thread and the copy
    stream while the current frame steps. Same fields bit for bit."""
    import concurrent.futures as cf
    import tempfile

    import numpy as np

    import blaze_amd as B

    s, _, _, gen = B.parse_scene(os.path.join(ROOT, "scenes", "plume128.blz"))
    n_frames = a.steps + a.warmup
    tmp = tempfile.mkdtemp(prefix="blz_part_")
    files = []
    for f in range(n_frames):
        gen.seed = 1000 + f
        gen.count = 100000
        path = os.path.join(tmp, f"anim_{f + 1:04d}.part")
        B.generate_particles(gen).tofile(path)
        files.append(path)

    def read(path):
        return np.fromfile(path, np.float32).reshape(-1, 8)

    def run(staged):
        g = B.Solver(s)
        t = np.float32(0.0)
        pool = cf.ThreadPoolExecutor(1)
        if staged:
            g.stage_source(read(files[0]))
        t0 = None
        for f in range(n_frames):
            if f == a.warmup:
                g.sync()
                t0 = time.perf_counter()
            if staged:
                g.add_source_staged()
                nxt = pool.submit(lambda p=files[f + 1]: g.stage_source(read(p))) if f + 1 < n_frames else None
                g.step(float(t))
                if nxt is not None:
                    nxt.result()
            else:
                g.add_source(read(files[f]))
                g.step(float(t))
            t = np.float32(t + np.float32(s.dt))
        g.sync()
        el = time.perf_counter() - t0
        pool.shutdown()
        return el, g.field(B.FIELD_T)

    e_sync, T_sync = run(False)
    e_stage, T_stage = run(True)
    for f in files:
        os.remove(f)
    os.rmdir(tmp)
    W, H, D = s.grid_res
    vs = W * H * D * a.steps
    out = {
        "metric": "voxel-steps/s, 128^3 plume with a new .part file per frame (100k particles)",
        "value": vs / e_stage,
        "unit": "voxel-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": e_stage * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded particle balls written as .part files)",
        "config": {"workload": "plume128, per-frame .part source files, staged uploads", "parallelism": "single GPU"},
        "synchronous": {"value": vs / e_sync, "ms_per_step": e_sync * 1e3 / a.steps},
        "speedup_vs_synchronous": e_sync / e_stage,
        "fields_bitwise_equal": bool(np.array_equal(T_sync.view(np.uint32), T_stage.view(np.uint32))),
    }
    print(json.dumps(out), flush=True)


def batched(a):
    """64 independent scenes, grids N^3 with N uniform in [32, 128] (rounded
    to multiples of 8, seed 5), source radius U[0.03, 0.1], buoyancy
    U[0.5, 2.0], learned warm start on; one captured-graph step for all,
    autotuned sweep tiles."""
    import numpy as np

    import blaze_amd as B
    import blaze_amd.warm_start as WS

    net = None if a.no_warm_start else WS.load_net()

    def run(warm, cap, steps, closed_top=False, tol=1e-4):
        sets, parts = batched_scenes(a.scenes, warm_start=warm, cap=cap, warm_every=a.warm_every,
                                     closed_top=closed_top, tol=tol)
        b = B.Batch(sets)
        if warm:
            b.set_warm_start(net)
        for q in range(a.scenes):
            b.scenes[q].add_source(parts[q])
        v = b.autotune(10)
        t = np.float32(0.0)
        for _ in range(a.warmup):
            b.step(float(t))
            t = np.float32(t + np.float32(1.0 / 24.0))
        t0 = time.perf_counter()
        for _ in range(steps):
            b.step(float(t))
            t = np.float32(t + np.float32(1.0 / 24.0))
        el = time.perf_counter() - t0
        res = [b.scenes[q].stats().residual for q in range(a.scenes)]
        sweeps = sum(b.scenes[q].stats().rbgs_sweeps_executed for q in range(a.scenes))
        log(f"[batched warm={warm} cap={cap} closed_top={closed_top} tol={tol:g}] {el * 1e3 / steps:.1f} ms per step, "
            f"residual max {max(res):.3e} mean {sum(res) / len(res):.3e}, sweeps executed in the last step "
            f"{sweeps}", a.quiet)
        return b, v, el, res, sweeps

    b, v, elapsed, res, _ = run(not a.no_warm_start, 100, a.steps)
    st = b.stats()
    vox = st.voxels
    out = {
        "metric": f"voxel-steps/s, {a.scenes} batched scenes of 32^3..128^3 (one captured-graph step, "
                  f"warm start {'off' if a.no_warm_start else 'on'})",
        "value": vox * a.steps / elapsed,
        "unit": "voxel-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": elapsed * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded particle balls)",
        "config": {"workload": f"{a.scenes} scenes, N uniform in [32,128] step 8 (seed 5), RBGS to 1e-4 cap 100",
                   "voxels": vox, "parallelism": "single GPU"},
        "batch": {"autotuned_variant": v, "graph_nodes": st.graph_nodes, "capture_ms": st.capture_ms,
                  "tune_ms_per_sweep": {st.tune_variant[k]: st.tune_ms[k] for k in range(B.BATCH_TUNE)},
                  "max_final_residual": max(res), "mean_final_residual": sum(res) / len(res)},
    }
    del b
    if not a.no_warm_start:
        # the same scenes and cap without the warm start: time and final
        # residuals at equal sweeps (the cap ends most solves of both)
        bb, _, el, rr, sw = run(False, 100, a.steps)
        out["plain_same_cap"] = {"ms_per_step": el * 1e3 / a.steps, "value": vox * a.steps / el,
                                 "max_final_residual": max(rr), "mean_final_residual": sum(rr) / len(rr),
                                 "graph_nodes": bb.stats().graph_nodes}
        out["warm_vs_plain"] = {"time_ratio": out["plain_same_cap"]["ms_per_step"] / out["ms_per_step"],
                                "mean_residual_ratio": out["plain_same_cap"]["mean_final_residual"]
                                / max(1e-300, out["batch"]["mean_final_residual"])}
        del bb
    if a.batched_tol_cap > 0 and not a.no_warm_start:
        # equal tolerance: every wall closed (no residual floor |mean(rhs)|
        # from an open top), the cap lifted, warm start on and off; where
        # every scene reaches the tolerance, the time per step is the time
        # to it
        legs = []
        for t in (float(x) for x in a.batched_tol.split(",")):
            tol = {"tolerance": t, "cap": a.batched_tol_cap, "steps": a.batched_tol_steps}
            for warm in (True, False):
                bt, _, el, rr, sw = run(warm, a.batched_tol_cap, a.batched_tol_steps, closed_top=True, tol=t)
                tol["warm" if warm else "plain"] = {
                    "ms_per_step": el * 1e3 / a.batched_tol_steps,
                    "scenes_at_tolerance": sum(1 for r in rr if r <= t),
                    "max_final_residual": max(rr), "mean_final_residual": sum(rr) / len(rr),
                    "sweeps_executed_last_step": sw, "graph_nodes": bt.stats().graph_nodes,
                    "graphs_run_last_step": bt.stats().graphs_run}
                del bt
            tol["time_ratio_plain_over_warm"] = tol["plain"]["ms_per_step"] / tol["warm"]["ms_per_step"]
            legs.append(tol)
        out["at_tolerance_closed_top"] = legs
    print(json.dumps(out), flush=True)


def batched_scenes(n=64, warm_start=True, cap=100, warm_every=1, closed_top=False, tol=1e-4):
    """The batched workload: n scenes, grids N^3 with N uniform in [32, 128]
    rounded to multiples of 8 (seed 5), source radius U[0.03, 0.1], buoyancy
    U[0.5, 2.0], solid walls with an open top (closed_top: closed too), RBGS
    to `tol` with a cap of `cap` sweeps, the warm start (when on) before
    the sweeps and every `warm_every` sweeps (1: before them only); returns
    the settings and the particles of each scene."""
    import numpy as np

    import blaze_amd as B

    rng = np.random.default_rng(5)
    sets, parts = [], []
    for q in range(n):
        n = int(np.clip(round(rng.uniform(32, 128) / 8) * 8, 32, 128))
        s = B.default_settings()
        for ax in range(3):
            s.grid_res[ax] = n
            s.domain_min[ax] = 0.0
            s.domain_max[ax] = 1.0
        s.dx = 1.0 / n
        s.dt = 1.0 / 24.0
        s.density = 1.0
        s.gravity = -9.8
        s.buoyancy = float(rng.uniform(0.5, 2.0))
        for ax, v in enumerate((1, 1, 1, int(closed_top), 1, 1)):
            s.closed_bounds[ax] = v
        s.residual_tol = tol
        s.max_iter_solve = cap
        s.residual_check_every = 10
        s.cfl_number = 0.0
        s.warm_start = warm_every if warm_start else 0
        g = B.ParticleGen()
        g.count = int(20000 * (n / 64) ** 3)
        g.seed = 1 + q
        g.center[0], g.center[1], g.center[2] = 0.5, 0.1, 0.5
        g.radius = float(rng.uniform(0.03, 0.1))
        g.pscale = 2.0 / n
        g.temp = 1.0
        sets.append(s)
        parts.append(B.generate_particles(g))
    return sets, parts


if __name__ == "__main__":
    main()
