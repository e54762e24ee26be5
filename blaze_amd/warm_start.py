"""Training of the learned warm start on data the solver generates.

No reference counterpart. A pair is taken at a pressure solve of a running
scene: the network input is the solver's own coarse input (the residual
the pressure leaves after the global correction, restricted and
normalised, and the solid mask: ``Solver.warm_start_global`` then
``Solver.warm_start_input``), as a warm-started solve meets it; the target
is the coarse Poisson correction for that input, n E - sum_nbr E = r on the
coarse grid with Neumann walls (solved exactly by the orthonormal DCT-II,
which diagonalises that operator, or, where the blocks are clipped, by the
library's own Gauss-Seidel on the coarse grid), mean removed, unit RMS.
Training runs on the GPU through the library's own forward, backward and
Adam kernels (``Net.train_step``).
"""
from __future__ import annotations

import os

import numpy as np

from . import (FIELD_P, FIELD_RHS, STAGE_ADVECT, STAGE_BUOYANCY, STAGE_COOLDOWN, STAGE_DIVERGENCE, STAGE_DRAG,
               STAGE_PRESSURE, STAGE_PROJECT_UPDATE, STAGE_TURBULENCE, STAGE_VORTICITY, STAGE_WIND, Net, Solver,
               default_settings, generate_particles, parse_scene)

PKG = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(PKG, "warm_start_weights.npy")


def _force_stages(s):
    st = []
    if abs(s.cooling_rate) > 0:
        st.append(STAGE_COOLDOWN)
    if s.drag > 0:
        st.append(STAGE_DRAG)
    if abs(s.vorticity_conf) > 0:
        st.append(STAGE_VORTICITY)
    if abs(s.buoyancy) > 0:
        st.append(STAGE_BUOYANCY)
    if s.wind_amp > 0:
        st.append(STAGE_WIND)
    if s.turbulence_amp > 0:
        st.append(STAGE_TURBULENCE)
    return st


def coarse_target(x, coarse_dims, sweeps):
    """Solve n E - sum_nbr E = x[0] on the coarse grid (Neumann, same
    Gauss-Seidel solver, zero start) and return E padded to the network
    grid with its mean over the real cells removed, scaled to unit RMS (the
    solver fits the correction's scale itself, so only its shape is
    learned)."""
    wc, hc, dc = coarse_dims
    _, dn, hn, wn = x.shape
    cs = default_settings()
    cs.grid_res[0], cs.grid_res[1], cs.grid_res[2] = wc, hc, dc
    cs.dx, cs.dt, cs.density = 1.0, 1.0, 1.0
    cs.max_iter_solve = sweeps
    cs.residual_tol = 0.0
    g = Solver(cs)
    g.set_field(FIELD_RHS, np.ascontiguousarray(x[0, :dc, :hc, :wc]))
    g.set_field(FIELD_P, np.zeros((dc, hc, wc), np.float32))
    g.run_stage(STAGE_PRESSURE, 0.0, 1.0)
    e = g.field(FIELD_P).astype(np.float64)
    e -= e.mean()
    rms = np.sqrt(np.mean(e * e))
    if rms > 0:
        e /= rms
    out = np.zeros((dn, hn, wn), np.float32)
    out[:dc, :hc, :wc] = e
    return out


def coarse_target_dct(x, coarse_dims):
    """The exact solution of n E - sum_nbr E = x[0] on the coarse grid
    (Neumann graph Laplacian, diagonalised by the orthonormal DCT-II:
    eigenvalues sum over axes of 2 - 2 cos(pi a / n)), the constant mode set
    to 0, padded to the network grid and scaled to unit RMS."""
    import scipy.fft as F

    wc, hc, dc = coarse_dims
    _, dn, hn, wn = x.shape
    r = x[0, :dc, :hc, :wc].astype(np.float64)
    lam = [2.0 - 2.0 * np.cos(np.pi * np.arange(n) / n) for n in (dc, hc, wc)]
    L = lam[0][:, None, None] + lam[1][None, :, None] + lam[2][None, None, :]
    c = F.dctn(r, type=2, norm="ortho")
    L[0, 0, 0] = 1.0
    c = c / L
    c[0, 0, 0] = 0.0
    e = F.idctn(c, type=2, norm="ortho")
    rms = np.sqrt(np.mean(e * e))
    if rms > 0:
        e /= rms
    out = np.zeros((dn, hn, wn), np.float32)
    out[:dc, :hc, :wc] = e
    return out


def scene_pairs(scene, frames, res=None, sweeps=2000, every=1, seed_net=None, log=None, points=1, between=50,
                after_global=False):
    """Run `scene` (a .blz path) for `frames` frames stage by stage; at each
    `every`-th frame record (input, target) at `points` moments of the
    pressure solve: before its sweeps and after each further `between`
    sweeps (the residuals a repeated correction meets). `res` overrides the
    grid (downsampled runs keep the domain and scale dx, and the particle
    count scales with the cells). after_global: the global correction is
    applied before each sample (the input the network meets after it) and
    the target is the exact coarse solution."""
    s, _, _, gen = parse_scene(scene)
    if res is not None:
        f = s.grid_res[0] / res[0]
        for a in range(3):
            s.grid_res[a] = res[a]
        s.dx = s.dx * f
        if gen is not None:
            gen.pscale = gen.pscale * f
            if after_global:
                gen.count = max(1, int(round(gen.count / f ** 3)))
    s.cfl_number = 0.0
    s.warm_start = 0
    if points > 1:
        s.residual_tol = 0.0
        s.max_iter_solve = between
    parts = generate_particles(gen)
    g = Solver(s)
    net = seed_net or Net(seed=0)
    g.set_warm_start(net)
    _, cdims = g.warm_start_dims()
    dt = float(s.dt)
    pairs = []
    t = np.float32(0.0)
    for fr in range(frames):
        frame = s.source_range[0] + fr
        if frame <= s.source_range[1]:
            g.add_source(parts if (s.source_animated == 0 or fr == 0) else None)
        for st in _force_stages(s):
            g.run_stage(st, float(t), dt)
        g.run_stage(STAGE_DIVERGENCE, float(t), dt)
        whole = all(s.grid_res[a] % cdims[a] == 0 for a in range(3))   # whole blocks: the DCT applies
        for c in range(points):
            if fr % every == 0:
                if after_global:
                    g.warm_start_global(dt)
                x = g.warm_start_input(dt)
                if np.isfinite(x).all() and np.abs(x[0]).max() > 0:
                    tgt = coarse_target_dct(x, cdims) if (after_global and whole) else coarse_target(x, cdims, sweeps)
                    pairs.append((x, tgt))
            if c < points - 1:
                g.run_stage(STAGE_PRESSURE, float(t), dt)
        g.run_stage(STAGE_PRESSURE, float(t), dt)
        g.run_stage(STAGE_PROJECT_UPDATE, float(t), dt)
        g.run_stage(STAGE_ADVECT, float(t), dt)
        t = np.float32(t + np.float32(dt))
        if log and fr % 20 == 0:
            log(f"[data] {os.path.basename(scene)} res {tuple(s.grid_res)} frame {fr}: {len(pairs)} pairs")
    g.set_warm_start(None)
    return pairs


def augment(pairs):
    """Mirror images along x and z (the Poisson operator and the walls'
    treatment are symmetric under them)."""
    out = []
    for x, y in pairs:
        out.append((x, y))
        out.append((np.ascontiguousarray(x[:, :, :, ::-1]), np.ascontiguousarray(y[:, :, ::-1])))
        out.append((np.ascontiguousarray(x[:, ::-1]), np.ascontiguousarray(y[::-1])))
    return out


def train(net, pairs, steps=1500, batch=4, lr=1e-3, seed=0, log=None):
    """Adam on mini-batches of equal-size pairs; cosine learning-rate decay.
    Returns the loss history."""
    rng = np.random.default_rng(seed)
    groups = {}
    for x, y in pairs:
        groups.setdefault(x.shape, []).append((x, y))
    keys = sorted(groups, key=lambda k: -np.prod(k))
    weights = np.array([len(groups[k]) * np.prod(k[1:]) for k in keys], np.float64)
    weights /= weights.sum()
    hist = []
    for it in range(steps):
        k = keys[rng.choice(len(keys), p=weights)]
        grp = groups[k]
        idx = rng.choice(len(grp), size=min(batch, len(grp)), replace=False)
        xb = np.stack([grp[i][0] for i in idx])
        yb = np.stack([grp[i][1] for i in idx])
        cur = lr * 0.5 * (1.0 + np.cos(np.pi * it / steps))
        hist.append(net.train_step(xb, yb, lr=cur))
        if log and (it % 100 == 0 or it == steps - 1):
            log(f"[train] step {it} loss {np.mean(hist[-100:]):.4e} lr {cur:.2e}")
    return hist


def relative_error(net, pairs):
    """Mean relative L2 error of the prediction over pairs, after the
    optimal scaling the solver applies (shape error: 0 exact, 1 useless)."""
    errs = []
    for x, y in pairs:
        p = net.forward(x[None])[0].astype(np.float64)
        p -= p.mean()
        yy = y.astype(np.float64)
        a = float(np.sum(p * yy) / max(np.sum(p * p), 1e-300))
        errs.append(np.linalg.norm(a * p - yy) / (np.linalg.norm(yy) + 1e-30))
    return float(np.mean(errs))


def load_net(path=WEIGHTS, seed=0):
    net = Net(seed=seed)
    if os.path.exists(path):
        net.set_params(np.load(path, allow_pickle=False))
    return net


def main(argv=None):
    """Generate pairs from the plume scenes (native and 2x-downsampled
    grids), train, report held-out error, save the FP32 weights."""
    import argparse
    import time

    root = os.path.dirname(PKG)
    ap = argparse.ArgumentParser(description="train the warm-start U-Net on solver-generated pairs")
    ap.add_argument("--frames", type=int, default=200)
    ap.add_argument("--every", type=int, default=1)
    ap.add_argument("--sweeps", type=int, default=2000, help="coarse Gauss-Seidel sweeps per target")
    ap.add_argument("--steps", type=int, default=12000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=3e-4)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--points", type=int, default=6, help="pairs per solve: before the sweeps and after each --between sweeps")
    ap.add_argument("--between", type=int, default=40)
    ap.add_argument("--before-global", action="store_true",
                    help="inputs without the global correction applied first (the network-first order)")
    ap.add_argument("--out", default=WEIGHTS)
    a = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)
    t0 = time.time()
    scenes = [(os.path.join(root, "scenes", "plume64.blz"), None), (os.path.join(root, "scenes", "plume64.blz"),
                                                                     (32, 32, 32)),
              (os.path.join(root, "scenes", "plume128.blz"), None), (os.path.join(root, "scenes", "plume128.blz"),
                                                                      (64, 64, 64))]
    pairs = []
    for path, res in scenes:
        pairs += scene_pairs(path, a.frames, res=res, sweeps=a.sweeps, every=a.every, log=log, points=a.points,
                             between=a.between, after_global=not a.before_global)
    log(f"[data] {len(pairs)} pairs in {time.time() - t0:.1f}s")
    rng = np.random.default_rng(a.seed)
    perm = rng.permutation(len(pairs))
    n_test = max(1, len(pairs) // 10)
    test = [pairs[i] for i in perm[:n_test]]
    tr = augment([pairs[i] for i in perm[n_test:]])
    if os.path.exists(WEIGHTS):
        log(f"[train] shipped weights' held-out relative error {relative_error(load_net(), test):.3f}")
    net = Net(seed=a.seed)
    log(f"[train] untrained held-out relative error {relative_error(net, test):.3f}")
    train(net, tr, steps=a.steps, batch=a.batch, lr=a.lr, seed=a.seed, log=log)
    log(f"[train] held-out relative error {relative_error(net, test):.3f} ({time.time() - t0:.1f}s)")
    np.save(a.out, net.params(), allow_pickle=False)
    log(f"[train] saved {a.out}")


if __name__ == "__main__":
    main()
