"""Pins of the oracle's force, noise and advection functions with non-zero
outputs (FP64 oracle, closed forms).

Each test builds a field whose result the mathematics fixes exactly, chosen
so that a plausible transcription mistake changes the answer: a sign flip,
the normalised N in place of grad|omega|, a dropped 1/2, a missing speed
clamp, corrected V-for-U/W reads in the curl noise, wrong RK3 weights, a
wrong fade or lerp in the Perlin noise. The Perlin noise is restated here
independently as the textbook improved-noise formula (Perlin 2002: fade
6t^5 - 15t^4 + 10t^3, trilinear blend of the eight corner dot products),
with the corner gradients taken from the lattice hash, which
tests/test_oracle_pins.py pins by its integer formula.
"""
import itertools
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(__file__))
from helpers import random_state  # noqa: E402
from helpers import settings as _settings  # noqa: E402

O = pytest.importorskip("oracle")
B = pytest.importorskip("blaze_amd")

F_T, F_U, F_V, F_W = 0, 1, 2, 3
STAGE_VORTICITY, STAGE_WIND, STAGE_TURBULENCE, STAGE_ADVECT = 2, 4, 5, 9


def settings(res, **kw):
    return _settings(B, res, **kw)


def f32(v):
    return np.float64(np.float32(v))


# ------------------------------------------- independent Perlin restatement
# gradient vectors of cuda_noise.cuh:179-201 (cases 0xC-0xF repeat earlier
# directions; 0xD and 0xF as written there)
GRAD = {0x0: (1, 1, 0), 0x1: (-1, 1, 0), 0x2: (1, -1, 0), 0x3: (-1, -1, 0), 0x4: (1, 0, 1),
        0x5: (-1, 0, 1), 0x6: (1, 0, -1), 0x7: (-1, 0, -1), 0x8: (0, 1, 1), 0x9: (0, -1, 1),
        0xA: (0, 1, -1), 0xB: (0, -1, -1), 0xC: (1, 1, 0), 0xD: (0, -1, 1), 0xE: (-1, 1, 0),
        0xF: (0, -1, -1)}


def fade(t):
    return ((6.0 * t - 15.0) * t + 10.0) * t ** 3


def corner_grad(c, seed):
    return np.array(GRAD[O.lib.orc_lattice_hash(*map(float, c), float(seed)) & 0xF], dtype=float)


def perlin_ref(p, scale, seed):
    """Improved Perlin noise at p * scale: sum over the 8 lattice corners of
    g_c . (x - c) times the product of fade weights (the trilinear blend)."""
    x = np.asarray(p, dtype=float) * scale
    x0 = np.floor(x)
    f = x - x0
    w1 = fade(f)
    total = 0.0
    for cx, cy, cz in itertools.product((0, 1), repeat=3):
        c = np.array([cx, cy, cz], dtype=float)
        wt = np.prod(np.where(c == 1, w1, 1.0 - w1))
        total += wt * corner_grad(x0 + c, seed) @ (f - c)
    return total


def test_fade_closed_form():
    """The fade polynomial of the independent restatement: 0 and 1 at the
    ends with zero slope and curvature, 1/2 in the middle."""
    assert fade(0.0) == 0.0 and fade(1.0) == 1.0 and fade(0.5) == 0.5
    h = 1e-4
    # flat ends: fade(h) = 10 h^3 + O(h^4), so fade(h) / h -> 0
    assert abs(fade(h)) / h < 1e-6 and abs(1 - fade(1 - h)) / h < 1e-6
    assert abs(fade(h) / h ** 3 - 10) < 1e-2


def test_perlin_interior_matches_improved_noise():
    """perlinNoise (cuda_noise.cuh:570-615) at fractional offsets inside
    lattice cells, away from the corners, equals the textbook formula: pins
    the fade weights and the lerp order, which the lattice-point tests
    cannot see."""
    rng = np.random.default_rng(20)
    for scale, seed in ((1.0, 123), (0.25, 7), (4.0, 456), (0.5, 789)):
        for _ in range(300):
            p = rng.uniform(0, 30, 3) / scale
            got = O.lib.orc_perlin_f64(*p, scale, seed)
            want = perlin_ref(p, scale, seed)
            assert abs(got - want) < 1e-12, (p, scale, seed, got, want)


def test_perlin_single_axis_offset_is_a_two_corner_blend():
    """At a lattice point moved by h along x only, the noise is
    g000_x h (1 - fade(h)) + g100_x (h - 1) fade(h): a hand-computable value
    that a missing fade (linear blend) or swapped lerp arguments change."""
    for corner in [(3, 5, 2), (10, 1, 7), (0, 4, 9)]:
        g0 = corner_grad(corner, 123)
        g1 = corner_grad((corner[0] + 1, corner[1], corner[2]), 123)
        for h in (0.125, 0.25, 0.375, 0.5, 0.8125):
            want = g0[0] * h * (1 - fade(h)) + g1[0] * (h - 1) * fade(h)
            got = O.lib.orc_perlin_f64(corner[0] + h, corner[1], corner[2], 1.0, 123)
            assert abs(got - want) < 1e-14, (corner, h, got, want)


# --------------------------------------------------------------- vorticity
def test_vorticity_confinement_closed_form():
    """vorticity_confinement_kernel (src/FluidSolver.cu:203-248) on
    U = W = 0, V = c0 + a x^2 / 2 (x = object coordinate of the V face):
    every trilinear sample hits V's x nodes and the central differences are
    exact for quadratics, so omega = (0, 0, a x), grad|omega| = (a, 0, 0)
    (un-normalised: the N of :226 is unused), speed = c0 + a x^2 / 2, and
    where |omega| + eps > speed omega is scaled to speed. The force
    conf dx (grad|omega| x omega) = (0, -conf dx a omega_c, 0) is split half
    to each face: an interior V face gains dt * f_y. With c0 = 1 and a = 4 the
    clamp is active for x > 1 - 1/sqrt(2) and inactive below, so a missing
    clamp, N for grad|omega| (a factor a), a sign flip or a dropped half all
    fail."""
    W, H, D = 32, 16, 16
    conf = f32(0.7)   # settings hold float32
    s = settings((W, H, D), vorticity_conf=conf)
    o = O.Oracle(s, fp64=True)
    dx = s.dx
    c0, a = 1.0, 4.0
    nx, ny, nz = o.dims(F_V)
    x = (np.arange(nx) + 0.5) * dx
    V0 = np.broadcast_to(c0 + 0.5 * a * x * x, (nz, ny, nx)).copy()
    o.set_field(F_U, np.zeros(o.dims(F_U)[::-1]))
    o.set_field(F_V, V0)
    o.set_field(F_W, np.zeros(o.dims(F_W)[::-1]))
    dt = f32(1 / 24)
    o.run_stage(STAGE_VORTICITY, 0.0, dt)
    dV = o.field(F_V) - V0
    # omega_c per x column of cells
    xc = (np.arange(W) + 0.5) * dx
    om = a * xc
    speed = c0 + 0.5 * a * xc * xc
    eps = f32(1e-20) * (1.0 / (dx * dt))
    onorm = om + eps
    om_c = np.where(onorm > speed, speed * om / onorm, om)
    fy = conf * dx * (-a * om_c)
    clamped = onorm > speed
    # cells whose twelve-sample stencils stay on exact V nodes: i in 2..W-4
    # (i = W-3 reaches node W-1, moved to W-1.001 by the sample clamp)
    i_ok = np.arange(2, W - 3)
    assert clamped[i_ok].any() and (~clamped[i_ok]).any()
    # interior faces j in 3..H-3 (both cells j-1, j active), k in 2..D-3
    got = dV[2:D - 2, 3:H - 2, :][:, :, i_ok]
    want = np.broadcast_to(dt * fy[i_ok], got.shape)
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()
    # faces with one active cell (j = 2 and j = H-2) gain half
    for j in (2, H - 2):
        assert np.abs(dV[2:D - 2, j, :][:, i_ok] - 0.5 * dt * fy[i_ok]).max() < 1e-12 * np.abs(fy).max()
    # no x or z force: U and W unchanged
    assert np.abs(o.field(F_U)).max() < 1e-14 and np.abs(o.field(F_W)).max() < 1e-14
    # outside the active box nothing changes
    assert not dV[:, :2, :].any() and not dV[:, H - 1:, :].any()


# -------------------------------------------------------------- curl noise
def test_curl_noise_closed_form_with_reference_reads():
    """add_curl_noise_kernel (src/FluidSolver.cu:271-293) on linear back
    buffers U_b = A + p x, V_b = b y, W_b = C + q z (x = i dx on U's faces,
    y = j dx on V's, z = k dx on W's). The reference differences U and W
    against V_b(i,j,k) (:276, :278), so for cell (i, j, k)
      dudx = (A + p (i+1) dx - b j dx)/dx, dvdy = b,
      dwdz = (C + q (k+1) dx - b j dx)/dx,
    curl = (dvdy - dwdz, dwdz - dudx, dudx - dvdy), rescaled to the length of
    the face-averaged back velocity (A + p (i+1/2) dx, b (j+1/2) dx,
    C + q (k+1/2) dx); each cell adds dt/2 of it to its two faces. Reading
    U_b and W_b as a correct difference would give curl (b - q, q - p, p - b)
    instead; a wrong sign, a missing face average or a different split shows
    as well."""
    W, H, D = 12, 10, 8
    s = settings((W, H, D))
    o = O.Oracle(s, fp64=True)
    dx = s.dx
    A, p, b, Cc, q = 0.8, 1.7, 2.5, -0.3, -0.9

    def lin(which, axis, c0, c1):
        nx, ny, nz = o.dims(which)
        kk, jj, ii = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        return c0 + c1 * (ii, jj, kk)[axis] * dx

    o.set_field_back(F_U, lin(F_U, 0, A, p))
    o.set_field_back(F_V, lin(F_V, 1, 0.0, b))
    o.set_field_back(F_W, lin(F_W, 2, Cc, q))
    st = random_state(s, seed=21)
    for w in (F_U, F_V, F_W):
        o.set_field(w, st[w])
    dt = f32(1 / 24)
    o.curl_noise(dt)
    kk, jj, ii = np.meshgrid(np.arange(D), np.arange(H), np.arange(W), indexing="ij")
    vb = b * jj * dx
    dudx = (A + p * (ii + 1) * dx - vb) / dx
    dvdy = np.full(ii.shape, b)
    dwdz = (Cc + q * (kk + 1) * dx - vb) / dx
    curl = np.stack([dvdy - dwdz, dwdz - dudx, dudx - dvdy])
    avg = np.stack([A + p * (ii + 0.5) * dx, b * (jj + 0.5) * dx, Cc + q * (kk + 0.5) * dx])
    eps = f32(1e-20) * (1.0 / (dx * dt))
    cn = curl * np.sqrt((avg ** 2).sum(0)) / (np.sqrt((curl ** 2).sum(0)) + eps)
    inc = 0.5 * dt * cn                                   # (3, D, H, W) per cell
    want = {w: st[w].astype(np.float64) for w in (F_U, F_V, F_W)}
    want[F_U][:, :, :-1] += inc[0]
    want[F_U][:, :, 1:] += inc[0]
    want[F_V][:, :-1, :] += inc[1]
    want[F_V][:, 1:, :] += inc[1]
    want[F_W][:-1] += inc[2]
    want[F_W][1:] += inc[2]
    for w in (F_U, F_V, F_W):
        got = o.field(w)
        assert np.abs(got - want[w]).max() < 1e-12 * (1 + np.abs(want[w]).max()), w
    # the reference reads matter here: the corrected differences differ
    assert np.abs(curl[0] - (b - q)).max() > 1.0


# -------------------------------------------------------------- turbulence
def test_turbulence_partial_mask_and_noise_closed_form():
    """compute_turbulence_kernel (src/FluidSolver.cu:295-335) with uniform
    T inside the temperature ramp and a uniform velocity inside the speed
    ramp: the mask is the constant (1 - (T - t0)/(t1 - t0)) ((|u| - v0)/
    (v1 - v0)), and each face's back value is mask * amp * noise at
    ((i + offset) * freq, j * freq, k * freq) for U (offset on j for V, on k
    for W), offset = fmod(time * (W + H + D)/3, 25000). With freq = 1 and
    offset = 1/4 the noise is the two-corner blend along one axis, computed
    here from the corner gradients."""
    W, H, D = 8, 8, 8
    amp, seed = f32(0.6), 7
    s = settings((W, H, D), turbulence_amp=amp, turbulence_scale=1.0, noise_seed=seed,
                 turb_mask_temp_ramp=(0.5, 2.5), turb_mask_vel_ramp=(0.0, 2.0))
    o = O.Oracle(s, fp64=True)
    T0, u0 = 1.0, (0.3, -0.4, 1.2)          # |u| = 1.3
    o.set_field(F_T, np.full(o.dims(F_T)[::-1], T0))
    for w, v in zip((F_U, F_V, F_W), u0):
        o.set_field(w, np.full(o.dims(w)[::-1], v))
    time = 1.0 / 32.0                       # offset = time * 8 = 0.25 exactly
    o.run_stage(STAGE_TURBULENCE, time, f32(1 / 24))
    mask = (1 - (T0 - 0.5) / 2.0) * (np.linalg.norm(u0) / 2.0)
    assert 0.2 < mask < 0.8
    off = 0.25
    for w, axis in ((F_U, 0), (F_V, 1), (F_W, 2)):
        back = o.field(w, back=True)
        nz, ny, nx = back.shape
        for k, j, i in [(0, 0, 0), (1, 2, 3), (nz - 1, ny - 1, nx - 1), (3, 5, 2)]:
            c = [i, j, k]
            c1 = list(c)
            c1[axis] += 1
            g0 = corner_grad(c, seed)[axis]
            g1 = corner_grad(c1, seed)[axis]
            noise = g0 * off * (1 - fade(off)) + g1 * (off - 1) * fade(off)
            want = mask * amp * noise
            assert abs(back[k, j, i] - want) < 1e-14, (w, (i, j, k), back[k, j, i], want)
        # and everywhere against the independent restatement
        kk, jj, ii = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        for k, j, i in zip(kk.ravel()[::7], jj.ravel()[::7], ii.ravel()[::7]):
            p = np.array([i, j, k], dtype=float)
            p[axis] += off
            assert abs(back[k, j, i] - mask * amp * perlin_ref(p, 1.0, seed)) < 1e-13


# -------------------------------------------------------------------- wind
@pytest.mark.parametrize("wind_dir", [0, 1, 2, 3])
def test_wind_turbulence_on_and_off_lattice(wind_dir):
    """add_wind_kernel (src/FluidSolver.cu:337-390) with windTurbAmp > 0.
    The noise position is (i dx, j dx, k dx + offset) * freq; with
    windTurbScale = dx the lattice points are the faces, so at time 0 the
    noise is 0 exactly and the border slabs take the closed form
    u (1 - dt) + amp dir dt. At time 1/64 (offset = dx/4 on z) every face
    takes tamp * the two-corner z blend of seeds 123 (U), 456 (V), 789 (W)."""
    W = H = D = 16
    amp, tamp = f32(0.4), 3.0
    s = settings((W, H, D), wind_dir=wind_dir, wind_amp=amp, wind_speed=1.0, wind_turb_amp=tamp,
                 wind_turb_scale=1.0 / 16)
    dt = f32(1 / 24)
    st = random_state(s, seed=22 + wind_dir)
    xwind = wind_dir <= 1
    dirv = (1.0 if wind_dir in (0, 2) else -1.0)
    for time, off in ((0.0, 0.0), (1.0 / 64.0, 0.25)):
        o = O.Oracle(s, fp64=True)
        for w, a in st.items():
            o.set_field(w, a)
        o.run_stage(STAGE_WIND, time, dt)
        for w, seed in ((F_U, 123), (F_V, 456), (F_W, 789)):
            got = o.field(w)
            a0 = st[w].astype(np.float64)
            nz, ny, nx = got.shape
            for k, j, i in itertools.product(range(nz), range(ny), range(nx)):
                border = (i < 2 or i >= W + 1 - 2) if xwind else (k < 2 or k >= D + 1 - 2)
                if not border:
                    assert got[k, j, i] == a0[k, j, i]
                    continue
                g0 = corner_grad((i, j, k), seed)[2]
                g1 = corner_grad((i, j, k + 1), seed)[2]
                noise = tamp * (g0 * off * (1 - fade(off)) + g1 * (off - 1) * fade(off))
                on_axis = (w == F_U and xwind) or (w == F_W and not xwind)
                want = a0[k, j, i] * (1 - dt) + (amp * ((dirv if on_axis else 0.0) + noise)) * dt
                assert abs(got[k, j, i] - want) < 1e-13, (time, w, (i, j, k), got[k, j, i], want)
            if off == 0.0:
                continue
            # the noise is not trivially zero off the lattice
            assert np.abs(got - a0 * (1 - dt)).max() > 1e-3


# --------------------------------------------------------------------- RK3
def test_rk3_backtrace_linear_velocity():
    """advect_RK3_kernel (src/FluidSolver.cu:431-456), Ralston's third-order
    scheme (weights 2/9, 3/9, 4/9 at stages 0, 1/2, 3/4): for u = a x,
    v = w = 0 every trilinear sample is exact and the back-trace lands at
    x0 (1 - h + h^2/2 - h^3/6), h = a dt (the cubic Taylor polynomial of
    exp(-h)). T = 1 + 2 x and U itself are linear, so the tricubic sample
    returns them at that point. Any other weights or stage fractions move
    the point by O(h^2)."""
    W, H, D = 32, 8, 8
    s = settings((W, H, D))
    o = O.Oracle(s, fp64=True)
    dx = s.dx
    a = 2.0
    dt = f32(1 / 24)
    h = a * dt
    r = 1 - h + h * h / 2 - h ** 3 / 6
    nxu = o.dims(F_U)[0]
    xu = np.arange(nxu) * dx
    o.set_field(F_U, np.broadcast_to(a * xu, o.dims(F_U)[::-1]).copy())
    o.set_field(F_V, np.zeros(o.dims(F_V)[::-1]))
    o.set_field(F_W, np.zeros(o.dims(F_W)[::-1]))
    xt = (np.arange(W) + 0.5) * dx
    o.set_field(F_T, np.broadcast_to(1 + 2 * xt, (D, H, W)).copy())
    o.run_stage(STAGE_ADVECT, 0.0, dt)
    T1 = o.field(F_T)
    # cells whose back-traced point keeps the 4-point stencil off the clamp
    gx = xt * r / dx - 0.5
    ok = (gx >= 1.0) & (gx <= W - 3)
    assert ok.sum() > 20
    want = 1 + 2 * xt * r
    assert np.abs(T1[:, :, ok] - want[ok]).max() < 1e-12
    U1 = o.field(F_U)
    gu = xu * r / dx
    # (the last face's own stage samples sit on U's sample clamp nxu - 1.001)
    oku = (gu >= 1.0) & (gu <= nxu - 3) & (xu / dx <= nxu - 1.001)
    assert np.abs(U1[:, :, oku] - a * xu[oku] * r).max() < 1e-12
    # the second-order truncation (1 - h + h^2/2) would differ by a x0 h^3/6
    assert np.abs(want[ok] - (1 + 2 * xt[ok] * (1 - h + h * h / 2))).min() > 1e-6
