"""Pins of the CPU oracle (the parity checker of the HIP path).

The reference has no tests, golden vectors or fixtures for the solver, so
each oracle function is pinned here by mathematics that a plausible
transcription mistake would break: closed forms, polynomial reproduction,
invariants, brute force against a textbook solver (scipy), and the
equivalence of the reference's per-slab Gauss-Seidel order with the
column order the GPU uses.
"""
import itertools
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(__file__))
from helpers import compare, random_state  # noqa: E402

O = pytest.importorskip("oracle")
B = pytest.importorskip("blaze_amd")

from helpers import settings as _settings  # noqa: E402


def settings(res=(16, 12, 10), **kw):
    return _settings(B, res, **kw)


# ---------------------------------------------------------------- cerp
def lagrange_weights(s):
    # Lagrange basis on nodes -1, 0, 1, 2 (independent closed form)
    return np.array([-s * (s - 1) * (s - 2) / 6, (s + 1) * (s - 1) * (s - 2) / 2,
                     -(s + 1) * s * (s - 2) / 2, (s + 1) * s * (s - 1) / 6])


def test_cerp_reproduces_cubics_fp64():
    """src/Grid.h:148-154: cubic Lagrange interpolation through nodes -1..2
    reproduces any cubic polynomial exactly."""
    rng = np.random.default_rng(0)
    for _ in range(200):
        c = rng.standard_normal(4)
        poly = lambda x: c[0] + c[1] * x + c[2] * x * x + c[3] * x ** 3  # noqa: E731
        q = [poly(x) for x in (-1.0, 0.0, 1.0, 2.0)]
        s = rng.uniform(0, 1)
        got = O.lib.orc_cerp_f64(*q, s)
        assert abs(got - poly(s)) <= 1e-12 * (1 + abs(poly(s)))
        w = lagrange_weights(s)
        assert abs(got - float(w @ np.array(q))) <= 1e-12 * (1 + abs(got))


def test_cerp_nodes_fp32():
    """At s = 0 the interpolant is q1 exactly (weights 0, 1, 0, 0)."""
    rng = np.random.default_rng(1)
    for _ in range(100):
        q = rng.standard_normal(4).astype(np.float32)
        assert np.float32(O.lib.orc_cerp_f32(*map(float, q), 0.0)) == q[1]


# --------------------------------------------------------- cubic pulse
def test_cubic_pulse_closed_form():
    """src/Grid.h:309-313: 1 - 3t^2 + 2t^3 for t = x/w <= 1, else 0."""
    w = 0.3
    for x in np.linspace(0, 0.45, 91):
        got = O.lib.orc_cubic_pulse_f32(float(x), w)
        t = x / w
        want = 0.0 if x > w else 1 - 3 * t * t + 2 * t ** 3
        assert abs(got - want) < 1e-6
    assert O.lib.orc_cubic_pulse_f32(0.0, w) == 1.0
    assert O.lib.orc_cubic_pulse_f32(w, w) == 0.0


# -------------------------------------------------------------- Perlin
def test_noise_hash_is_bijective_and_mixes():
    """The 32-bit integer mix of cuda_noise.cuh:46-56 is a composition of
    invertible steps: no collisions on a large input range; single-bit
    input flips change about half of the output bits (a dropped or wrong
    step degrades this)."""
    n = 1 << 16
    outs = np.array([O.lib.orc_noise_hash(i) for i in range(n)], dtype=np.uint64)
    assert len(np.unique(outs)) == n
    rng = np.random.default_rng(2)
    flips = []
    for _ in range(400):
        x = int(rng.integers(0, 2 ** 32))
        b = int(rng.integers(0, 32))
        d = O.lib.orc_noise_hash(x) ^ O.lib.orc_noise_hash(x ^ (1 << b))
        flips.append(bin(d).count("1"))
    assert 13 < np.mean(flips) < 19


def test_lattice_hash_float_sum_exact_for_small_lattice():
    """randomIntGrid (cuda_noise.cuh:120-123) sums in float; for small
    lattice points the float sum is exact and equals integer arithmetic."""
    for x, y, z in itertools.product(range(-2, 5), range(0, 4), range(0, 3)):
        for seed in (0, 7, 123):
            v = x * 1723 + y * 93241 + z * 149812 + 3824 + seed
            want = O.lib.orc_noise_hash(max(v, 0))  # negative saturates to 0
            assert O.lib.orc_lattice_hash(float(x), float(y), float(z), float(seed)) == want


def test_perlin_zero_on_lattice_and_bounded():
    """Gradient noise vanishes at lattice points and stays bounded."""
    for scale, seed in ((1.0, 123), (0.25, 7), (4.0, 456)):
        for x, y, z in itertools.product(range(0, 12, 3), range(0, 9, 4), range(0, 7, 2)):
            p = np.array([x, y, z]) / scale
            assert O.lib.orc_perlin_f64(*p, scale, seed) == 0.0
        rng = np.random.default_rng(seed)
        vals = [O.lib.orc_perlin_f64(*rng.uniform(0, 40, 3), scale, seed) for _ in range(2000)]
        assert max(abs(v) for v in vals) < 1.5
        assert np.std(vals) > 0.05


GRAD = {0x0: (1, 1, 0), 0x1: (-1, 1, 0), 0x2: (1, -1, 0), 0x3: (-1, -1, 0), 0x4: (1, 0, 1),
        0x5: (-1, 0, 1), 0x6: (1, 0, -1), 0x7: (-1, 0, -1), 0x8: (0, 1, 1), 0x9: (0, -1, 1),
        0xA: (0, 1, -1), 0xB: (0, -1, -1), 0xC: (1, 1, 0), 0xD: (0, -1, 1), 0xE: (-1, 1, 0),
        0xF: (0, -1, -1)}


def test_perlin_gradient_at_lattice_point():
    """Near a lattice corner only that corner's gradient term survives (the
    fade weights are flat there), so n(p0 + h e) / h equals the component of
    the hashed gradient vector (grad, cuda_noise.cuh:179-201, with its
    0xD / 0xF cases)."""
    h = 1e-4
    for corner in [(3, 5, 2), (10, 1, 7), (0, 0, 0)]:
        hv = O.lib.orc_lattice_hash(*map(float, corner), 123.0) & 0xF
        g = GRAD[hv]
        for a in range(3):
            p = np.array(corner, dtype=float)
            p[a] += h
            d = O.lib.orc_perlin_f64(*p, 1.0, 123) / h
            assert abs(d - g[a]) < 1e-3, (corner, a, d, g)


# ------------------------------------------------------------ sampling
def linear_field(s, which, a, b, c, d):
    nx, ny, nz = O.Oracle(s).dims(which)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return a * i + b * j + c * k + d


def test_trilinear_exact_for_linear_fields_and_clamps():
    """src/Grid.h:94-108: trilinear sampling reproduces linear fields inside
    the clamp box [0, N-1.001] and clamps grid coordinates outside it."""
    s = settings((16, 12, 10))
    o = O.Oracle(s, fp64=True)
    f = linear_field(s, 0, 0.5, -1.25, 2.0, 3.0)
    o.set_field(0, f)
    rng = np.random.default_rng(3)
    dx = s.dx
    for _ in range(200):
        g = rng.uniform([0, 0, 0], [14.9, 10.9, 8.9])
        pos = (g + 0.5) * dx           # T origin offset 0.5 cell
        want = 0.5 * g[0] - 1.25 * g[1] + 2.0 * g[2] + 3.0
        assert abs(o.sample(0, pos) - want) < 1e-9
    # outside: clamped to the grid coordinate box
    pos = (np.array([-3.0, 20.0, 4.0]) + 0.5) * dx
    want = 0.5 * 0 - 1.25 * (12 - 1.001) + 2.0 * 4.0 + 3.0
    assert abs(o.sample(0, pos) - want) < 1e-6


def test_tricubic_exact_for_cubic_fields():
    """src/Grid.h:165-222: tricubic (x, then y, then z) reproduces separable
    cubic fields where no stencil index is clamped; the clamped variant
    agrees on positive fields."""
    s = settings((16, 12, 10))
    o = O.Oracle(s, fp64=True)
    nx, ny, nz = o.dims(0)
    P = lambda x: 2 + 0.1 * x + 0.02 * x * x + 0.001 * x ** 3  # noqa: E731
    Q = lambda y: 1 + 0.3 * y + 0.01 * y * y + 0.0005 * y ** 3  # noqa: E731
    R = lambda z: 3 - 0.2 * z + 0.05 * z * z  # noqa: E731
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    o.set_field(0, P(i) * Q(j) * R(k))
    rng = np.random.default_rng(4)
    for _ in range(200):
        g = rng.uniform([1, 1, 1], [nx - 3, ny - 3, nz - 3])
        pos = (g + 0.5) * s.dx
        want = P(g[0]) * Q(g[1]) * R(g[2])
        assert abs(o.sample(0, pos, cubic=True) - want) < 1e-9 * abs(want)
        assert abs(o.sample(0, pos, cubic=True, clamp=True) - want) < 1e-9 * abs(want)


# ----------------------------------------------------------------- RBGS
def test_rbgs_column_order_equals_reference_slab_order():
    """gs_solve_kernel (src/FluidSolver.cu:526-598) updates, per slab k, all
    red cells then all black cells. Red cells read only black neighbours in
    their slab and their own column, so this equals red columns sweeping
    k = 0..D-1 followed by black columns: bit-identical in FP32."""
    for res, cb in [((16, 12, 10), (1, 1, 1, 1, 1, 1)), ((10, 14, 8), (0, 1, 1, 0, 1, 0))]:
        s = settings(res, max_iter_solve=7, closed_bounds=cb)
        st = random_state(s, seed=5)
        a, b = O.Oracle(s), O.Oracle(s)
        b.set_interleaved_rbgs(True)
        for w in (5, 6):
            a.set_field(w, st[w])
            b.set_field(w, st[w])
        a.run_stage(7, 0.0, 1 / 24)
        b.run_stage(7, 0.0, 1 / 24)
        compare("rbgs order", a.field(5), b.field(5))


def laplacian(W, H, D):
    import scipy.sparse as sp

    n = W * H * D
    idx = lambda i, j, k: i + j * W + k * W * H  # noqa: E731
    rows, cols, vals = [], [], []
    for k in range(D):
        for j in range(H):
            for i in range(W):
                c = idx(i, j, k)
                cnt = 0
                for di, dj, dk in ((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)):
                    a, b_, e = i + di, j + dj, k + dk
                    if 0 <= a < W and 0 <= b_ < H and 0 <= e < D:
                        rows.append(c)
                        cols.append(idx(a, b_, e))
                        vals.append(-1.0)
                        cnt += 1
                rows.append(c)
                cols.append(c)
                vals.append(float(cnt))
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


def test_rbgs_converges_to_scipy_solution():
    """The Gauss-Seidel iteration solves n p - sum(p_nbr) = (rho dx^2/dt) rhs
    with Neumann walls (:533-561): after many FP64 sweeps on a consistent
    right-hand side, p matches a sparse least-squares solve up to a
    constant."""
    import scipy.sparse.linalg as spla

    W, H, D = 8, 6, 6
    s = settings((W, H, D), max_iter_solve=3000, density=1.3)
    rng = np.random.default_rng(6)
    rhs = rng.standard_normal((D, H, W))
    rhs -= rhs.mean()
    o = O.Oracle(s, fp64=True)
    o.set_field(6, rhs)
    o.set_field(5, np.zeros_like(rhs))
    dt = 1 / 24
    o.run_stage(7, 0.0, dt)
    p = o.field(5).ravel()
    scale = (np.float64(np.float32(1.3)) * s.dx * s.dx) / np.float64(np.float32(dt))
    A = laplacian(W, H, D)
    want = spla.lsqr(A, scale * rhs.ravel(), atol=1e-14, btol=1e-14, iter_lim=20000)[0]
    p = p - p.mean()
    want = want - want.mean()
    # the oracle computes scale from the float32 settings and dt
    assert np.abs(p - want).max() < 1e-8 * max(1.0, np.abs(want).max())


def test_residual_equals_post_update_divergence():
    """The FP64 residual is the post-update divergence: with all walls
    closed, |rhs + c (sum p_nbr - n p)| equals |div u| of the velocity after
    pressure_gradient_update_kernel (:601-669), computed independently from
    the face velocities, to FP32 rounding."""
    s = settings((12, 10, 8), max_iter_solve=15, residual_tol=1e-30, residual_check_every=15)
    st = random_state(s, seed=7)
    o = O.Oracle(s)
    for w in (1, 2, 3):
        o.set_field(w, st[w])
    dt = 1 / 24
    o.run_stage(6, 0.0, dt)   # divergence
    o.run_stage(7, 0.0, dt)   # 15 sweeps, residual reported
    res = o.stats()["residual"]
    o.run_stage(8, 0.0, dt)   # velocity update
    U, V, Wf = (o.field(w).astype(np.float64) for w in (1, 2, 3))
    div = (U[:, :, 1:] - U[:, :, :-1] + V[:, 1:, :] - V[:, :-1, :] + Wf[1:, :, :] - Wf[:-1, :, :]) / s.dx
    assert res > 0
    assert abs(np.abs(div).max() - res) < 1e-4 * res + 1e-3


def test_projection_closed_box_invariants():
    """Closed box: wall faces are exactly 0 after the update; the interior
    divergence falls as sweeps increase."""
    out = []
    for iters in (5, 60):
        s = settings((12, 10, 8), max_iter_solve=iters)
        st = random_state(s, seed=8)
        o = O.Oracle(s, fp64=True)
        for w in (1, 2, 3):
            o.set_field(w, st[w])
        o.set_field(5, np.zeros_like(st[5]))
        for stage in (6, 7, 8):
            o.run_stage(stage, 0.0, 1 / 24)
        U, V, Wf = o.field(1), o.field(2), o.field(3)
        assert not U[:, :, 0].any() and not U[:, :, -1].any()
        assert not V[:, 0, :].any() and not V[:, -1, :].any()
        assert not Wf[0].any() and not Wf[-1].any()
        div = (U[:, :, 1:] - U[:, :, :-1] + V[:, 1:, :] - V[:, :-1, :] + Wf[1:] - Wf[:-1]) / s.dx
        out.append(np.abs(div).max())
    assert out[1] < 0.5 * out[0]


def test_divergence_closed_form_linear_velocity():
    """compute_divergence_kernel (:475-515): for u = a x, v = b y, w = c z the
    interior rhs is -(a+b+c); a closed wall face is treated as zero flux."""
    s = settings((10, 8, 6), closed_bounds=(1, 0, 0, 1, 0, 0))
    o = O.Oracle(s, fp64=True)
    a, b, c = 1.5, -0.5, 2.0
    dx = s.dx
    for w, coef in ((1, a), (2, b), (3, c)):
        nx, ny, nz = o.dims(w)
        k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        g = {1: i, 2: j, 3: k}[w]
        o.set_field(w, coef * g * dx)
    o.run_stage(6, 0.0, 1 / 24)
    rhs = o.field(6)
    assert np.allclose(rhs[1:-1, 1:-1, 1:-1], -(a + b + c), atol=1e-12)
    # xMin closed (cb[0]): the U(0) = 0 face term removed -> -(a*1dx/dx) ... at i = 0 U(0) = 0 anyway
    # yMax closed (cb[3]) is grid j = 0: the V(0) face is dropped; V(0) = 0 here as well
    # xMax open: cell W-1 keeps -(a+b+c)
    assert np.allclose(rhs[1:-1, 1:-1, -1], -(a + b + c), atol=1e-12)


# --------------------------------------------------------------- forces
def test_cooling_drag_buoyancy_closed_forms():
    """:118-156: T *= 1 - dt c; U,V,W *= 1 - dt drag; V += dt (beta T~ g),
    T~ = T for a uniform T field."""
    s = settings((8, 8, 8), cooling_rate=0.9, drag=0.5, buoyancy=0.7)
    st = random_state(s, seed=9)
    o = O.Oracle(s)
    for w, a in st.items():
        o.set_field(w, a)
    dt = np.float32(1 / 24)
    o.run_stage(0, 0.0, dt)
    compare("cool", o.field(0), st[0] * np.float32(1 - dt * np.float32(0.9)))
    o.run_stage(1, 0.0, dt)
    f = np.float32(1 - dt * np.float32(0.5))
    for w in (1, 2, 3):
        compare(f"drag{w}", o.field(w), st[w] * f)
    o.set_field(0, np.full_like(st[0], 2.0))
    v0 = o.field(2)
    o.run_stage(3, 0.0, dt)
    force = np.float32(np.float32(0.7) * np.float32(2.0)) * np.float32(s.gravity)
    compare("buoyancy", o.field(2), v0 + dt * force)


def test_vorticity_zero_for_uniform_velocity():
    """:203-248: a uniform velocity has no vorticity, so confinement adds
    nothing."""
    s = settings((12, 12, 12), vorticity_conf=0.5)
    o = O.Oracle(s, fp64=True)
    for w, v in ((1, 0.3), (2, -0.2), (3, 0.7)):
        o.set_field(w, np.full(o.dims(w)[::-1], v))
    o.run_stage(2, 0.0, 1 / 24)
    for w, v in ((1, 0.3), (2, -0.2), (3, 0.7)):
        assert np.abs(o.field(w) - v).max() < 1e-15


def test_vorticity_solid_rotation_has_no_confinement():
    """Solid rotation about the z axis has constant |omega|, so its gradient
    (and the confinement force) vanishes away from the clamped border."""
    s = settings((16, 16, 8), vorticity_conf=0.5)
    o = O.Oracle(s, fp64=True)
    dx = s.dx
    Om = 2.0
    c = 8 * dx
    nx, ny, nz = o.dims(1)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    U0 = -Om * ((j + 0.5) * dx - c)
    nx, ny, nz = o.dims(2)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    V0 = Om * ((i + 0.5) * dx - c)
    o.set_field(1, U0)
    o.set_field(2, V0)
    o.set_field(3, np.zeros(o.dims(3)[::-1]))
    o.run_stage(2, 0.0, 1 / 24)
    dU = o.field(1) - U0
    dV = o.field(2) - V0
    assert np.abs(dU[3:-3, 3:-3, 3:-3]).max() < 1e-9
    assert np.abs(dV[3:-3, 3:-3, 3:-3]).max() < 1e-9


def test_wind_without_turbulence_closed_form():
    """:337-390 with windTurbAmp = 0: border slabs blend u <- u(1-dt) +
    amp dir dt on the wind axis and decay the other components; the
    interior is untouched."""
    s = settings((12, 8, 8), wind_dir=1, wind_amp=0.4, wind_speed=0.2, wind_turb_amp=0.0,
                 wind_turb_scale=0.5)
    st = random_state(s, seed=10)
    o = O.Oracle(s, fp64=True)
    for w, a in st.items():
        o.set_field(w, a)
    dt = np.float64(np.float32(1 / 24))
    amp = np.float64(np.float32(0.4))
    o.run_stage(4, 0.0, dt)
    U = o.field(1)
    U0 = st[1].astype(np.float64)
    for i in (0, 1, 11, 12):
        assert np.allclose(U[:, :, i], U0[:, :, i] * (1 - dt) + (amp * -1.0) * dt, atol=1e-12)
    assert np.array_equal(U[:, :, 2:11], U0[:, :, 2:11])


def test_turbulence_masked_out_changes_nothing():
    """:295-335: with the temperature ramp below T everywhere the mask is 0,
    the noise field is 0, and the curl step adds 0."""
    s = settings((10, 10, 10), turbulence_amp=2.0, turbulence_scale=0.5,
                 turb_mask_temp_ramp=(0.0, 0.5), turb_mask_vel_ramp=(0.0, 1.0))
    st = random_state(s, seed=11)
    st[0] = np.abs(st[0]) + 1.0
    o = O.Oracle(s, fp64=True)
    for w, a in st.items():
        o.set_field(w, a)
    o.run_stage(5, 1.0, 1 / 24)
    for w in (1, 2, 3):
        assert np.array_equal(o.field(w), st[w].astype(np.float64))
        assert not o.field(w, back=True).any()


# ------------------------------------------------------------ advection
def test_advection_uniform_integer_shift_fp64():
    """:431-456: a uniform velocity makes the RK3 back-trace pos - dt v; with
    dt v an integer number of cells the tricubic sample hits a node, so T is
    the shifted field. The last cell of each axis is excluded: the sample
    clamp N - 1.001 (src/Grid.h:167-169) moves it off the node."""
    s = settings((16, 12, 10))
    o = O.Oracle(s, fp64=True)
    rng = np.random.default_rng(12)
    T0 = rng.uniform(0.5, 1.5, (10, 12, 16))
    o.set_field(0, T0)
    dt = np.float64(np.float32(1 / 24))
    vx = 2 * s.dx / dt
    o.set_field(1, np.full(o.dims(1)[::-1], vx))
    o.set_field(2, np.zeros(o.dims(2)[::-1]))
    o.set_field(3, np.zeros(o.dims(3)[::-1]))
    o.run_stage(9, 0.0, dt)
    T1 = o.field(0)
    assert np.abs(T1[:-1, :-1, 2:-1] - T0[:-1, :-1, :-3]).max() < 1e-9


# ----------------------------------------------------------------- P2G
def test_p2g_single_particle_closed_form():
    """sphereToGrid (src/Grid.h:315-335) + set_source (:99-116): one particle
    adds value*mask to cells within pscale/2 of its centre (V negated: the
    grid y axis is flipped), and the front is overwritten only where
    |sum| > 1e-5."""
    s = settings((16, 16, 16))
    o = O.Oracle(s, fp64=True)
    base = {w: np.full(o.dims(w)[::-1], 7.0) for w in (0, 1, 2, 3)}
    for w, a in base.items():
        o.set_field(w, a)
    p = np.array([[0.52, 0.47, 0.55, 1.0, 2.0, 0.0, 5 * s.dx, 3.0]], np.float32)
    o.add_source(p)
    T = o.field(0)
    dx = s.dx
    r = np.float64(np.float32(p[0, 6]) / 2)
    k, j, i = np.meshgrid(np.arange(16), np.arange(16), np.arange(16), indexing="ij")
    pos = np.array([p[0, 0], 1.0 - np.float64(p[0, 1]), p[0, 2]], dtype=np.float64)  # object space (y flipped)
    d = np.sqrt(((i + 0.5) * dx - pos[0]) ** 2 + ((j + 0.5) * dx - pos[1]) ** 2 + ((k + 0.5) * dx - pos[2]) ** 2)
    t = d / r
    mask = np.where(d > r, 0.0, 1 - 3 * t * t + 2 * t ** 3)
    want = np.where(3.0 * mask > 1e-5, 3.0 * mask, 7.0)
    assert np.abs(T - want).max() < 1e-12
    V = o.field(2)
    assert (V < 0).any() and V.min() >= -2.0 and (V[V != 7.0] <= 0).all()
    W = o.field(3)
    assert np.array_equal(W, base[3])   # vz = 0: channel skipped


# ------------------------------------------------------ precision, order
def test_fp32_oracle_close_to_fp64():
    """Same frames in FP32 and FP64: the FP32 rounding error stays small
    (the FP64 run is the accuracy reference of the FP32 path)."""
    s, _, _, gen = B.parse_scene(os.path.join(os.path.dirname(__file__), "..", "scenes", "plume64.blz"))
    s.grid_res[0] = s.grid_res[1] = s.grid_res[2] = 24
    s.dx = 1.0 / 24
    s.max_iter_solve = 40
    gen.count = 3000
    gen.pscale = 0.75 / 24
    parts = B.generate_particles(gen)
    a, b = O.Oracle(s), O.Oracle(s, fp64=True)
    t = np.float32(0)
    for it in range(3):
        a.add_source(parts if it == 0 else None)
        b.add_source(parts if it == 0 else None)
        a.step(float(t))
        b.step(float(t))
        t = np.float32(t + np.float32(s.dt))
    for w in (0, 1, 2, 3):
        x, y = a.field(w).astype(np.float64), b.field(w)
        rel = np.linalg.norm(x - y) / (np.linalg.norm(y) + 1e-30)
        assert rel < 1e-4, (w, rel)


def test_openmp_matches_serial():
    """The oracle's threaded loops are order-free (gathers, per-column RBGS):
    1 thread and all threads give identical bits."""
    s, _, _, gen = B.parse_scene(os.path.join(os.path.dirname(__file__), "..", "scenes", "plume128.blz"))
    s.grid_res[0] = s.grid_res[1] = s.grid_res[2] = 20
    s.dx = 1.0 / 20
    s.max_iter_solve = 10
    gen.count = 2000
    gen.pscale = 1.5 / 20
    parts = B.generate_particles(gen)
    outs = []
    for threads in (1, max(2, O.max_threads())):
        O.set_threads(threads)
        o = O.Oracle(s)
        o.add_source(parts)
        o.step(0.0)
        o.step(float(np.float32(s.dt)))
        outs.append([o.field(w) for w in (0, 1, 2, 3, 5)])
    O.set_threads(os.cpu_count() or 1)
    for x, y in zip(*outs):
        compare("omp", x, y)


def test_cfl_substep_count():
    """n = ceil(max|u| dt / (cfl dx)) clamped to [1, max_substeps]."""
    s = settings((8, 8, 8), cfl_number=2.0, max_substeps=5, max_iter_solve=1)
    for umax, want in ((0.0, 1), (2 * 2.0 * s.dx / s.dt * 0.99, 2), (2 * 2.0 * s.dx / s.dt * 1.01, 3),
                       (1e3, 5)):
        o = O.Oracle(s)
        U = np.zeros(o.dims(1)[::-1])
        U[3, 3, 3] = -umax
        o.set_field(1, U)
        o.step(0.0)
        assert o.stats()["substeps"] == want, (umax, o.stats())


def _poisson_case(res, seed, **kw):
    s = settings(res, closed_bounds=(1, 1, 1, 1, 1, 1), residual_check_every=10, **kw)
    W, H, D = res
    rng = np.random.default_rng(seed)
    rhs = rng.standard_normal((D, H, W)).astype(np.float32)
    rhs -= np.float32(rhs.mean())
    return s, rhs


def test_adaptive_check_spacing_reaches_the_tolerance():
    """residual_check_adaptive: the solve still ends at the first evaluated
    residual <= tol, with few residual evaluations, and no earlier than the
    every-10-sweeps solve (the checks it skips cannot end it sooner)."""
    s, rhs = _poisson_case((24, 20, 16), 3, max_iter_solve=20000, residual_tol=1e-3)
    runs = {}
    for adaptive in (0, 1):
        s.residual_check_adaptive = adaptive
        o = O.Oracle(s)
        o.set_field(6, rhs)
        o.set_field(5, np.zeros_like(rhs))
        o.run_stage(7, 0.0, 1 / 24)
        runs[adaptive] = o.stats()
    fixed, adap = runs[0], runs[1]
    print(f"[adaptive] fixed {fixed['rbgs_iters_last']} sweeps, adaptive {adap['rbgs_iters_last']}", flush=True)
    assert fixed["residual"] <= 1e-3 and adap["residual"] <= 1e-3
    assert fixed["rbgs_iters_last"] <= adap["rbgs_iters_last"] <= 2 * fixed["rbgs_iters_last"]
    assert adap["rbgs_iters_last"] % 10 == 0


def test_fixed_point_stop_is_a_fixed_point():
    """With an unreachable tolerance the adaptive solve ends well before its
    cap once the residual has not changed over 60 sweeps, and there the FP32
    iteration is indeed stationary: more sweeps leave the pressure bit for
    bit unchanged."""
    s, rhs = _poisson_case((8, 8, 8), 5, max_iter_solve=200000, residual_tol=1e-30)
    s.residual_check_adaptive = 1
    o = O.Oracle(s)
    o.set_field(6, rhs)
    o.set_field(5, np.zeros_like(rhs))
    o.run_stage(7, 0.0, 1 / 24)
    st = o.stats()
    print(f"[stall] stopped after {st['rbgs_iters_last']} sweeps at residual {st['residual']:.3e}", flush=True)
    assert st["rbgs_iters_last"] < 200000 and st["residual"] > 1e-30
    p = o.field(5)
    s2 = s.copy()
    s2.residual_check_adaptive = 0
    s2.residual_tol = 0.0
    s2.max_iter_solve = 100
    o2 = O.Oracle(s2)
    o2.set_field(6, rhs)
    o2.set_field(5, p)
    o2.run_stage(7, 0.0, 1 / 24)
    compare("fixed point", o2.field(5), p)
