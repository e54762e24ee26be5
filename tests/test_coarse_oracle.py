"""CPU pins of the global-correction oracle (oracle/coarse_oracle.py): each
check is fixed by mathematics independent of the oracle's own formulas."""
import itertools

import numpy as np
import pytest

C = pytest.importorskip("oracle.coarse_oracle")


def inject(x, shape, F):
    """Piecewise-constant prolongation P (each block's value on its cells)."""
    D, H, W = shape
    out = np.repeat(np.repeat(np.repeat(x, F, 0), F, 1), F, 2)
    return out[:D, :H, :W]


@pytest.mark.parametrize("shape,F", [((10, 7, 5), 4), ((9, 6, 5), 2), ((8, 8, 8), 1)])
def test_galerkin_operator_is_PtAP(shape, F):
    """A_G x equals P^T A P x with P built by repetition and P^T by block
    sums, on grids whose far blocks are clipped."""
    rng = np.random.default_rng(1)
    G = C.block_grid(shape, F)
    x = rng.standard_normal(G)
    ref = C.block_sums(C.laplacian(inject(x, shape, F)), F)
    np.testing.assert_allclose(C.galerkin_apply(x, shape, F), ref, rtol=1e-12, atol=1e-12)


def test_coarse_solve_matches_dense_pseudo_inverse():
    shape, F = (9, 6, 5), 2
    G = C.block_grid(shape, F)
    n = int(np.prod(G))
    A = np.zeros((n, n))
    for q in range(n):
        u = np.zeros(n)
        u[q] = 1.0
        A[:, q] = C.galerkin_apply(u.reshape(G), shape, F).ravel()
    assert np.allclose(A, A.T)
    rng = np.random.default_rng(2)
    r = rng.standard_normal(shape)
    b = C.coarse_rhs(r, F)
    assert abs(b.sum()) < 1e-10 * np.abs(b).sum()   # consistent: orthogonal to the null space
    x = C.coarse_solve(b, shape, F, iters=5000, rtol2=1e-30)
    xd = np.linalg.pinv(A) @ b.ravel()
    np.testing.assert_allclose(x.ravel() - x.mean(), xd - xd.mean(), rtol=1e-6, atol=1e-9 * np.abs(xd).max())


def test_unit_blocks_give_the_exact_solution():
    """F = 1: the block grid is the fine grid, so the correction solves the
    (consistent) problem exactly and its scale is 1."""
    rng = np.random.default_rng(3)
    shape = (12, 10, 9)
    rhs = rng.standard_normal(shape)
    rhs -= rhs.mean()
    p = rng.standard_normal(shape)
    scale = 0.37
    r0 = np.abs(C.residual(p, rhs, scale)).max()
    pc, al = C.global_correction(p, rhs, scale, 1, iters=5000, rtol2=1e-30)
    assert abs(al - 1.0) < 1e-8
    assert np.abs(C.residual(pc, rhs, scale)).max() < 1e-8 * r0


def test_energy_scale_minimises_error_energy():
    """With a known solution p*, the scale minimises ||p* - p - a e||_A over a
    brute-force grid of a."""
    rng = np.random.default_rng(4)
    shape = (10, 8, 6)
    ps = rng.standard_normal(shape)
    scale = 1.0
    rhs = C.laplacian(ps) / scale
    p = ps + rng.standard_normal(shape)
    e = rng.standard_normal(shape)
    r = C.residual(p, rhs, scale)
    al = C.energy_alpha(r, e)

    def energy(a):
        d = ps - p - a * e
        return float((d * C.laplacian(d)).sum())

    grid = np.linspace(al - 1.0, al + 1.0, 2001)
    best = grid[np.argmin([energy(a) for a in grid])]
    assert abs(best - al) <= 1e-3
    assert energy(al) <= energy(0.0)


def test_prolongation_reproduces_linear_functions():
    """Block values of a linear function (at block centres) prolongate to the
    same linear function on every fine cell between the first and last
    block centres."""
    shape, F = (16, 12, 8), 4
    G = C.block_grid(shape, F)
    cz, cy, cx = [(np.arange(g) + 0.5) * F - 0.5 for g in G]
    lin = lambda z, y, x: 0.3 * z - 1.2 * y + 2.0 * x + 0.7
    x = lin(cz[:, None, None], cy[None, :, None], cx[None, None, :])
    out = C.prolong(x, shape, F)
    for k, j, i in itertools.product(range(shape[0]), range(shape[1]), range(shape[2])):
        if cz[0] <= k <= cz[-1] and cy[0] <= j <= cy[-1] and cx[0] <= i <= cx[-1]:
            assert abs(out[k, j, i] - lin(k, j, i)) < 1e-9


def test_block_size_rule():
    assert C.block_size((256, 256, 256)) == 16
    assert C.block_size((128, 64, 32)) == 8
    assert C.block_size((16, 9, 3)) == 1
    assert C.block_size((17, 4, 4)) == 2


def test_dct_training_target_solves_the_coarse_problem():
    """The network's training target (blaze_amd/warm_start.py
    coarse_target_dct): on the coarse grid, n E - sum_nbr E equals the
    input's residual channel minus its mean, up to the unit-RMS scale, for
    uneven grid sizes and a padded network grid."""
    W = pytest.importorskip("blaze_amd.warm_start")
    rng = np.random.default_rng(4)
    for (wc, hc, dc), (wn, hn, dn) in (((8, 6, 5), (8, 8, 8)), ((16, 16, 16), (16, 16, 16))):
        x = np.zeros((2, dn, hn, wn), np.float32)
        x[0, :dc, :hc, :wc] = rng.standard_normal((dc, hc, wc))
        e = W.coarse_target_dct(x, (wc, hc, dc)).astype(np.float64)[:dc, :hc, :wc]
        lap = C.laplacian(e)
        r = x[0, :dc, :hc, :wc].astype(np.float64)
        r = r - r.mean()
        k = float((lap * r).sum() / (lap * lap).sum())
        assert k > 0
        assert np.abs(k * lap - r).max() < 1e-4 * np.abs(r).max()
        assert abs(e.mean()) < 1e-6 and abs(np.sqrt((e * e).mean()) - 1.0) < 1e-5


@pytest.mark.parametrize("shape", [(10, 8, 6), (7, 5, 9)])
def test_scale_after_a_first_correction_by_symmetry(shape):
    """The network's scale is taken against r' = r - a1 A e1 without forming
    A e1 per cell: <r', e2> = <r, e2> - a1 <e1, A e2> (A symmetric) and
    sum r' = sum r (A e1 sums to zero over the grid). Both identities hold
    for the Neumann graph Laplacian; a sign or transpose slip fails here."""
    rng = np.random.default_rng(11)
    r = rng.standard_normal(shape)
    e1 = rng.standard_normal(shape)
    e2 = rng.standard_normal(shape)
    a1 = 0.37
    rp = r - a1 * C.laplacian(e1)
    direct = float((rp * e2).sum())
    via_sym = float((r * e2).sum()) - a1 * float((e1 * C.laplacian(e2)).sum())
    assert abs(direct - via_sym) <= 1e-10 * (abs(direct) + 1.0)
    assert abs(float(C.laplacian(e1).sum())) <= 1e-10 * float(np.abs(e1).sum())
    assert abs(float(rp.sum()) - float(r.sum())) <= 1e-10 * float(np.abs(r).sum())
    # and the resulting energy scale equals the one taken against r' itself
    den = float((e2 * C.laplacian(e2)).sum())
    n = e2.size
    al_sym = (via_sym - float(r.sum()) * float(e2.sum()) / n) / den
    assert abs(al_sym - C.energy_alpha(rp, e2)) <= 1e-10 * (abs(al_sym) + 1.0)
