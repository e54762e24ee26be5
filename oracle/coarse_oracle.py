"""CPU oracle of the warm start's global correction (test infrastructure
only: the product never imports it). Restates, in float64 numpy, what
blaze_amd/csrc/k_warm.hip does on the GPU after the network's correction:

  r  = scale*rhs + sum_nbr p - n p           (A = graph Laplacian, Neumann)
  b  = block sums of r over blocks of F^3 cells (clipped at the far faces)
  b -= (sum b / (W H D)) * |block|           (drop the constant part of r)
  A_G x = b, A_G = P^T A P for piecewise-constant P: the graph Laplacian of
         the block grid with edge weight = fine faces shared by two blocks
  e  = trilinear prolongation of x (block centres), clamped at the ends
  p += alpha (e - mean e), alpha = <r, e - mean e> / <e, A e>
                                            (least error A-norm; the constant
                                            part changes no residual)

There is no reference counterpart (the reference starts each solve from the
previous pressure). The oracle is pinned by its own tests: the Galerkin
operator against P^T A P built explicitly, the coarse solve against a dense
least-squares solve, and, with F = 1 (block grid = fine grid), the whole
correction against the exact solution (the residual of a consistent
problem vanishes after it).

Arrays are (D, H, W), x fastest, like the library's natural order.
"""
import numpy as np


def laplacian(p):
    """A p for the graph Laplacian with Neumann ends: n p - sum_nbr p."""
    p = np.asarray(p, np.float64)
    a = np.zeros_like(p)
    for ax in range(3):
        d = np.diff(p, axis=ax)
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[ax] = slice(0, -1)
        hi[ax] = slice(1, None)
        a[tuple(lo)] -= d
        a[tuple(hi)] += d
    return a


def residual(p, rhs, scale):
    return scale * np.asarray(rhs, np.float64) - laplacian(p)


def block_grid(shape, F):
    D, H, W = shape
    return (-(-D // F), -(-H // F), -(-W // F))


def block_sums(r, F):
    D, H, W = r.shape
    Gd, Gh, Gw = block_grid(r.shape, F)
    out = np.zeros((Gd, Gh, Gw))
    for K in range(Gd):
        for J in range(Gh):
            for I in range(Gw):
                out[K, J, I] = r[F * K:F * K + F, F * J:F * J + F, F * I:F * I + F].sum()
    return out


def block_volumes(shape, F):
    D, H, W = shape
    ext = [np.minimum(F, n - F * np.arange(g)) for n, g in zip((D, H, W), block_grid(shape, F))]
    return ext[0][:, None, None] * ext[1][None, :, None] * ext[2][None, None, :], ext


def galerkin_apply(x, shape, F):
    """A_G x with edge weights = shared fine faces."""
    _, (ez, ey, ex) = block_volumes(shape, F)
    w = [ey[None, :, None] * ex[None, None, :],   # faces between z-neighbours
         ez[:, None, None] * ex[None, None, :],   # y
         ez[:, None, None] * ey[None, :, None]]   # x
    x = np.asarray(x, np.float64)
    a = np.zeros_like(x)
    for ax in range(3):
        d = np.diff(x, axis=ax)
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[ax] = slice(0, -1)
        hi[ax] = slice(1, None)
        wl = np.broadcast_to(w[ax], x.shape)[tuple(lo)]
        a[tuple(lo)] -= wl * d
        a[tuple(hi)] += wl * d
    return a


def coarse_rhs(r, F):
    vol, _ = block_volumes(r.shape, F)
    b = block_sums(r, F)
    return b - b.sum() / r.size * vol


def coarse_solve(b, shape, F, iters=400, rtol2=1e-8):
    """Conjugate gradients from 0 (as the kernel: stop at |r|^2 <= rtol2 |r0|^2)."""
    x = np.zeros_like(b)
    r = b.copy()
    d = r.copy()
    rr = float((r * r).sum())
    stop = rr * rtol2
    for _ in range(iters):
        if not (rr > stop and rr > 0.0):
            break
        ad = galerkin_apply(d, shape, F)
        pap = float((d * ad).sum())
        if not pap > 0.0:
            break
        al = rr / pap
        x += al * d
        r -= al * ad
        rn = float((r * r).sum())
        d = r + (rn / rr) * d
        rr = rn
    return x


def prolong(x, shape, F):
    """Trilinear at block centres (i + 0.5)/F - 0.5, clamped to the grid."""
    out = np.asarray(x, np.float64)
    for ax, n in enumerate(shape):
        g = out.shape[ax]
        f = np.clip((np.arange(n) + 0.5) / F - 0.5, 0.0, g - 1)
        a0 = f.astype(int)
        a1 = np.minimum(a0 + 1, g - 1)
        w = f - a0
        sh = [1, 1, 1]
        sh[ax] = -1
        w = w.reshape(sh)
        out = np.take(out, a0, axis=ax) * (1.0 - w) + np.take(out, a1, axis=ax) * w
    return out


def energy_alpha(r, e):
    ae = laplacian(e)
    den = float((e * ae).sum())
    if not den > 0.0:
        return 0.0
    return float((r * (e - e.mean())).sum()) / den


def global_correction(p, rhs, scale, F, iters=400, rtol2=1e-8):
    """The corrected pressure and the scale used."""
    r = residual(p, rhs, scale)
    x = coarse_solve(coarse_rhs(r, F), r.shape, F, iters, rtol2)
    e = prolong(x, r.shape, F)
    al = energy_alpha(r, e)
    return np.asarray(p, np.float64) + al * (e - e.mean()), al


def block_size(shape, cap=16):
    """Smallest power of two F with every axis at most cap * F."""
    F = 1
    while max(shape) > cap * F:
        F *= 2
    return F
