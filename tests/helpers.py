"""Shared helpers of the test suite (settings builders, random states,
comparison with a readable failure message)."""
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SCENES = os.path.join(ROOT, "scenes")

FIELD_NAMES = ["T", "U", "V", "W", "density", "P", "rhs"]


def settings(blaze, res=(32, 24, 20), **kw):
    """blz_settings for a unit-ish box of the given resolution."""
    s = blaze.default_settings()
    W, H, D = res
    s.grid_res[0], s.grid_res[1], s.grid_res[2] = W, H, D
    s.dx = 1.0 / W
    s.dt = 1.0 / 24.0
    for a, n in enumerate(res):
        s.domain_min[a] = 0.0
        s.domain_max[a] = n * s.dx
    for k, v in kw.items():
        if isinstance(v, (list, tuple)):
            arr = getattr(s, k)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(s, k, v)
    return s


def dims(s, which):
    W, H, D = s.grid_res
    return {0: (W, H, D), 1: (W + 1, H, D), 2: (W, H + 1, D), 3: (W, H, D + 1), 4: (W, H, D),
            5: (W, H, D), 6: (W, H, D)}[which]


def random_state(s, seed=0, vel_scale=1.0, smoke=False):
    """Smooth-ish random fields: T >= 0, velocities of magnitude ~vel_scale."""
    rng = np.random.default_rng(seed)
    out = {}
    for w in (0, 1, 2, 3, 5, 6) + ((4,) if smoke else ()):
        nx, ny, nz = dims(s, w)
        a = rng.standard_normal((nz, ny, nx)).astype(np.float32)
        if w in (0, 4):
            a = np.abs(a)
        elif w in (1, 2, 3):
            a = a * np.float32(vel_scale)
        out[w] = a
    return out


def compare(name, got, want, rtol_l2=0.0):
    """Bitwise (rtol_l2 == 0) or relative-L2 comparison with a useful message."""
    got = np.asarray(got)
    want = np.asarray(want)
    assert got.shape == want.shape, f"{name}: shape {got.shape} vs {want.shape}"
    if rtol_l2 == 0.0:
        same = (got.view(np.uint32) == want.astype(np.float32).view(np.uint32)) | (
            np.isnan(got) & np.isnan(want))
        if not same.all():
            bad = np.argwhere(~same)
            k = tuple(bad[0])
            diff = np.abs(got.astype(np.float64) - want.astype(np.float64))
            raise AssertionError(
                f"{name}: {len(bad)} of {got.size} values differ; first at (k,j,i)={k}: "
                f"got {got[k]!r} want {want[k]!r}; max |diff| {np.nanmax(diff):.3e}")
        return 0.0
    num = np.linalg.norm((got.astype(np.float64) - want.astype(np.float64)).ravel())
    den = np.linalg.norm(want.astype(np.float64).ravel()) + 1e-30
    rel = num / den
    assert rel <= rtol_l2, f"{name}: relative L2 {rel:.3e} > {rtol_l2:.1e}"
    return rel
