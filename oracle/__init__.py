"""CPU oracle of the Blaze solver step — TEST INFRASTRUCTURE ONLY.

ctypes access to ``oracle/liboracle.so`` (a plain C++ restatement of the
reference kernels, see blaze_oracle.hpp). Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg may import
this package; the product (``blaze_amd``) never calls it.

The oracle takes the same C ``blz_settings`` struct as the product; it is
declared here again (ctypes) so that this package does not import the
product.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")


class _Settings(C.Structure):
    _fields_ = [
        ("grid_res", C.c_int * 3), ("dx", C.c_float), ("dt", C.c_float),
        ("domain_min", C.c_float * 3), ("domain_max", C.c_float * 3),
        ("closed_bounds", C.c_int * 6), ("max_iter_solve", C.c_int),
        ("density", C.c_float), ("gravity", C.c_float), ("buoyancy", C.c_float),
        ("cooling_rate", C.c_float), ("vorticity_conf", C.c_float), ("drag", C.c_float),
        ("wind_dir", C.c_int), ("wind_amp", C.c_float), ("wind_speed", C.c_float),
        ("wind_turb_amp", C.c_float), ("wind_turb_scale", C.c_float),
        ("turbulence_amp", C.c_float), ("turbulence_scale", C.c_float),
        ("turb_mask_temp_ramp", C.c_float * 2), ("turb_mask_vel_ramp", C.c_float * 2),
        ("source_animated", C.c_int), ("source_range", C.c_int * 2),
        ("source_max_particle_count", C.c_int),
        ("noise_seed", C.c_int), ("residual_tol", C.c_float), ("residual_check_every", C.c_int),
        ("cfl_number", C.c_float), ("max_substeps", C.c_int), ("smoke_density", C.c_float),
        ("warm_start", C.c_int), ("residual_check_adaptive", C.c_int),
    ]


class _RenderSettings(C.Structure):
    """blz_render_settings (include/blaze.h), declared again here."""
    _fields_ = [
        ("width", C.c_int), ("height", C.c_int), ("cam_focal", C.c_float), ("cam_aperture", C.c_float),
        ("cam_u", C.c_float * 3), ("cam_v", C.c_float * 3), ("cam_w", C.c_float * 3), ("cam_trans", C.c_float * 3),
        ("p_step", C.c_float), ("s_step", C.c_float), ("cutoff", C.c_float), ("light_count", C.c_int),
        ("light_dir", (C.c_float * 3) * 10), ("light_color", (C.c_float * 3) * 10),
        ("light_intensity", C.c_float * 10), ("light_exposure", C.c_float * 10), ("light_samples", C.c_int * 10),
        ("density_scale", C.c_float), ("shadow_density_scale", C.c_float),
        ("volume_color", C.c_float * 3), ("scattering", C.c_float * 3),
        ("absorption", C.c_float), ("gain", C.c_float), ("emission_scale", C.c_float),
        ("ramp_size", C.c_int), ("ramp_keys", C.c_float * 10), ("ramp_colors", (C.c_float * 3) * 10),
        ("temp_range", C.c_float * 2), ("ms_scale", C.c_float), ("ms_mask", C.c_float * 2),
        ("ms_color", C.c_float * 3), ("ms_blur_iter", C.c_int),
    ]


def _lib():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"oracle: {LIB_PATH} missing; run make -C oracle")
    lib = C.CDLL(LIB_PATH)
    vp, P, d, f, i = C.c_void_p, C.POINTER, C.c_double, C.c_float, C.c_int
    sig = {
        "orc_create": ([P(_Settings), i], vp),
        "orc_destroy": ([vp], None),
        "orc_set_threads": ([i], None),
        "orc_max_threads": ([], i),
        "orc_set_rbgs_order": ([vp, i], None),
        "orc_add_source": ([vp, P(f), i], i),
        "orc_step": ([vp, f], i),
        "orc_run_stage": ([vp, i, f, f], i),
        "orc_field_dims": ([vp, i, P(i)], i),
        "orc_copy_field": ([vp, i, i, P(d)], i),
        "orc_set_field": ([vp, i, P(d)], i),
        "orc_set_field_back": ([vp, i, P(d)], i),
        "orc_curl_noise": ([vp, f], i),
        "orc_get_stats": ([vp, P(i), P(i), P(i), P(d), P(d)], i),
        "orc_residual": ([vp, f], d),
        "orc_cerp_f32": ([f, f, f, f, f], f),
        "orc_cerp_f64": ([d, d, d, d, d], d),
        "orc_cubic_pulse_f32": ([f, f], f),
        "orc_perlin_f32": ([f, f, f, f, i], f),
        "orc_perlin_f64": ([d, d, d, d, i], d),
        "orc_noise_hash": ([C.c_uint], C.c_uint),
        "orc_lattice_hash": ([f, f, f, f], C.c_uint),
        "orc_sample": ([vp, i, i, i, d, d, d], d),
        "orc_render": ([P(f), i, i, i, f, P(f), P(f), P(_RenderSettings), P(f), i, i, i], i),
        "orc_bexp": ([f], f),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib


lib = _lib()


def _as_c_settings(s) -> _Settings:
    """Accept any ctypes struct with the blz_settings layout."""
    out = _Settings()
    C.memmove(C.byref(out), C.byref(s), C.sizeof(_Settings))
    return out


def set_threads(n: int):
    lib.orc_set_threads(int(n))


def max_threads() -> int:
    return lib.orc_max_threads()


class Oracle:
    """Serial CPU solver. fp64=True runs the same code in double precision."""

    def __init__(self, settings, fp64: bool = False):
        self._s = _as_c_settings(settings)
        self.W, self.H, self.D = list(self._s.grid_res)
        self.fp64 = fp64
        self._h = lib.orc_create(C.byref(self._s), 1 if fp64 else 0)
        if not self._h:
            raise MemoryError("orc_create failed")

    def __del__(self):
        if getattr(self, "_h", None):
            lib.orc_destroy(self._h)
            self._h = None

    def set_interleaved_rbgs(self, on: bool):
        lib.orc_set_rbgs_order(self._h, 1 if on else 0)

    def add_source(self, part8=None):
        if part8 is None:
            lib.orc_add_source(self._h, None, 0)
            return
        a = np.ascontiguousarray(part8, dtype=np.float32).reshape(-1, 8)
        lib.orc_add_source(self._h, a.ctypes.data_as(C.POINTER(C.c_float)), a.shape[0])

    def step(self, time: float):
        lib.orc_step(self._h, C.c_float(time))

    def run_stage(self, stage: int, time: float, dt: float):
        rc = lib.orc_run_stage(self._h, stage, C.c_float(time), C.c_float(dt))
        if rc:
            raise ValueError(f"orc_run_stage({stage}) -> {rc}")

    def dims(self, which: int):
        d = (C.c_int * 3)()
        if lib.orc_field_dims(self._h, which, d):
            raise ValueError(f"field {which} unavailable")
        return tuple(d)

    def field(self, which: int, back: bool = False) -> np.ndarray:
        nx, ny, nz = self.dims(which)
        out = np.empty((nz, ny, nx), dtype=np.float64)
        if lib.orc_copy_field(self._h, which, 1 if back else 0, out.ctypes.data_as(C.POINTER(C.c_double))):
            raise ValueError(f"field {which} unavailable")
        return out if self.fp64 else out.astype(np.float32)

    def set_field(self, which: int, arr):
        nx, ny, nz = self.dims(which)
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(nz, ny, nx)
        if lib.orc_set_field(self._h, which, a.ctypes.data_as(C.POINTER(C.c_double))):
            raise ValueError(f"field {which} unavailable")

    def set_field_back(self, which: int, arr):
        nx, ny, nz = self.dims(which)
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(nz, ny, nx)
        if lib.orc_set_field_back(self._h, which, a.ctypes.data_as(C.POINTER(C.c_double))):
            raise ValueError(f"field {which} unavailable")

    def curl_noise(self, dt: float):
        """The curl-noise half of the turbulence stage on the current back
        buffers (pin tests)."""
        lib.orc_curl_noise(self._h, C.c_float(dt))

    def stats(self):
        sub, tot, last = C.c_int(), C.c_int(), C.c_int()
        r, r0 = C.c_double(), C.c_double()
        lib.orc_get_stats(self._h, C.byref(sub), C.byref(tot), C.byref(last), C.byref(r), C.byref(r0))
        return {"substeps": sub.value, "rbgs_iters": tot.value, "rbgs_iters_last": last.value,
                "residual": r.value, "residual_initial": r0.value}

    def residual(self, dt: float) -> float:
        return lib.orc_residual(self._h, C.c_float(dt))

    def sample(self, which: int, pos, cubic=False, clamp=False) -> float:
        return lib.orc_sample(self._h, which, int(cubic), int(clamp), *[float(v) for v in pos])


def render(T, dx, dom_min, dom_max, settings, rows=None):
    """Image (height, width, 3) float32 of the temperature grid T (D, H, W)
    with the render settings (any ctypes struct of the blz_render_settings
    layout); rows = (row0, row1[, stride]) renders only rows row0, row0 +
    stride, ... below row1 (the others stay 0)."""
    rs = _RenderSettings()
    C.memmove(C.byref(rs), C.byref(settings), C.sizeof(_RenderSettings))
    T = np.ascontiguousarray(T, np.float32)
    D, H, W = T.shape
    out = np.zeros((rs.height, rs.width, 3), np.float32)
    lo = (C.c_float * 3)(*dom_min)
    hi = (C.c_float * 3)(*dom_max)
    r0, r1, st = (tuple(rows) + (1,))[:3] if rows is not None else (0, rs.height, 1)
    fp = C.POINTER(C.c_float)
    rc = lib.orc_render(T.ctypes.data_as(fp), W, H, D, C.c_float(dx), lo, hi, C.byref(rs),
                        out.ctypes.data_as(fp), int(r0), int(r1), int(st))
    if rc != 0:
        raise ValueError("orc_render: invalid arguments")
    return out


def bexp(x: float) -> float:
    return lib.orc_bexp(C.c_float(x))
