"""blaze_amd — MI355X (gfx950) smoke/fire solver step behind Blaze's interface.

Python mirror of the C ABI in ``include/blaze.h`` (ctypes over the in-tree
``libblaze_amd.so``). It follows the reference's ``FluidSolver`` /
``SceneSettings`` / ``Parser`` interface (asiroisvigneux/Blaze
src/FluidSolver.h:38-110, src/SceneSettings.h:32-71, src/Parser.h:41-211):

* :func:`parse_scene` reads a ``.blz`` file into :class:`Settings`;
* :class:`Solver` owns the device state: ``add_source`` (addSource),
  ``step`` (step), ``field`` / ``set_field`` (state access for tests and
  consumers such as the renderer, which reads ``T`` like getTempGrid).

Every compute call runs HIP kernels; there is no CPU fallback. Importing
this package raises if the shared library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libblaze_amd.so")

FIELD_T, FIELD_U, FIELD_V, FIELD_W, FIELD_DENSITY, FIELD_P, FIELD_RHS = range(7)
FIELDS = {"T": FIELD_T, "U": FIELD_U, "V": FIELD_V, "W": FIELD_W, "density": FIELD_DENSITY,
          "P": FIELD_P, "rhs": FIELD_RHS}
(STAGE_COOLDOWN, STAGE_DRAG, STAGE_VORTICITY, STAGE_BUOYANCY, STAGE_WIND, STAGE_TURBULENCE,
 STAGE_DIVERGENCE, STAGE_PRESSURE, STAGE_PROJECT_UPDATE, STAGE_ADVECT) = range(10)
NUM_STAGES = 10
BATCH_TUNE = 10  # BLZ_BATCH_TUNE: batch autotuner candidates
STAGE_NAMES = ["cooldown", "drag", "vorticity", "buoyancy", "wind", "turbulence", "divergence",
               "pressure", "project_update", "advect"]


class Settings(C.Structure):
    """C mirror of the solver subset of SceneSettings (blz_settings)."""

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

    def copy(self) -> "Settings":
        s = Settings()
        C.pointer(s)[0] = self
        return s


class Stats(C.Structure):
    _fields_ = [
        ("substeps", C.c_int), ("rbgs_iters", C.c_int), ("rbgs_iters_last", C.c_int),
        ("residual", C.c_double), ("residual_initial", C.c_double),
        ("warm_start_used", C.c_int), ("warm_start_fallbacks", C.c_int),
        ("stage_ms", C.c_double * NUM_STAGES), ("source_ms", C.c_double), ("step_ms", C.c_double),
        ("rbgs_kernel_ms", C.c_double), ("rbgs_launches", C.c_int), ("residual_checks", C.c_int),
        ("rbgs_sweeps_timed", C.c_int), ("rbgs_kernel", C.c_int), ("rbgs_sweeps_executed", C.c_int),
    ]


TRACE_CHECK, TRACE_KEPT, TRACE_REJECTED = 0, 1, 2


class TraceEntry(C.Structure):
    """blz_trace_entry: one FP64 residual evaluation of a solve."""
    _fields_ = [("iters", C.c_int), ("kind", C.c_int), ("res", C.c_double)]


class BatchStats(C.Structure):
    _fields_ = [
        ("scenes", C.c_int), ("variant", C.c_int), ("graph_nodes", C.c_int), ("graphs_run", C.c_int),
        ("voxels", C.c_longlong), ("capture_ms", C.c_double), ("step_ms", C.c_double),
        ("tune_variant", C.c_int * BATCH_TUNE), ("tune_ms", C.c_double * BATCH_TUNE),
    ]


class RenderSettings(C.Structure):
    """blz_render_settings: camera, lights, sampling, shader, colour ramp and
    multi-scatter terms of the renderer (the .blz render keys)."""
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


class ParticleGen(C.Structure):
    _fields_ = [
        ("count", C.c_int), ("seed", C.c_ulonglong), ("center", C.c_float * 3),
        ("radius", C.c_float), ("pscale", C.c_float), ("temp", C.c_float),
        ("vel", C.c_float * 3), ("radial_vel", C.c_float),
    ]


E_INVALID, E_HIP, E_NOMEM, E_IO, E_PARSE, E_NUMERIC = -1, -2, -3, -4, -5, -6


class BlazeError(RuntimeError):
    """A failed C-ABI call; ``code`` is its BLZ_E_* status."""

    def __init__(self, msg: str, code: int = 0):
        super().__init__(msg)
        self.code = code


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"blaze_amd: {LIB_PATH} is missing; build it first (python -c "
            "'import __graft_entry__ as g; g.build()' or make -C blaze_amd/csrc)")
    lib = C.CDLL(LIB_PATH)
    P = C.POINTER
    vp = C.c_void_p
    sig = {
        "blz_abi_version": ([], C.c_int),
        "blz_settings_default": ([P(Settings)], C.c_int),
        "blz_parse_scene": ([C.c_char_p, P(Settings), C.c_char_p, C.c_int, C.c_char_p, C.c_int,
                             P(ParticleGen), P(C.c_int)], C.c_int),
        "blz_read_particles": ([C.c_char_p, P(C.c_float), C.c_int, P(C.c_int)], C.c_int),
        "blz_count_particles": ([C.c_char_p, P(C.c_int)], C.c_int),
        "blz_generate_particles": ([P(ParticleGen), P(C.c_float)], C.c_int),
        "blz_create": ([P(Settings), P(vp)], C.c_int),
        "blz_destroy": ([vp], C.c_int),
        "blz_add_source": ([vp, P(C.c_float), C.c_int], C.c_int),
        "blz_step": ([vp, C.c_float], C.c_int),
        "blz_step_async": ([vp, C.c_float], C.c_int),
        "blz_sync": ([vp], C.c_int),
        "blz_run_stage": ([vp, C.c_int, C.c_float, C.c_float], C.c_int),
        "blz_get_field": ([vp, C.c_int, P(vp), P(C.c_int)], C.c_int),
        "blz_copy_field": ([vp, C.c_int, P(C.c_float)], C.c_int),
        "blz_set_field": ([vp, C.c_int, P(C.c_float)], C.c_int),
        "blz_get_stats": ([vp, P(Stats)], C.c_int),
        "blz_set_timing": ([vp, C.c_int], C.c_int),
        "blz_set_rbgs_variant": ([vp, C.c_int], C.c_int),
        "blz_set_residual_trace": ([vp, C.c_int], C.c_int),
        "blz_set_solve_limits": ([vp, C.c_float, C.c_int, C.c_int], C.c_int),
        "blz_get_residual_trace": ([vp, C.c_int, P(TraceEntry), C.c_int, P(C.c_int)], C.c_int),
        "blz_selftest_division": ([C.c_int, P(C.c_ulonglong)], C.c_int),
        "blz_last_error": ([], C.c_char_p),
        "blz_net_create": ([C.c_ulonglong, P(vp)], C.c_int),
        "blz_net_destroy": ([vp], C.c_int),
        "blz_net_num_params": ([vp, P(C.c_longlong)], C.c_int),
        "blz_net_get_params": ([vp, P(C.c_float)], C.c_int),
        "blz_net_set_params": ([vp, P(C.c_float)], C.c_int),
        "blz_net_get_grads": ([vp, P(C.c_float)], C.c_int),
        "blz_net_forward": ([vp, P(C.c_float), P(C.c_int), P(C.c_float)], C.c_int),
        "blz_net_train_step": ([vp, P(C.c_float), P(C.c_float), P(C.c_int), C.c_float, C.c_int,
                                P(C.c_double)], C.c_int),
        "blz_set_warm_start": ([vp, vp], C.c_int),
        "blz_set_warm_start_parts": ([vp, C.c_int], C.c_int),
        "blz_render_settings_default": ([P(RenderSettings)], C.c_int),
        "blz_parse_render": ([C.c_char_p, P(RenderSettings)], C.c_int),
        "blz_render": ([vp, P(RenderSettings), P(C.c_float)], C.c_int),
        "blz_render_noalloc": ([vp, P(RenderSettings), P(C.c_float)], C.c_int),
        "blz_stage_source": ([vp, P(C.c_float), C.c_int], C.c_int),
        "blz_add_source_staged": ([vp], C.c_int),
        "blz_write_image": ([C.c_char_p, P(C.c_float), C.c_int, C.c_int], C.c_int),
        "blz_warm_start_input": ([vp, C.c_float, P(C.c_float), P(C.c_int), P(C.c_int)], C.c_int),
        "blz_autotune_rbgs": ([vp, C.c_int, P(C.c_int), P(C.c_float)], C.c_int),
        "blz_warm_start_apply": ([vp, C.c_float, P(C.c_float), P(C.c_double), P(C.c_double), P(C.c_int)], C.c_int),
        "blz_warm_start_global": ([vp, C.c_float], C.c_int),
        "blz_warm_start_input_global": ([vp, C.c_float, P(C.c_float)], C.c_int),
        "blz_batch_create": ([P(Settings), C.c_int, P(vp)], C.c_int),
        "blz_batch_destroy": ([vp], C.c_int),
        "blz_batch_solver": ([vp, C.c_int, P(vp)], C.c_int),
        "blz_batch_set_warm_start": ([vp, vp], C.c_int),
        "blz_batch_autotune": ([vp, C.c_int, P(C.c_int)], C.c_int),
        "blz_batch_set_variant": ([vp, C.c_int], C.c_int),
        "blz_batch_step": ([vp, C.c_float], C.c_int),
        "blz_batch_get_stats": ([vp, P(BatchStats)], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib, sig


_lib, ABI_SYMBOLS = _load()


def _check(rc: int, what: str):
    if rc != 0:
        msg = _lib.blz_last_error().decode(errors="replace")
        raise BlazeError(f"{what} failed ({rc}): {msg}", rc)


def default_settings() -> Settings:
    s = Settings()
    _check(_lib.blz_settings_default(C.byref(s)), "blz_settings_default")
    return s


def parse_scene(path: str):
    """.blz -> (Settings, scene_dir, particle_file, ParticleGen or None)."""
    s = Settings()
    d = C.create_string_buffer(4096)
    pf = C.create_string_buffer(4096)
    g = ParticleGen()
    has = C.c_int(0)
    _check(_lib.blz_parse_scene(path.encode(), C.byref(s), d, 4096, pf, 4096, C.byref(g), C.byref(has)),
           f"parse_scene({path})")
    return s, d.value.decode(), pf.value.decode(), (g if has.value else None)


def default_render_settings() -> RenderSettings:
    r = RenderSettings()
    _check(_lib.blz_render_settings_default(C.byref(r)), "blz_render_settings_default")
    return r


def parse_render(path: str) -> RenderSettings:
    """The render keys of a .blz (camera, lights, shader, ramp, multi-scatter)."""
    r = RenderSettings()
    _check(_lib.blz_parse_render(path.encode(), C.byref(r)), f"parse_render({path})")
    return r


def write_image(path: str, img: np.ndarray):
    """(height, width, 3) float image -> .ppm / .png / .exr by extension."""
    a = np.ascontiguousarray(img, np.float32)
    h, w, _ = a.shape
    _check(_lib.blz_write_image(path.encode(), a.ctypes.data_as(C.POINTER(C.c_float)), w, h), f"write_image({path})")


def generate_particles(gen: ParticleGen) -> np.ndarray:
    out = np.zeros((gen.count, 8), dtype=np.float32)
    _check(_lib.blz_generate_particles(C.byref(gen), out.ctypes.data_as(C.POINTER(C.c_float))),
           "generate_particles")
    return out


def read_particles(path: str) -> np.ndarray:
    n = C.c_int(0)
    _check(_lib.blz_count_particles(path.encode(), C.byref(n)), "count_particles")
    out = np.zeros((n.value, 8), dtype=np.float32)
    got = C.c_int(0)
    _check(_lib.blz_read_particles(path.encode(), out.ctypes.data_as(C.POINTER(C.c_float)), n.value,
                                   C.byref(got)), "read_particles")
    return out[: got.value]


def field_dims(s: Settings, which: int):
    W, H, D = s.grid_res
    return {FIELD_T: (W, H, D), FIELD_U: (W + 1, H, D), FIELD_V: (W, H + 1, D),
            FIELD_W: (W, H, D + 1), FIELD_DENSITY: (W, H, D), FIELD_P: (W, H, D),
            FIELD_RHS: (W, H, D)}[which]


class Solver:
    """One device solver (blz_solver). Arrays are returned as (D, H, W)-shaped
    float32 views of the x-fastest layout."""

    def __init__(self, settings: Settings):
        self.settings = settings.copy()
        self._h = C.c_void_p()
        _check(_lib.blz_create(C.byref(self.settings), C.byref(self._h)), "blz_create")

    def close(self):
        if self._h:
            _lib.blz_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def add_source(self, part8=None):
        if part8 is None:
            _check(_lib.blz_add_source(self._h, None, 0), "blz_add_source")
            return
        a = np.ascontiguousarray(part8, dtype=np.float32).reshape(-1, 8)
        _check(_lib.blz_add_source(self._h, a.ctypes.data_as(C.POINTER(C.c_float)), a.shape[0]),
               "blz_add_source")

    def step(self, time: float):
        _check(_lib.blz_step(self._h, C.c_float(time)), "blz_step")

    def step_async(self, time: float):
        _check(_lib.blz_step_async(self._h, C.c_float(time)), "blz_step_async")

    def sync(self):
        _check(_lib.blz_sync(self._h), "blz_sync")

    def run_stage(self, stage: int, time: float, dt: float):
        _check(_lib.blz_run_stage(self._h, stage, C.c_float(time), C.c_float(dt)), "blz_run_stage")

    def field(self, which) -> np.ndarray:
        which = FIELDS.get(which, which)
        nx, ny, nz = field_dims(self.settings, which)
        out = np.empty((nz, ny, nx), dtype=np.float32)
        _check(_lib.blz_copy_field(self._h, which, out.ctypes.data_as(C.POINTER(C.c_float))),
               "blz_copy_field")
        return out

    def set_field(self, which, arr):
        which = FIELDS.get(which, which)
        nx, ny, nz = field_dims(self.settings, which)
        a = np.ascontiguousarray(arr, dtype=np.float32).reshape(nz, ny, nx)
        _check(_lib.blz_set_field(self._h, which, a.ctypes.data_as(C.POINTER(C.c_float))),
               "blz_set_field")

    def device_field(self, which):
        """(device pointer, (nx, ny, nz)) of a front buffer (getTempGrid for T)."""
        which = FIELDS.get(which, which)
        p = C.c_void_p()
        dims = (C.c_int * 3)()
        _check(_lib.blz_get_field(self._h, which, C.byref(p), dims), "blz_get_field")
        return p.value, tuple(dims)

    def stats(self) -> Stats:
        st = Stats()
        _check(_lib.blz_get_stats(self._h, C.byref(st)), "blz_get_stats")
        return st

    def set_timing(self, on: bool):
        _check(_lib.blz_set_timing(self._h, 1 if on else 0), "blz_set_timing")

    def set_rbgs_variant(self, v: int):
        _check(_lib.blz_set_rbgs_variant(self._h, v), "blz_set_rbgs_variant")

    def set_solve_limits(self, residual_tol: float, max_iter_solve: int, residual_check_every: int = 10):
        """Change the solve's tolerance, sweep cap and check interval for the
        following steps."""
        _check(_lib.blz_set_solve_limits(self._h, C.c_float(residual_tol), int(max_iter_solve),
                                         int(residual_check_every)), "blz_set_solve_limits")

    def set_residual_trace(self, max_entries: int):
        """Record up to max_entries residual evaluations per solve (0: off)."""
        _check(_lib.blz_set_residual_trace(self._h, int(max_entries)), "blz_set_residual_trace")

    def residual_trace(self, solve: int):
        """History of solve `solve` of the last step: list of (sweeps, kind,
        residual), kind TRACE_CHECK / TRACE_KEPT / TRACE_REJECTED."""
        n = C.c_int()
        _check(_lib.blz_get_residual_trace(self._h, solve, None, 0, C.byref(n)), "blz_get_residual_trace")
        buf = (TraceEntry * max(1, n.value))()
        _check(_lib.blz_get_residual_trace(self._h, solve, buf, n.value, C.byref(n)), "blz_get_residual_trace")
        return [(buf[i].iters, buf[i].kind, buf[i].res) for i in range(n.value)]

    def autotune(self, sweeps: int = 20):
        """Time every RBGS variant on this grid and keep the fastest; returns
        (variant, {variant: ms per sweep})."""
        v = C.c_int()
        cand = [2, 4, 8, 102, 103, 202, 203, 402, 403, 1103, 1104, 1105, 1106, 1203, 1204, 1403, 1404]
        ms = (C.c_float * len(cand))()
        _check(_lib.blz_autotune_rbgs(self._h, sweeps, C.byref(v), ms), "blz_autotune_rbgs")
        return v.value, {c: ms[i] for i, c in enumerate(cand) if ms[i] > 0}

    def stage_source(self, part8: np.ndarray):
        """Start the upload of the next frame's particles (N, 8) on the copy
        stream; safe to call from another thread while step() runs."""
        a = np.ascontiguousarray(part8, np.float32)
        self._staged = a
        _check(_lib.blz_stage_source(self._h, a.ctypes.data_as(C.POINTER(C.c_float)), int(a.shape[0])),
               "blz_stage_source")

    def add_source_staged(self):
        """addSource with the frame staged last."""
        _check(_lib.blz_add_source_staged(self._h), "blz_add_source_staged")

    def render(self, settings: RenderSettings) -> np.ndarray:
        """(height, width, 3) float32 image of the current temperature."""
        out = np.empty((settings.height, settings.width, 3), np.float32)
        _check(_lib.blz_render(self._h, C.byref(settings), out.ctypes.data_as(C.POINTER(C.c_float))), "blz_render")
        return out

    def render_noalloc(self, settings: RenderSettings) -> np.ndarray:
        """render() through blz_render_noalloc (the library's static scratch)."""
        out = np.empty((settings.height, settings.width, 3), np.float32)
        _check(_lib.blz_render_noalloc(self._h, C.byref(settings), out.ctypes.data_as(C.POINTER(C.c_float))),
               "blz_render_noalloc")
        return out

    def set_warm_start(self, net, network: bool = True, global_correction: bool = True, parts: int = 0):
        """Attach a Net (None detaches); used when settings.warm_start != 0
        and residual_tol > 0. network / global_correction select the parts
        of the correction (at least one); parts != 0 gives the bits of
        blz_set_warm_start_parts directly (bit 2: exact coarse solve in place
        of the network, bit 3: network before the global correction)."""
        self._net = net
        _check(_lib.blz_set_warm_start(self._h, net._h if net is not None else None), "blz_set_warm_start")
        if net is not None:
            bits = parts or (int(network) | (2 * int(global_correction)))
            _check(_lib.blz_set_warm_start_parts(self._h, bits), "blz_set_warm_start_parts")

    def warm_start_dims(self):
        nd, cd = (C.c_int * 3)(), (C.c_int * 3)()
        _check(_lib.blz_warm_start_input(self._h, C.c_float(1.0), None, nd, cd), "blz_warm_start_input")
        return tuple(nd), tuple(cd)

    def warm_start_apply(self, dt: float, pred: np.ndarray):
        """Apply a given coarse prediction (Dn, Hn, Wn) as the learned
        correction would be; returns (residual before, after, kept)."""
        a = np.ascontiguousarray(pred, np.float32)
        b, f, k = C.c_double(), C.c_double(), C.c_int()
        _check(_lib.blz_warm_start_apply(self._h, C.c_float(dt), a.ctypes.data_as(C.POINTER(C.c_float)),
                                         C.byref(b), C.byref(f), C.byref(k)), "blz_warm_start_apply")
        return b.value, f.value, bool(k.value)

    def warm_start_global(self, dt: float):
        """Apply the global correction alone, unjudged (training data)."""
        _check(_lib.blz_warm_start_global(self._h, C.c_float(dt)), "blz_warm_start_global")

    def warm_start_input(self, dt: float, after_global: bool = False) -> np.ndarray:
        """(2, Dn, Hn, Wn) float32: normalised coarse residual and solid mask.
        after_global: the input a correction of the default order feeds the
        network (the residual the global correction would leave; the
        pressure is not changed)."""
        (wn, hn, dn), _ = self.warm_start_dims()
        out = np.empty((2, dn, hn, wn), np.float32)
        ptr = out.ctypes.data_as(C.POINTER(C.c_float))
        if after_global:
            _check(_lib.blz_warm_start_input_global(self._h, C.c_float(dt), ptr), "blz_warm_start_input_global")
        else:
            _check(_lib.blz_warm_start_input(self._h, C.c_float(dt), ptr, None, None), "blz_warm_start_input")
        return out


class _Borrowed(Solver):
    """A scene handle owned by a Batch (not destroyed here)."""

    def __init__(self, h, settings, owner):
        self._h = h
        self.settings = settings.copy()
        self._owner = owner

    def __del__(self):
        pass


class Batch:
    """Many scenes stepped together (blz_batch): one captured graph replay per
    step, all scenes' sweeps and residual checks in single launches."""

    def __init__(self, settings_list):
        arr = (Settings * len(settings_list))(*[x.copy() for x in settings_list])
        self._h = C.c_void_p()
        _check(_lib.blz_batch_create(arr, len(settings_list), C.byref(self._h)), "blz_batch_create")
        self.scenes = []
        for i, st in enumerate(settings_list):
            h = C.c_void_p()
            _check(_lib.blz_batch_solver(self._h, i, C.byref(h)), "blz_batch_solver")
            self.scenes.append(_Borrowed(h, st, self))
        self._net = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.blz_batch_destroy(h)
            self._h = None

    def set_warm_start(self, net):
        self._net = net
        _check(_lib.blz_batch_set_warm_start(self._h, net._h if net is not None else None),
               "blz_batch_set_warm_start")

    def autotune(self, sweeps: int = 20) -> int:
        v = C.c_int()
        _check(_lib.blz_batch_autotune(self._h, sweeps, C.byref(v)), "blz_batch_autotune")
        return v.value

    def set_variant(self, v: int):
        """Sweep with candidate v of the autotuner (stats().tune_variant)."""
        _check(_lib.blz_batch_set_variant(self._h, v), "blz_batch_set_variant")

    def step(self, time: float):
        _check(_lib.blz_batch_step(self._h, C.c_float(time)), "blz_batch_step")

    def stats(self) -> BatchStats:
        st = BatchStats()
        _check(_lib.blz_batch_get_stats(self._h, C.byref(st)), "blz_batch_get_stats")
        return st


class Net:
    """The warm-start U-Net (blz_net): parameters live on the device; inputs
    are (B, 2, D, H, W) float32 arrays (residual, mask), outputs (B, D, H, W)."""

    def __init__(self, seed: int = 0):
        self._h = C.c_void_p()
        _check(_lib.blz_net_create(C.c_ulonglong(seed), C.byref(self._h)), "blz_net_create")
        n = C.c_longlong()
        _check(_lib.blz_net_num_params(self._h, C.byref(n)), "blz_net_num_params")
        self.num_params = n.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.blz_net_destroy(h)
            self._h = None

    def params(self) -> np.ndarray:
        out = np.empty(self.num_params, np.float32)
        _check(_lib.blz_net_get_params(self._h, out.ctypes.data_as(C.POINTER(C.c_float))), "blz_net_get_params")
        return out

    def set_params(self, p: np.ndarray):
        a = np.ascontiguousarray(p, np.float32)
        assert a.size == self.num_params
        _check(_lib.blz_net_set_params(self._h, a.ctypes.data_as(C.POINTER(C.c_float))), "blz_net_set_params")

    def grads(self) -> np.ndarray:
        out = np.empty(self.num_params, np.float32)
        _check(_lib.blz_net_get_grads(self._h, out.ctypes.data_as(C.POINTER(C.c_float))), "blz_net_get_grads")
        return out

    @staticmethod
    def _prep(x):
        x = np.ascontiguousarray(x, np.float32)
        assert x.ndim == 5 and x.shape[1] == 2, "input must be (B, 2, D, H, W)"
        B, _, D, H, W = x.shape
        # channel-major per batch -> two planes of B*D*H*W values
        planes = np.ascontiguousarray(x.transpose(1, 0, 2, 3, 4))
        dims = (C.c_int * 4)(B, D, H, W)
        return planes, dims, (B, D, H, W)

    def forward(self, x: np.ndarray) -> np.ndarray:
        planes, dims, shp = self._prep(x)
        out = np.empty(shp, np.float32)
        _check(_lib.blz_net_forward(self._h, planes.ctypes.data_as(C.POINTER(C.c_float)), dims,
                                    out.ctypes.data_as(C.POINTER(C.c_float))), "blz_net_forward")
        return out

    def train_step(self, x: np.ndarray, target: np.ndarray, lr: float = 1e-3, update: bool = True) -> float:
        planes, dims, shp = self._prep(x)
        t = np.ascontiguousarray(target, np.float32)
        assert t.shape == shp
        loss = C.c_double()
        _check(_lib.blz_net_train_step(self._h, planes.ctypes.data_as(C.POINTER(C.c_float)),
                                       t.ctypes.data_as(C.POINTER(C.c_float)), dims, C.c_float(lr),
                                       1 if update else 0, C.byref(loss)), "blz_net_train_step")
        return loss.value


def selftest_division(divisor: int) -> int:
    """Mismatches of the solver's device division by `divisor` (3..6) against
    IEEE division over all 2^32 floats (expected 0)."""
    n = C.c_ulonglong(0)
    _check(_lib.blz_selftest_division(divisor, C.byref(n)), "blz_selftest_division")
    return n.value


def abi_version() -> int:
    return _lib.blz_abi_version()
