"""Python host binding of libgwm_hip.so (the MI355X dry + warm-rain step).

A thin ctypes layer over the C-ABI in include/gwm_hip.h, used by the tests
and bench.py.  There is no CPU fallback: if the HIP library is missing or no
GPU is present, the calls raise.
"""
from __future__ import annotations

import atexit
import ctypes
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
REPO = os.path.dirname(PKG)
LIB_PATH = os.path.join(PKG, "lib", "libgwm_hip.so")
HEADER = os.path.join(REPO, "include", "gwm_hip.h")

F32, F64 = 0, 1
# gwm_fp32_math functions (GWM_MATH_*)
MATH_POW_EOS, MATH_POW_EXNER, MATH_EXP, MATH_LOG1P, MATH_EXPM1, MATH_RCP, MATH_POW1PM1 = range(7)
RHO, RHO_THETA, MOM_U, MOM_V, MOM_W, PRESSURE = 0, 1, 2, 3, 4, 5
TEND_RHO, TEND_RHO_THETA, TEND_U, TEND_V, TEND_W = 6, 7, 8, 9, 10
PRECIP, TRACER0 = 11, 16
BOUNDARY_NONE, BOUNDARY_STRIPS, BOUNDARY_RELAX = 0, 1, 2

PLAN_VIEW_NAMES = [
    "terrain_height", "z_center", "rho_d", "theta_m", "air_pressure", "air_temperature",
    "u_velocity", "v_velocity", "wind_speed", "w_velocity", "specific_humidity",
    "relative_humidity", "dewpoint", "cloud_water_mixing_ratio", "rain_water_mixing_ratio",
    "total_condensate", "column_cloud_water", "column_rain_water", "column_total_condensate",
    "column_rain_fraction", "synthetic_reflectivity", "accumulated_surface_precipitation",
    "mean_surface_precipitation_rate",
]
PLAN_VIEW_FIELDS = len(PLAN_VIEW_NAMES)

SUMMARY_KEYS = [
    "total_dry_mass", "total_rho_theta_m", "total_horizontal_momentum_x",
    "total_horizontal_momentum_y", "total_vertical_momentum", "rho_min", "rho_max",
    "theta_min", "theta_max", "u_min", "u_max", "v_min", "v_max", "w_min", "w_max",
    "vapor_path", "cloud_path", "rain_path", "precip_sum", "precip_mean", "precip_max",
    "precip_cells", "precipitating_cells", "qv_min", "qv_max", "qc_min", "qc_max",
    "qr_min", "qr_max", "nonfinite_values",
]


class GridDesc(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("halo", ctypes.c_int), ("periodic_x", ctypes.c_int),
                ("periodic_y", ctypes.c_int), ("dx", ctypes.c_double),
                ("dy", ctypes.c_double), ("z_top", ctypes.c_double),
                ("terrain_taper_eta", ctypes.c_double)]


class StepConfig(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("fast_substeps", ctypes.c_int),
                ("gravity", ctypes.c_double),
                ("fast_update_horizontal_momentum", ctypes.c_int),
                ("fast_update_density", ctypes.c_int)]


def step_config(dt=2.0, fast_substeps=1, gravity=9.81, horizontal=True, density=True):
    """DryStepperConfig defaults (include/gwm/dycore/dry_core.hpp:50-56)."""
    return StepConfig(dt, fast_substeps, gravity, int(horizontal), int(density))


class WarmRainConfig(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("condensation_relaxation", ctypes.c_double),
                ("cloud_autoconversion_threshold", ctypes.c_double),
                ("cloud_autoconversion_rate", ctypes.c_double),
                ("rain_evaporation_rate", ctypes.c_double),
                ("rain_terminal_velocity", ctypes.c_double),
                ("sedimentation_layer_depth_m", ctypes.c_double),
                ("enable_latent_heating", ctypes.c_int),
                ("accretion_rate", ctypes.c_double)]


class SurrogateNorm(ctypes.Structure):
    _fields_ = [("x_mean", ctypes.c_float * 7), ("x_std", ctypes.c_float * 7),
                ("y_mean", ctypes.c_float * 4), ("y_std", ctypes.c_float * 4),
                ("box_lo", ctypes.c_float * 7), ("box_hi", ctypes.c_float * 7)]


SURROGATE_IN, SURROGATE_OUT, SURROGATE_PARAMS = 7, 4, 9092


def warm_rain_config(dt=2.0, condensation_relaxation=1.0, cloud_autoconversion_threshold=1.0e-4,
                     cloud_autoconversion_rate=1.0e-3, rain_evaporation_rate=5.0e-3,
                     rain_terminal_velocity=0.0, sedimentation_layer_depth_m=0.0,
                     enable_latent_heating=True, accretion_rate=0.0):
    """WarmRainConfig defaults (include/gwm/physics/warm_rain.hpp:14-23)."""
    return WarmRainConfig(dt, condensation_relaxation, cloud_autoconversion_threshold,
                          cloud_autoconversion_rate, rain_evaporation_rate,
                          rain_terminal_velocity, sedimentation_layer_depth_m,
                          int(enable_latent_heating), accretion_rate)


BOUNDARY_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p)


class BoundaryConfig(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int), ("relax_width", ctypes.c_double),
                ("callback", BOUNDARY_FN), ("user", ctypes.c_void_p)]


class GwmError(RuntimeError):
    pass


_lib = None
_shutting_down = False


@atexit.register
def _mark_shutdown():
    # Objects collected during interpreter shutdown leak their device memory
    # to the process exit instead of calling into a runtime that may be
    # tearing down.
    global _shutting_down
    _shutting_down = True


def header_functions():
    """Names of every function declared in include/gwm_hip.h."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(gwm_[a-z0-9_]+)\s*\(", text)) - {"gwm_boundary_fn"})


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GwmError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        vp, ip, dp, cp = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_char_p
        pvp = ctypes.POINTER(ctypes.c_void_p)
        L.gwm_last_error.restype = cp
        L.gwm_ctx_create.argtypes = [ip, ip, pvp]
        L.gwm_ctx_destroy.argtypes = [vp]
        L.gwm_sync.argtypes = [vp]
        L.gwm_ctx_precision.argtypes = [vp]
        L.gwm_ctx_set_kernel_timing.argtypes = [vp, ip]
        L.gwm_ctx_kernel_stats.argtypes = [vp, cp, ctypes.POINTER(ctypes.c_long),
                                           ctypes.POINTER(ctypes.c_double)]
        L.gwm_ctx_reset_kernel_stats.argtypes = [vp]
        L.gwm_ctx_set_chunks.argtypes = [vp, ip]
        L.gwm_state_create_perturbation.argtypes = [vp, vp, ip, pvp]
        L.gwm_state_is_perturbation.argtypes = [vp]
        L.gwm_state_download_f64.argtypes = [vp, ip, vp]
        L.gwm_ctx_tuning_report.argtypes = [vp, cp, ctypes.c_size_t]
        L.gwm_fp32_math.argtypes = [vp, ip, vp, vp, ctypes.c_long]
        L.gwm_grid_create.argtypes = [vp, ctypes.POINTER(GridDesc), vp, vp, pvp]
        L.gwm_grid_destroy.argtypes = [vp]
        L.gwm_grid_download_metric.argtypes = [vp, ip, vp]
        L.gwm_grid_terrain_active.argtypes = [vp]
        L.gwm_terrain_cosine_mountain.argtypes = [ip, ip, dp, dp, dp, dp, dp, vp]
        L.gwm_terrain_agnesi_ridge.argtypes = [ip, ip, dp, dp, dp, dp, vp]
        L.gwm_stretched_levels.argtypes = [ip, dp, dp, dp, vp]
        L.gwm_state_create.argtypes = [vp, vp, ip, pvp]
        L.gwm_state_destroy.argtypes = [vp]
        L.gwm_state_field_size.restype = ctypes.c_long
        L.gwm_state_field_size.argtypes = [vp, ip]
        for n in ("gwm_state_upload", "gwm_state_download"):
            getattr(L, n).argtypes = [vp, ip, vp]
        L.gwm_state_fill.argtypes = [vp, ip, dp]
        L.gwm_state_copy.argtypes = [vp, vp]
        L.gwm_state_capture_boundary.argtypes = [vp]
        L.gwm_halo_fill.argtypes = [vp, ip]
        L.gwm_sync_interfaces.argtypes = [vp, ip]
        L.gwm_halo_exchange_layout.argtypes = [vp, ip, vp, ip, ip, ip, vp]
        L.gwm_device_alloc.argtypes = [vp, ctypes.c_size_t, pvp]
        L.gwm_device_free.argtypes = [vp, vp]
        L.gwm_memcpy_h2d.argtypes = [vp, vp, vp, ctypes.c_size_t]
        L.gwm_memcpy_d2h.argtypes = [vp, vp, vp, ctypes.c_size_t]
        L.gwm_pressure.argtypes = [vp]
        L.gwm_slow_tendencies.argtypes = [vp, ctypes.POINTER(StepConfig)]
        L.gwm_fast_modes.argtypes = [vp, ctypes.POINTER(StepConfig)]
        L.gwm_ssprk3_step.argtypes = [vp, ctypes.POINTER(StepConfig), ctypes.POINTER(BoundaryConfig)]
        L.gwm_tracers_ssprk3.argtypes = [vp, ctypes.POINTER(StepConfig), ip]
        L.gwm_tracers_ssprk3_callback.argtypes = [vp, ctypes.POINTER(StepConfig), ip,
                                                  BOUNDARY_FN, vp]
        L.gwm_warm_rain.argtypes = [vp, ctypes.POINTER(WarmRainConfig), ip]
        L.gwm_apply_boundary.argtypes = [vp, ip, dp]
        L.gwm_moist_step.argtypes = [vp, ctypes.POINTER(StepConfig),
                                     ctypes.POINTER(WarmRainConfig), ctypes.POINTER(BoundaryConfig)]
        L.gwm_budget_sums.argtypes = [vp, vp]
        L.gwm_budget_sums_noalloc.argtypes = [vp, vp]
        L.gwm_state_set_boundary.argtypes = [vp, ip, vp]
        L.gwm_state_checkpoint_write.argtypes = [vp, cp, cp]
        L.gwm_state_checkpoint_read.argtypes = [vp, cp]
        L.gwm_state_set_tracer_strip_mask.argtypes = [vp, ctypes.c_uint]
        L.gwm_plan_view.argtypes = [vp, ip, ip, dp, ip, vp, ctypes.POINTER(ctypes.c_int)]
        L.gwm_init_constant.argtypes = [vp, dp, dp, dp, dp, dp]
        L.gwm_init_hydrostatic_rest.argtypes = [vp, dp, dp]
        L.gwm_init_bubble.argtypes = [vp, vp, ip]
        L.gwm_init_mountain_wave.argtypes = [vp, dp, dp, dp, dp, dp]
        L.gwm_init_constant_n.argtypes = [vp, dp, dp, dp, dp, dp]
        L.gwm_init_convection.argtypes = [vp, dp, ctypes.c_ulonglong, dp, dp]
        ull, lp = ctypes.c_ulonglong, ctypes.c_long
        L.gwm_surrogate_create.argtypes = [vp, ull, pvp]
        L.gwm_surrogate_destroy.argtypes = [vp]
        L.gwm_surrogate_set_norm.argtypes = [vp, ctypes.POINTER(SurrogateNorm)]
        L.gwm_surrogate_get_norm.argtypes = [vp, ctypes.POINTER(SurrogateNorm)]
        L.gwm_surrogate_set_params.argtypes = [vp, vp]
        L.gwm_surrogate_get_params.argtypes = [vp, vp]
        L.gwm_surrogate_sample.argtypes = [vp, ctypes.POINTER(WarmRainConfig), ull, lp, vp, vp, vp]
        L.gwm_surrogate_train.argtypes = [vp, vp, vp, lp, ip, ip, dp, dp, vp]
        L.gwm_surrogate_loss_grad.argtypes = [vp, vp, vp, ip, vp, vp]
        L.gwm_surrogate_infer.argtypes = [vp, vp, lp, vp]
        L.gwm_warm_rain_surrogate.argtypes = [vp, ctypes.POINTER(WarmRainConfig), vp, ip,
                                              ctypes.POINTER(ctypes.c_longlong)]
        L.gwm_surrogate_last_counts.argtypes = [vp, ctypes.POINTER(ctypes.c_longlong),
                                                ctypes.POINTER(ctypes.c_longlong)]
        L.gwm_moist_step_surrogate.argtypes = [vp, ctypes.POINTER(StepConfig),
                                               ctypes.POINTER(WarmRainConfig), vp,
                                               ctypes.POINTER(BoundaryConfig)]
        L.gwm_moist_steps_graph.argtypes = [vp, ctypes.POINTER(StepConfig),
                                            ctypes.POINTER(WarmRainConfig), vp, ip]
        L.gwm_ensemble_tuning_report.argtypes = [cp, ctypes.c_size_t]
        for n in ("gwm_ensemble_steps", "gwm_ensemble_steps_streams"):
            getattr(L, n).argtypes = [vp, ip, ctypes.POINTER(StepConfig),
                                      ctypes.POINTER(WarmRainConfig), vp, ip]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def check(rc):
    if rc != 0:
        raise GwmError(lib().gwm_last_error().decode())


def cosine_mountain(nx, ny, height, cx=0.5, cy=0.5, hx=0.15, hy=0.15):
    out = np.zeros((ny, nx), np.float32)
    check(lib().gwm_terrain_cosine_mountain(nx, ny, cx, cy, hx, hy, height, _ptr(out)))
    return out


def agnesi_ridge(nx, ny, dx, center_m, half_width_m, height_m):
    out = np.zeros((ny, nx), np.float32)
    check(lib().gwm_terrain_agnesi_ridge(nx, ny, dx, center_m, half_width_m, height_m, _ptr(out)))
    return out


def stretched_levels(nz, dz_bottom, dz_top, z_top):
    out = np.zeros(nz + 1, np.float64)
    check(lib().gwm_stretched_levels(nz, dz_bottom, dz_top, z_top, _ptr(out)))
    return out


class Context:
    def __init__(self, device=0, precision=F32):
        self.h = ctypes.c_void_p()
        check(lib().gwm_ctx_create(device, precision, ctypes.byref(self.h)))
        self.precision = precision
        self.dtype = np.float64 if precision == F64 else np.float32

    def close(self):
        if self.h and not _shutting_down:
            lib().gwm_ctx_destroy(self.h)
        self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        check(lib().gwm_sync(self.h))

    def set_kernel_timing(self, on):
        check(lib().gwm_ctx_set_kernel_timing(self.h, int(on)))

    def kernel_stats(self, name):
        n = ctypes.c_long()
        ms = ctypes.c_double()
        check(lib().gwm_ctx_kernel_stats(self.h, name.encode(), ctypes.byref(n), ctypes.byref(ms)))
        return n.value, ms.value

    def reset_kernel_stats(self):
        check(lib().gwm_ctx_reset_kernel_stats(self.h))

    def set_chunks(self, chunks):
        """Fix the level-chunk count of the column kernels (0: autotune)."""
        check(lib().gwm_ctx_set_chunks(self.h, int(chunks)))

    def fp32_math(self, fn, x):
        """The step kernels' FP32 helper `fn` (MATH_*) applied to x on the device."""
        x = np.ascontiguousarray(x, np.float32)
        dx = DeviceArray(self, x.shape)
        dy = DeviceArray(self, x.shape)
        dx.upload(x)
        check(lib().gwm_fp32_math(self.h, int(fn), dx.ptr, dy.ptr, x.size))
        return dy.download()

    def tuning_report(self):
        """{launch shape key: chunk count} chosen by the autotuner."""
        buf = ctypes.create_string_buffer(8192)
        check(lib().gwm_ctx_tuning_report(self.h, buf, len(buf)))
        out = {}
        for item in buf.value.decode().split(";"):
            if item:
                k, v = item.rsplit("=", 1)
                out[k] = int(v)
        return out


class Grid:
    def __init__(self, ctx, nx, ny, nz, halo=1, periodic_x=True, periodic_y=True, dx=1000.0,
                 dy=1000.0, z_top=16000.0, z_interfaces=None, terrain=None, taper_eta=0.25):
        self.ctx = ctx
        self.nx, self.ny, self.nz, self.halo = nx, ny, nz, halo
        self.desc = GridDesc(nx, ny, nz, halo, int(periodic_x), int(periodic_y), dx, dy, z_top,
                             taper_eta)
        zi = None if z_interfaces is None else np.ascontiguousarray(z_interfaces, np.float64)
        tr = None if terrain is None else np.ascontiguousarray(terrain, np.float32).reshape(-1)
        self.h = ctypes.c_void_p()
        check(lib().gwm_grid_create(ctx.h, ctypes.byref(self.desc),
                                    None if zi is None else _ptr(zi),
                                    None if tr is None else _ptr(tr), ctypes.byref(self.h)))

    def __del__(self):
        if getattr(self, "h", None) and not _shutting_down:
            lib().gwm_grid_destroy(self.h)
        self.h = None

    def metric(self, which):
        nx, ny, nz = self.nx, self.ny, self.nz
        shapes = {0: (nz + 1, ny, nx), 1: (nz, ny, nx), 2: (nz, ny, nx), 3: (nz + 1, ny, nx),
                  4: (ny, nx), 5: (ny, nx), 6: (ny, nx), 7: (nz + 1,), 8: (nz,), 9: (nz + 1,),
                  10: (nz,)}
        dt = np.float64 if which >= 7 else self.ctx.dtype
        out = np.empty(shapes[which], dt)
        check(lib().gwm_grid_download_metric(self.h, which, _ptr(out)))
        return out

    @property
    def terrain_active(self):
        return bool(lib().gwm_grid_terrain_active(self.h))


class State:
    """Device state; storage() returns the halo-padded arrays (nz, py, px)."""

    def __init__(self, grid, n_tracers=0, perturbation=False):
        """perturbation=True: the FP32 perturbation form (rho and rho*theta
        stored as departures from FP64 bases: level means on flat, doubly
        periodic grids; on terrain the 3-D state the base was built from,
        dry states only, open or periodic sides)."""
        self.grid = grid
        self.ctx = grid.ctx
        self.n_tracers = n_tracers
        self.h = ctypes.c_void_p()
        create = lib().gwm_state_create_perturbation if perturbation else lib().gwm_state_create
        check(create(self.ctx.h, grid.h, n_tracers, ctypes.byref(self.h)))
        self.perturbation = bool(lib().gwm_state_is_perturbation(self.h))

    def __del__(self):
        if getattr(self, "h", None) and not _shutting_down:
            lib().gwm_state_destroy(self.h)
        self.h = None

    def storage_shape(self, field):
        g = self.grid
        h = g.halo
        fx = g.nx + (1 if field in (MOM_U, TEND_U) else 0)
        fy = g.ny + (1 if field in (MOM_V, TEND_V) else 0)
        fz = g.nz + (1 if field in (MOM_W, TEND_W) else 0)
        if field == PRECIP:
            return (g.ny, g.nx)
        return (fz, fy + 2 * h, fx + 2 * h)

    def storage(self, field):
        out = np.empty(self.storage_shape(field), self.ctx.dtype)
        check(lib().gwm_state_download(self.h, field, _ptr(out)))
        return out

    def set_storage(self, field, arr):
        a = np.ascontiguousarray(arr, self.ctx.dtype)
        assert a.shape == self.storage_shape(field), (a.shape, self.storage_shape(field))
        check(lib().gwm_state_upload(self.h, field, _ptr(a)))

    def get(self, field):
        """Owned values (no halo), shaped like the oracle's global arrays."""
        s = self.storage(field)
        if field == PRECIP:
            return s[None]
        h = self.grid.halo
        return np.ascontiguousarray(s[:, h:s.shape[1] - h, h:s.shape[2] - h])

    def get_f64(self, field):
        """Owned values in FP64 (perturbation form: departure + FP64 base)."""
        shape = self.storage_shape(field)
        out = np.empty(shape, np.float64)
        check(lib().gwm_state_download_f64(self.h, field, _ptr(out)))
        if field == PRECIP:
            return out[None]
        h = self.grid.halo
        return np.ascontiguousarray(out[:, h:out.shape[1] - h, h:out.shape[2] - h])

    def set(self, field, arr):
        if field == PRECIP:
            a = np.ascontiguousarray(np.reshape(arr, self.storage_shape(field)), self.ctx.dtype)
            check(lib().gwm_state_upload(self.h, field, _ptr(a)))
            return
        s = self.storage(field)
        h = self.grid.halo
        s[:, h:s.shape[1] - h, h:s.shape[2] - h] = arr
        self.set_storage(field, s)

    def set_boundary(self, field, arr):
        """Owned values of one field of the lateral-boundary reference
        (gwm_state_set_boundary; capture_boundary first)."""
        s = np.zeros(self.storage_shape(field), self.ctx.dtype)
        h = self.grid.halo
        s[:, h:s.shape[1] - h, h:s.shape[2] - h] = arr
        check(lib().gwm_state_set_boundary(self.h, field, _ptr(np.ascontiguousarray(s))))

    def checkpoint_write(self, path, metadata=None):
        """Restart dump (gwm_state_checkpoint_write) and its manifest
        path + '.json'; metadata: any JSON-serialisable value."""
        import json
        meta = None if metadata is None else json.dumps(metadata).encode()
        check(lib().gwm_state_checkpoint_write(self.h, os.fsencode(path), meta))

    def checkpoint_read(self, path):
        check(lib().gwm_state_checkpoint_read(self.h, os.fsencode(path)))

    def set_tracer_strip_mask(self, mask):
        check(lib().gwm_state_set_tracer_strip_mask(self.h, int(mask)))

    def fill(self, field, value):
        check(lib().gwm_state_fill(self.h, field, float(value)))

    def halo_fill(self, field):
        check(lib().gwm_halo_fill(self.h, field))

    def sync_interfaces(self, field):
        check(lib().gwm_sync_interfaces(self.h, field))

    def capture_boundary(self):
        check(lib().gwm_state_capture_boundary(self.h))

    # operators
    def pressure(self):
        check(lib().gwm_pressure(self.h))

    def slow_tendencies(self, cfg):
        check(lib().gwm_slow_tendencies(self.h, ctypes.byref(cfg)))

    def fast_modes(self, cfg):
        check(lib().gwm_fast_modes(self.h, ctypes.byref(cfg)))

    def ssprk3_step(self, cfg, boundary=None):
        b = None if boundary is None else ctypes.byref(boundary)
        check(lib().gwm_ssprk3_step(self.h, ctypes.byref(cfg), b))

    def tracers_ssprk3(self, cfg, boundary_mode=BOUNDARY_NONE):
        check(lib().gwm_tracers_ssprk3(self.h, ctypes.byref(cfg), boundary_mode))

    def tracers_ssprk3_callback(self, cfg, fn, boundary_mode=BOUNDARY_NONE):
        """Tracer SSPRK3 with a host boundary plug-in: fn(stage, sim_time) is
        called on each stage input with a StageView of the stage's tracer
        fields and returns nothing (an exception fails the step)."""
        errors = []

        def tramp(view, t, _user):
            try:
                fn(StageView(self, view), t)
                return 0
            except Exception as e:  # reported after the call returns
                errors.append(e)
                return 1

        cb = BOUNDARY_FN(tramp)
        rc = lib().gwm_tracers_ssprk3_callback(self.h, ctypes.byref(cfg), boundary_mode, cb, None)
        if errors:
            raise errors[0]
        check(rc)

    def warm_rain(self, cfg, metric_fallout=True):
        check(lib().gwm_warm_rain(self.h, ctypes.byref(cfg), int(metric_fallout)))

    def apply_boundary(self, mode, width=0.0):
        check(lib().gwm_apply_boundary(self.h, mode, float(width)))

    def moist_step(self, cfg, wr, boundary=None):
        b = None if boundary is None else ctypes.byref(boundary)
        check(lib().gwm_moist_step(self.h, ctypes.byref(cfg), ctypes.byref(wr), b))

    def moist_steps_graph(self, cfg, wr, nsteps, surrogate=None):
        """nsteps moist steps replayed from a captured two-step HIP graph."""
        check(lib().gwm_moist_steps_graph(self.h, ctypes.byref(cfg), ctypes.byref(wr),
                                          None if surrogate is None else surrogate.h,
                                          int(nsteps)))

    def copy_from(self, other):
        """Device-to-device copy of every field of another state of the grid."""
        check(lib().gwm_state_copy(self.h, other.h))

    def warm_rain_surrogate(self, cfg, surrogate, metric_fallout=True):
        """Warm rain with the surrogate source terms; returns the number of
        cells that fell back to the plain scheme."""
        n = ctypes.c_longlong()
        check(lib().gwm_warm_rain_surrogate(self.h, ctypes.byref(cfg), surrogate.h,
                                            int(metric_fallout), ctypes.byref(n)))
        return n.value

    def moist_step_surrogate(self, cfg, wr, surrogate, boundary=None):
        b = None if boundary is None else ctypes.byref(boundary)
        check(lib().gwm_moist_step_surrogate(self.h, ctypes.byref(cfg), ctypes.byref(wr),
                                             surrogate.h, b))

    def plan_view(self, slice_k, steps, dt, with_precip=True):
        """Plan-view products of gwm_plan_view: (field count, [23][ny][nx]
        array in the order of PLAN_VIEW_NAMES; planes past the count are 0)."""
        out = np.zeros((PLAN_VIEW_FIELDS, self.grid.ny, self.grid.nx), np.float64)
        nf = ctypes.c_int(0)
        check(lib().gwm_plan_view(self.h, int(slice_k), int(steps), float(dt), int(with_precip),
                                  _ptr(out), ctypes.byref(nf)))
        return nf.value, out

    def summary(self, noalloc=False):
        """Budget sums (gwm_budget_sums; noalloc=True: the allocation-free
        gwm_budget_sums_noalloc)."""
        out = np.zeros(32, np.float64)
        fn = lib().gwm_budget_sums_noalloc if noalloc else lib().gwm_budget_sums
        check(fn(self.h, _ptr(out)))
        return dict(zip(SUMMARY_KEYS, out.tolist()))

    # initial states
    def init_constant(self, rho, theta, u=0.0, v=0.0, w=0.0):
        check(lib().gwm_init_constant(self.h, rho, theta, u, v, w))

    def init_hydrostatic_rest(self, rho_surface, theta_ref):
        check(lib().gwm_init_hydrostatic_rest(self.h, rho_surface, theta_ref))

    def init_bubble(self, rho_background=1.0, theta_background=300.0, u=0.0, v=0.0, w=0.0,
                    center=(0.5, 0.5, 0.5), radius=(0.1, 0.1, 0.1), theta_perturbation=2.0,
                    sign=+1):
        p = np.array([rho_background, theta_background, u, v, w, *center, *radius,
                      theta_perturbation], np.float64)
        check(lib().gwm_init_bubble(self.h, _ptr(p), int(sign)))

    def init_mountain_wave(self, rho_surface=1.18, theta_ref=300.0, u=15.0, v=0.0, w=0.0):
        check(lib().gwm_init_mountain_wave(self.h, rho_surface, theta_ref, u, v, w))

    def init_constant_n(self, theta_surface=288.0, brunt_vaisala=0.01, p_surface=1.0e5, u=10.0,
                        v=0.0):
        check(lib().gwm_init_constant_n(self.h, theta_surface, brunt_vaisala, p_surface, u, v))

    def init_convection(self, p_surface=1.0e5, seed=0, noise_amplitude=0.5, noise_top_m=1000.0):
        check(lib().gwm_init_convection(self.h, p_surface, seed, noise_amplitude, noise_top_m))


class DeviceArray:
    """A device allocation of the context (raw bytes), with host copies."""

    def __init__(self, ctx, shape, dtype=np.float32):
        self.ctx = ctx
        self.shape = tuple(int(v) for v in np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = ctypes.c_void_p()
        check(lib().gwm_device_alloc(ctx.h, self.nbytes, ctypes.byref(self.ptr)))

    def upload(self, arr):
        a = np.ascontiguousarray(arr, self.dtype)
        if a.nbytes != self.nbytes:
            raise GwmError("DeviceArray.upload: size mismatch")
        check(lib().gwm_memcpy_h2d(self.ctx.h, self.ptr, _ptr(a), a.nbytes))

    def download(self):
        out = np.empty(self.shape, self.dtype)
        check(lib().gwm_memcpy_d2h(self.ctx.h, _ptr(out), self.ptr, out.nbytes))
        return out

    def download_rows(self, rows):
        """Host copy of the first `rows` rows."""
        rows = min(int(rows), self.shape[0])
        out = np.empty((rows,) + self.shape[1:], self.dtype)
        check(lib().gwm_memcpy_d2h(self.ctx.h, _ptr(out), self.ptr, out.nbytes))
        return out

    def download_range(self, row, rows):
        """Host copy of `rows` rows from row `row` on."""
        row = int(row)
        rows = int(rows)
        if row < 0 or rows < 0 or row + rows > self.shape[0]:
            raise ValueError("download_range: rows outside the array")
        out = np.empty((rows,) + self.shape[1:], self.dtype)
        check(lib().gwm_memcpy_d2h(self.ctx.h, _ptr(out), self.offset(row), out.nbytes))
        return out

    def offset(self, rows):
        """Pointer to row `rows` of a 2-D array."""
        row_bytes = self.nbytes // self.shape[0]
        return ctypes.c_void_p(self.ptr.value + rows * row_bytes)

    def free(self):
        if self.ptr and not _shutting_down:
            lib().gwm_device_free(self.ctx.h, self.ptr)
        self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class StageView:
    """A stage state handed to a boundary plug-in (valid during the call):
    the owner's grid and tracer count, the stage's fields."""

    def __init__(self, owner, handle):
        self.owner = owner
        self.h = ctypes.c_void_p(handle)

    def get(self, field):
        out = np.empty(self.owner.storage_shape(field), self.owner.ctx.dtype)
        check(lib().gwm_halo_fill(self.h, field))
        check(lib().gwm_state_download(self.h, field, _ptr(out)))
        h = self.owner.grid.halo
        return out[:, h:out.shape[1] - h, h:out.shape[2] - h]

    def set_storage(self, field, arr):
        a = np.ascontiguousarray(arr, self.owner.ctx.dtype)
        check(lib().gwm_state_upload(self.h, field, _ptr(a)))

    def storage(self, field):
        out = np.empty(self.owner.storage_shape(field), self.owner.ctx.dtype)
        check(lib().gwm_state_download(self.h, field, _ptr(out)))
        return out


class Surrogate:
    """The warm-rain surrogate MLP (parameters on the device)."""

    def __init__(self, ctx, seed=0):
        self.ctx = ctx
        self.h = ctypes.c_void_p()
        check(lib().gwm_surrogate_create(ctx.h, int(seed), ctypes.byref(self.h)))

    def __del__(self):
        if getattr(self, "h", None) and not _shutting_down:
            lib().gwm_surrogate_destroy(self.h)
        self.h = None

    def set_norm(self, x_mean, x_std, y_mean, y_std, box_lo, box_hi):
        n = SurrogateNorm()
        for name, v, k in (("x_mean", x_mean, 7), ("x_std", x_std, 7), ("y_mean", y_mean, 4),
                           ("y_std", y_std, 4), ("box_lo", box_lo, 7), ("box_hi", box_hi, 7)):
            getattr(n, name)[:] = [float(t) for t in np.asarray(v, np.float64).reshape(k)]
        check(lib().gwm_surrogate_set_norm(self.h, ctypes.byref(n)))

    def norm(self):
        n = SurrogateNorm()
        check(lib().gwm_surrogate_get_norm(self.h, ctypes.byref(n)))
        return {k: np.array(getattr(n, k)[:], np.float32) for k, _ in SurrogateNorm._fields_}

    def params(self):
        out = np.empty(SURROGATE_PARAMS, np.float32)
        check(lib().gwm_surrogate_get_params(self.h, _ptr(out)))
        return out

    def set_params(self, p):
        a = np.ascontiguousarray(p, np.float32)
        if a.size != SURROGATE_PARAMS:
            raise GwmError("surrogate parameters: wrong size")
        check(lib().gwm_surrogate_set_params(self.h, _ptr(a)))

    def train(self, x, y, n, batch=4096, epochs=3, lr=2e-3, lr_final=1e-4, x_row=0):
        losses = np.zeros(epochs, np.float64)
        check(lib().gwm_surrogate_train(self.h, x.offset(x_row), y.offset(x_row), int(n),
                                        int(batch), int(epochs), float(lr), float(lr_final),
                                        _ptr(losses)))
        return losses

    def loss_grad(self, x, y, n):
        loss = ctypes.c_double()
        g = np.empty(SURROGATE_PARAMS, np.float32)
        check(lib().gwm_surrogate_loss_grad(self.h, x.ptr, y.ptr, int(n), ctypes.byref(loss),
                                            _ptr(g)))
        return loss.value, g

    def infer(self, x, n, y):
        check(lib().gwm_surrogate_infer(self.h, x.ptr, int(n), y.ptr))

    def last_counts(self):
        """(fallback cells, network cells) of the last surrogate warm-rain
        call on this network's context."""
        fb, nw = ctypes.c_longlong(), ctypes.c_longlong()
        check(lib().gwm_surrogate_last_counts(self.h, ctypes.byref(fb), ctypes.byref(nw)))
        return fb.value, nw.value


def surrogate_sample(state, cfg, seed, n, x, y, cell=None, row=0):
    """Training pairs from a state into device arrays x [.., 7], y [.., 4]
    (starting at `row`); `cell` optionally receives the cell indices."""
    check(lib().gwm_surrogate_sample(state.h, ctypes.byref(cfg), int(seed), int(n),
                                     x.offset(row), y.offset(row),
                                     None if cell is None else cell.offset(row)))


def ensemble_tuning_report():
    """Step-pair times of the batched ensemble's level-chunk candidates, per
    ensemble signature (gwm_ensemble_tuning_report)."""
    buf = ctypes.create_string_buffer(65536)
    check(lib().gwm_ensemble_tuning_report(buf, len(buf)))
    return buf.value.decode()


def ensemble_steps(states, cfg, wr, nsteps, surrogates=None, streams=False):
    """Advance states nsteps moist steps together: one launch per kernel of
    the step over all members (gwm_ensemble_steps), or with streams=True one
    context (stream) per member, concurrently (gwm_ensemble_steps_streams)."""
    n = len(states)
    arr = (ctypes.c_void_p * n)(*[st.h.value for st in states])
    sur = None
    if surrogates is not None:
        sur = (ctypes.c_void_p * n)(*[None if m is None else m.h.value for m in surrogates])
    fn = lib().gwm_ensemble_steps_streams if streams else lib().gwm_ensemble_steps
    check(fn(arr, n, ctypes.byref(cfg), ctypes.byref(wr), sur, int(nsteps)))
