"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# GWM_ORACLE_LIB points the checker at another build of this source (the
# mutation tests load deliberately broken copies of the oracle that way).
LIB_PATH = os.environ.get("GWM_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")

RHO, RHO_THETA, MOM_U, MOM_V, MOM_W, PRESSURE = 0, 1, 2, 3, 4, 5
TEND_RHO, TEND_RHO_THETA, TEND_U, TEND_V, TEND_W = 6, 7, 8, 9, 10
PRECIP, TRACER0 = 11, 16

METRIC_Z_INTERFACES, METRIC_Z_CENTERS, METRIC_INV_DZ_CELL = 0, 1, 2
METRIC_INV_DZ_FACE, METRIC_TERRAIN, METRIC_SLOPE_X, METRIC_SLOPE_Y = 3, 4, 5, 6
METRIC_FLAT_Z_INTERFACES, METRIC_FLAT_Z_CENTERS = 7, 8
METRIC_WEIGHT_INTERFACES, METRIC_WEIGHT_CENTERS = 9, 10

BOUNDARY_NONE, BOUNDARY_STRIPS, BOUNDARY_RELAX = 0, 1, 2

SUMMARY_KEYS = [
    "total_dry_mass", "total_rho_theta_m", "total_horizontal_momentum_x",
    "total_horizontal_momentum_y", "total_vertical_momentum", "rho_min", "rho_max",
    "theta_min", "theta_max", "u_min", "u_max", "v_min", "v_max", "w_min", "w_max",
    "vapor_path", "cloud_path", "rain_path", "precip_sum", "precip_mean", "precip_max",
    "precip_cells", "precipitating_cells", "qv_min", "qv_max", "qc_min", "qc_max",
    "qr_min", "qr_max",
]


PLAN_VIEW_NAMES = [
    "terrain_height", "z_center", "rho_d", "theta_m", "air_pressure", "air_temperature",
    "u_velocity", "v_velocity", "wind_speed", "w_velocity", "specific_humidity",
    "relative_humidity", "dewpoint", "cloud_water_mixing_ratio", "rain_water_mixing_ratio",
    "total_condensate", "column_cloud_water", "column_rain_water", "column_total_condensate",
    "column_rain_fraction", "synthetic_reflectivity", "accumulated_surface_precipitation",
    "mean_surface_precipitation_rate",
]
PLAN_VIEW_FIELDS = len(PLAN_VIEW_NAMES)


class StepCfg(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("fast_substeps", ctypes.c_int),
                ("gravity", ctypes.c_double),
                ("fast_update_horizontal_momentum", ctypes.c_int),
                ("fast_update_density", ctypes.c_int)]


def step_cfg(dt=2.0, fast_substeps=1, gravity=9.81, horizontal=True, density=True):
    return StepCfg(dt, fast_substeps, gravity, int(horizontal), int(density))


class WarmRainCfg(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("condensation_relaxation", ctypes.c_double),
                ("cloud_autoconversion_threshold", ctypes.c_double),
                ("cloud_autoconversion_rate", ctypes.c_double),
                ("rain_evaporation_rate", ctypes.c_double),
                ("rain_terminal_velocity", ctypes.c_double),
                ("sedimentation_layer_depth_m", ctypes.c_double),
                ("enable_latent_heating", ctypes.c_int),
                ("accretion_rate", ctypes.c_double)]


def warm_rain_cfg(dt=2.0, condensation_relaxation=1.0, cloud_autoconversion_threshold=1.0e-4,
                  cloud_autoconversion_rate=1.0e-3, rain_evaporation_rate=5.0e-3,
                  rain_terminal_velocity=0.0, sedimentation_layer_depth_m=0.0,
                  enable_latent_heating=True, accretion_rate=0.0):
    """Defaults of WarmRainConfig (include/gwm/physics/warm_rain.hpp:14-23)."""
    return WarmRainCfg(dt, condensation_relaxation, cloud_autoconversion_threshold,
                       cloud_autoconversion_rate, rain_evaporation_rate,
                       rain_terminal_velocity, sedimentation_layer_depth_m,
                       int(enable_latent_heating), accretion_rate)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-s", "-C", HERE], check=True)
        L = ctypes.CDLL(LIB_PATH)
        vp, ip, dp = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.orc_world_create.restype = vp
        L.orc_world_create.argtypes = [ip, ip, ip, ip, ip, ip, ip, ip, ip, dp, dp, dp, vp, vp, dp, ip]
        L.orc_world_destroy.argtypes = [vp]
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_num_ranks.argtypes = [vp]
        L.orc_rank_desc.argtypes = [vp, ip, vp]
        L.orc_neighbors.argtypes = [vp, ip, vp]
        L.orc_rank_storage_size.restype = ctypes.c_long
        L.orc_rank_storage_size.argtypes = [vp, ip, ip, vp]
        for name in ("orc_get_rank_storage", "orc_set_rank_storage"):
            getattr(L, name).argtypes = [vp, ip, ip, vp]
        for name in ("orc_get_global", "orc_set_global", "orc_get_metric"):
            getattr(L, name).argtypes = [vp, ip, vp]
        L.orc_fill.argtypes = [vp, ip, dp]
        for name in ("orc_exchange", "orc_sync_interfaces"):
            getattr(L, name).argtypes = [vp, ip]
        for name in ("orc_pressure", "orc_momentum_flux_only", "orc_pressure_gradient_only",
                     "orc_set_boundary_from_current"):
            getattr(L, name).argtypes = [vp]
        for name in ("orc_slow_tendencies", "orc_fast_modes"):
            getattr(L, name).argtypes = [vp, ctypes.POINTER(StepCfg)]
        L.orc_ssprk3_step.argtypes = [vp, ctypes.POINTER(StepCfg), ip, dp, dp]
        L.orc_tracers_ssprk3.argtypes = [vp, ctypes.POINTER(StepCfg), ip]
        L.orc_warm_rain.argtypes = [vp, ctypes.POINTER(WarmRainCfg), ip]
        L.orc_summary.argtypes = [vp, vp]
        L.orc_init_constant.argtypes = [vp, dp, dp, dp, dp, dp]
        L.orc_init_hydrostatic_rest.argtypes = [vp, dp, dp]
        L.orc_init_bubble.argtypes = [vp, vp, ip]
        L.orc_init_mountain_wave.argtypes = [vp, dp, dp, dp, dp, dp]
        L.orc_column_reference.argtypes = [ip, ip, vp, vp, dp, dp, vp, vp]
        L.orc_thermo.restype = dp
        L.orc_thermo.argtypes = [ip, ip, dp, dp, dp]
        L.orc_apply_boundary.argtypes = [vp, ip, dp]
        L.orc_set_tracer_strip_mask.argtypes = [vp, ctypes.c_uint]
        L.orc_plan_view.argtypes = [vp, ip, ip, dp, ip, vp]
        L.orc_time_steps.restype = dp
        L.orc_time_steps.argtypes = [vp, ctypes.POINTER(StepCfg), ctypes.POINTER(WarmRainCfg),
                                     ip, ip, ip]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class OracleError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise OracleError(lib().orc_last_error().decode())


class World:
    """A global grid, its virtual-rank layout and the state, on the CPU."""

    def __init__(self, nx, ny, nz, halo=1, ranks_x=1, ranks_y=1, periodic_x=True,
                 periodic_y=True, dx=1000.0, dy=1000.0, z_top=16000.0, z_interfaces=None,
                 terrain=None, taper_eta=0.25, n_tracers=0, precision=8):
        self.dtype = np.float64 if precision == 8 else np.float32
        self.nx, self.ny, self.nz, self.halo = nx, ny, nz, halo
        self._zi = None if z_interfaces is None else np.ascontiguousarray(z_interfaces, np.float64)
        self._tr = None if terrain is None else np.ascontiguousarray(terrain, np.float32).reshape(-1)
        self.n_tracers = n_tracers
        self.h = lib().orc_world_create(
            precision, nx, ny, nz, halo, ranks_x, ranks_y, int(periodic_x), int(periodic_y),
            float(dx), float(dy), float(z_top),
            None if self._zi is None else _ptr(self._zi),
            None if self._tr is None else _ptr(self._tr), float(taper_eta), n_tracers)
        if not self.h:
            raise OracleError(lib().orc_last_error().decode())

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_world_destroy(self.h)
            self.h = None

    # --- layout -----------------------------------------------------------
    @property
    def num_ranks(self):
        return lib().orc_num_ranks(self.h)

    def rank_desc(self, rank):
        out = np.zeros(13, np.int32)
        _check(lib().orc_rank_desc(self.h, rank, _ptr(out)))
        keys = ["rank", "coord_x", "coord_y", "ranks_x", "ranks_y", "global_nx", "global_ny",
                "nz", "halo", "i_begin", "i_end", "j_begin", "j_end"]
        return dict(zip(keys, out.tolist()))

    def neighbors(self, rank):
        out = np.zeros(4, np.int32)
        _check(lib().orc_neighbors(self.h, rank, _ptr(out)))
        return dict(zip(["west", "east", "south", "north"], out.tolist()))

    # --- storage ----------------------------------------------------------
    def rank_storage(self, rank, field):
        """Full halo-padded storage of a field on a rank, shaped (nz, py, px)."""
        dims = np.zeros(4, np.int32)
        n = lib().orc_rank_storage_size(self.h, rank, field, _ptr(dims))
        if n < 0:
            raise OracleError(lib().orc_last_error().decode())
        out = np.empty(n, self.dtype)
        _check(lib().orc_get_rank_storage(self.h, rank, field, _ptr(out)))
        fx, fy, fz, h = dims.tolist()
        return out.reshape(fz, fy + 2 * h, fx + 2 * h)

    def set_rank_storage(self, rank, field, arr):
        a = np.ascontiguousarray(arr, self.dtype)
        _check(lib().orc_set_rank_storage(self.h, rank, field, _ptr(a)))

    def global_shape(self, field):
        gx = self.nx + (1 if field in (MOM_U, TEND_U) else 0)
        gy = self.ny + (1 if field in (MOM_V, TEND_V) else 0)
        gz = self.nz + (1 if field in (MOM_W, TEND_W) else 0)
        if field == PRECIP:
            gz = 1
        return (gz, gy, gx)

    def get(self, field):
        out = np.empty(self.global_shape(field), self.dtype)
        _check(lib().orc_get_global(self.h, field, _ptr(out)))
        return out

    def set(self, field, arr):
        a = np.ascontiguousarray(np.broadcast_to(arr, self.global_shape(field)), self.dtype)
        _check(lib().orc_set_global(self.h, field, _ptr(a)))

    def fill(self, field, value):
        _check(lib().orc_fill(self.h, field, float(value)))

    def metric(self, which):
        nx, ny, nz = self.nx, self.ny, self.nz
        shapes = {0: (nz + 1, ny, nx), 1: (nz, ny, nx), 2: (nz, ny, nx), 3: (nz + 1, ny, nx),
                  4: (ny, nx), 5: (ny, nx), 6: (ny, nx), 7: (nz + 1,), 8: (nz,), 9: (nz + 1,),
                  10: (nz,)}
        dt = np.float64 if which >= 7 else self.dtype
        out = np.empty(shapes[which], dt)
        _check(lib().orc_get_metric(self.h, which, _ptr(out)))
        return out

    # --- operators --------------------------------------------------------
    def exchange(self, field):
        _check(lib().orc_exchange(self.h, field))

    def sync_interfaces(self, field):
        _check(lib().orc_sync_interfaces(self.h, field))

    def pressure(self):
        _check(lib().orc_pressure(self.h))

    def slow_tendencies(self, cfg):
        _check(lib().orc_slow_tendencies(self.h, ctypes.byref(cfg)))

    def momentum_flux_only(self):
        _check(lib().orc_momentum_flux_only(self.h))

    def pressure_gradient_only(self):
        _check(lib().orc_pressure_gradient_only(self.h))

    def fast_modes(self, cfg):
        _check(lib().orc_fast_modes(self.h, ctypes.byref(cfg)))

    def set_boundary_from_current(self):
        _check(lib().orc_set_boundary_from_current(self.h))

    def set_tracer_strip_mask(self, mask):
        """Bit t set: tracer t takes the strip boundary (default all)."""
        _check(lib().orc_set_tracer_strip_mask(self.h, ctypes.c_uint(mask)))

    def apply_boundary(self, mode, width=0.0):
        _check(lib().orc_apply_boundary(self.h, mode, float(width)))

    def ssprk3_step(self, cfg, boundary=BOUNDARY_NONE, relax_width=0.0, relax_tau=0.0):
        _check(lib().orc_ssprk3_step(self.h, ctypes.byref(cfg), boundary, relax_width, relax_tau))

    def tracers_ssprk3(self, cfg, boundary=BOUNDARY_NONE):
        _check(lib().orc_tracers_ssprk3(self.h, ctypes.byref(cfg), boundary))

    def warm_rain(self, cfg, metric_fallout=True):
        _check(lib().orc_warm_rain(self.h, ctypes.byref(cfg), int(metric_fallout)))

    def summary(self):
        out = np.zeros(32, np.float64)
        _check(lib().orc_summary(self.h, _ptr(out)))
        return dict(zip(SUMMARY_KEYS, out.tolist()))

    def plan_view(self, slice_k, steps, dt, with_precip=True):
        """Plan-view products: (field count, [23][ny][nx] array)."""
        out = np.zeros((PLAN_VIEW_FIELDS, self.ny, self.nx), np.float64)
        nf = lib().orc_plan_view(self.h, int(slice_k), int(steps), float(dt), int(with_precip),
                                 _ptr(out))
        if nf < 0:
            raise OracleError(lib().orc_last_error().decode())
        return nf, out

    def time_steps(self, cfg, wr=None, moist=False, steps=1, threads=0):
        wr = wr or warm_rain_cfg()
        s = lib().orc_time_steps(self.h, ctypes.byref(cfg), ctypes.byref(wr), int(moist),
                                 steps, threads)
        if s < 0:
            raise OracleError(lib().orc_last_error().decode())
        return s

    # --- initial states ---------------------------------------------------
    def init_constant(self, rho, theta, u=0.0, v=0.0, w=0.0):
        _check(lib().orc_init_constant(self.h, rho, theta, u, v, w))

    def init_hydrostatic_rest(self, rho_surface, theta_ref):
        _check(lib().orc_init_hydrostatic_rest(self.h, rho_surface, theta_ref))

    def init_bubble(self, rho_background=1.0, theta_background=300.0, u=0.0, v=0.0, w=0.0,
                    center=(0.5, 0.5, 0.5), radius=(0.1, 0.1, 0.1), theta_perturbation=2.0,
                    sign=+1):
        p = np.array([rho_background, theta_background, u, v, w, *center, *radius,
                      theta_perturbation], np.float64)
        _check(lib().orc_init_bubble(self.h, _ptr(p), int(sign)))

    def init_mountain_wave(self, rho_surface=1.18, theta_ref=300.0, u=15.0, v=0.0, w=0.0):
        _check(lib().orc_init_mountain_wave(self.h, rho_surface, theta_ref, u, v, w))


def column_reference(theta, zc, p_bottom, gravity=9.81, precision=8):
    """K11 column hydrostatic reference of one column: (p_ref, rho_ref)."""
    th = np.ascontiguousarray(theta, np.float64)
    z = np.ascontiguousarray(zc, np.float64)
    p = np.zeros_like(th)
    r = np.zeros_like(th)
    _check(lib().orc_column_reference(precision, th.size, _ptr(th), _ptr(z), float(gravity),
                                      float(p_bottom), _ptr(p), _ptr(r)))
    return p, r
