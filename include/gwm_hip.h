/*
 * gwm_hip.h -- C-ABI of the MI355X (gfx950) dry + warm-rain time step.
 *
 * This is the drop-in boundary for the reference's operator and state
 * contracts (FahrenheitResearch/gpu-wm-next).  The reference exposes a C++
 * API over CUDA managed memory; the C++ host mirror in
 * gpu-wm-next_amd/host/ restores those signatures on top of this ABI.  Each
 * entry point cites the reference interface it replaces.
 *
 * Conventions
 *  - Plain C types only.  Every function returns 0 on success and a nonzero
 *    GWM_E_* code on failure; gwm_last_error() then returns a message
 *    (thread-local), mirroring gwm::require / GWM_CUDA_CHECK
 *    (include/gwm/core/types.hpp:24-28, include/gwm/core/cuda_utils.hpp:10-24).
 *  - One context per host thread, one HIP stream per context.  Calls are
 *    stream-ordered; gwm_sync() waits.  Uploads and downloads are
 *    synchronous.
 *  - Device fields use exactly the reference storage layout
 *    (include/gwm/state/field3d.hpp:49-54):
 *      idx = (k*(ny+2h) + j+h)*(nx+2h) + i+h, x fastest, no halo in k,
 *    with C-grid faces sized (nx+1,ny,nz), (nx,ny+1,nz), (nx,ny,nz+1)
 *    (include/gwm/state/face_field.hpp:16-23).  Host arrays passed to
 *    upload/download use that same layout, halo included.
 *  - One GPU holds the whole horizontal domain (a 1x1 layout).  A virtual-
 *    rank split of the reference is tiling only; gwm_halo_exchange_layout()
 *    reproduces the multi-subdomain halo semantics bit-exactly for callers
 *    that hold per-rank arrays.
 */
#ifndef GWM_HIP_H
#define GWM_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
  GWM_OK = 0,
  GWM_E_INVALID = 1,   /* bad argument or shape mismatch (gwm::require) */
  GWM_E_HIP = 2,       /* HIP runtime error (GWM_CUDA_CHECK) */
  GWM_E_NOMEM = 3,
  GWM_E_UNSUPPORTED = 4
};

/* ---- precision --------------------------------------------------------- */
enum { GWM_F32 = 0, GWM_F64 = 1 };

/* ---- field identifiers (state, scratch outputs, tracers) --------------- */
enum {
  GWM_RHO = 0,          /* rho_d                    DryState::rho_d        */
  GWM_RHO_THETA = 1,    /* rho_d * theta_m          DryState::rho_theta_m  */
  GWM_MOM_U = 2,        /* x-face mass flux         DryState::mom_u        */
  GWM_MOM_V = 3,        /* y-face mass flux         DryState::mom_v        */
  GWM_MOM_W = 4,        /* z-face mass flux         DryState::mom_w        */
  GWM_PRESSURE = 5,     /* output of gwm_pressure()                         */
  GWM_TEND_RHO = 6,     /* DrySlowTendencies (include/gwm/dycore/dry_core.hpp:36-48) */
  GWM_TEND_RHO_THETA = 7,
  GWM_TEND_U = 8,
  GWM_TEND_V = 9,
  GWM_TEND_W = 10,
  GWM_PRECIP = 11,      /* surface bucket, (nx, ny) float, no halo, kg m-2 */
  GWM_TRACER0 = 16      /* tracer t is GWM_TRACER0 + t (rho_d * q)         */
};

/* Warm-rain tracer order (include/gwm/state/tracer_registry.hpp:57-63). */
enum { GWM_QV = 0, GWM_QC = 1, GWM_QR = 2 };

/* ---- opaque handles ---------------------------------------------------- */
typedef struct gwm_ctx gwm_ctx;
typedef struct gwm_grid gwm_grid;
typedef struct gwm_state gwm_state;

/* ---- configuration structs -------------------------------------------- */

/* GridMetrics::make_hybrid_height (include/gwm/domain/grid_metrics.hpp:87-92)
 * plus the RectilinearDomainConfig shape (idealized_domain_builder.hpp:12-31). */
typedef struct {
  int nx, ny, nz, halo;
  int periodic_x, periodic_y;
  double dx, dy;
  double z_top;             /* used when z_interfaces is NULL */
  double terrain_taper_eta; /* reference default 0.25 */
} gwm_grid_desc;

/* DryStepperConfig (include/gwm/dycore/dry_core.hpp:50-56). */
typedef struct {
  double dt;
  int fast_substeps;
  double gravity;
  int fast_update_horizontal_momentum;
  int fast_update_density;
} gwm_step_config;

/* WarmRainConfig (include/gwm/physics/warm_rain.hpp:14-23).  accretion_rate
 * is an extension (0 reproduces the reference). */
typedef struct {
  double dt;
  double condensation_relaxation;
  double cloud_autoconversion_threshold;
  double cloud_autoconversion_rate;
  double rain_evaporation_rate;
  double rain_terminal_velocity;
  double sedimentation_layer_depth_m;
  int enable_latent_heating;
  double accretion_rate;
} gwm_warm_rain_config;

/* Boundary handling inside the stepper.  NONE == NullBoundaryUpdater
 * (src/dycore/dry_core.cu:336-339); STRIPS == apply_reference_boundaries
 * (src/dycore/boundary_conditions.cpp:111-161) against the state captured by
 * gwm_state_capture_boundary(); RELAX is the relaxation-zone extension. */
enum { GWM_BOUNDARY_NONE = 0, GWM_BOUNDARY_STRIPS = 1, GWM_BOUNDARY_RELAX = 2 };

/* Host callback run once per RK stage, like BoundaryUpdater::apply
 * (include/gwm/dycore/dry_core.hpp:58-64).  `stage` is a view of the stage
 * state: gwm_state_download/upload and gwm_apply_boundary work on it. */
typedef int (*gwm_boundary_fn)(gwm_state* stage, double sim_time, void* user);

typedef struct {
  int mode;             /* GWM_BOUNDARY_* */
  double relax_width;   /* cells, RELAX only */
  gwm_boundary_fn callback; /* optional, runs before `mode` (it may refresh the
                               boundary reference with gwm_state_set_boundary) */
  void* user;
} gwm_boundary_config;

/* Summary slots written by gwm_budget_sums (double[GWM_SUMMARY_N]).  Sums are
 * compensated (Neumaier) FP64 device reductions; keys follow
 * summarize_dry_states (src/dycore/dry_diagnostics.cpp:11-96) and
 * summarize_runtime_state (src/core/runtime_summary.cpp:27-178). */
enum {
  GWM_S_TOTAL_DRY_MASS = 0,
  GWM_S_TOTAL_RHO_THETA = 1,
  GWM_S_TOTAL_MOM_X = 2,
  GWM_S_TOTAL_MOM_Y = 3,
  GWM_S_TOTAL_MOM_Z = 4,
  GWM_S_RHO_MIN = 5,
  GWM_S_RHO_MAX = 6,
  GWM_S_THETA_MIN = 7,
  GWM_S_THETA_MAX = 8,
  GWM_S_U_MIN = 9,
  GWM_S_U_MAX = 10,
  GWM_S_V_MIN = 11,
  GWM_S_V_MAX = 12,
  GWM_S_W_MIN = 13,
  GWM_S_W_MAX = 14,
  GWM_S_VAPOR_PATH = 15,
  GWM_S_CLOUD_PATH = 16,
  GWM_S_RAIN_PATH = 17,
  GWM_S_PRECIP_SUM = 18,
  GWM_S_PRECIP_MEAN = 19,
  GWM_S_PRECIP_MAX = 20,
  GWM_S_PRECIP_CELLS = 21,
  GWM_S_PRECIPITATING_CELLS = 22,
  GWM_S_TRACER_QMIN0 = 23, /* min/max mixing ratio of tracers 0..2 at 23..28 */
  /* Count of NaN / infinite values over rho, rho*theta, the momenta, the
   * tracers and the precipitation bucket.  Sums and extremes cover the finite
   * values; the extremes of a field holding any non-finite value are NaN. */
  GWM_S_NONFINITE = 29,
  GWM_SUMMARY_N = 32
};

/* ---- context ------------------------------------------------------------ */
const char* gwm_last_error(void);
int gwm_ctx_create(int device, int precision, gwm_ctx** out);
int gwm_ctx_destroy(gwm_ctx* ctx);
int gwm_sync(gwm_ctx* ctx);
int gwm_ctx_precision(const gwm_ctx* ctx);
/* Per-kernel HIP-event timing on the context stream (off by default). */
int gwm_ctx_set_kernel_timing(gwm_ctx* ctx, int enable);
/* Kernel names are stable strings ("fast_substep", "slow_stage", ...).
 * Returns launches and total milliseconds recorded since the last reset. */
int gwm_ctx_kernel_stats(gwm_ctx* ctx, const char* name, long* launches, double* total_ms);
int gwm_ctx_reset_kernel_stats(gwm_ctx* ctx);
/* Column kernels split each column into level chunks, and the flat fast
 * kernel runs tiles of 8 or 4 rows (64 columns wide); by default both are
 * chosen per kernel, grid shape and precision by timing the candidates
 * (chunks {1, 2, 4, 8}, and 3, 6, 12 for the slow and tracer stages, times
 * the tile heights) at first use (the
 * autotuner).  chunks > 0 fixes the launch code: the chunk count in the low
 * 4 bits, bit 4 set for 4-row tiles, bit 6 set to keep every fast substep
 * marching upwards (by default every other substep of a flat whole-column
 * launch marches down); 0 returns to the autotuner.  Results do not depend
 * on the choice.  No counterpart in the reference (tuning
 * control of this implementation). */
int gwm_ctx_set_chunks(gwm_ctx* ctx, int chunks);
/* "kernel:NXxNYxNZ:prec:variant=code;..." for every tuned launch shape
 * (code as for gwm_ctx_set_chunks). */
int gwm_ctx_tuning_report(gwm_ctx* ctx, char* buf, size_t size);

/* Diagnostic, no counterpart in the reference: evaluates one of the FP32
 * arithmetic helpers of the step kernels elementwise on the device, so their
 * accuracy can be measured (dev_y[i] = f(dev_x[i]), n floats each):
 *   GWM_MATH_POW_EOS    x^(1/(1-kappa))  (equation of state)
 *   GWM_MATH_POW_EXNER  x^kappa          (Exner function)
 *   GWM_MATH_EXP        exp(x)
 *   GWM_MATH_LOG1P      log(1+x)
 *   GWM_MATH_EXPM1      exp(x)-1
 *   GWM_MATH_RCP        1/x
 *   GWM_MATH_POW1PM1    (1+x)^(1/(1-kappa)) - 1  (pressure departure)      */
enum {
  GWM_MATH_POW_EOS = 0,
  GWM_MATH_POW_EXNER = 1,
  GWM_MATH_EXP = 2,
  GWM_MATH_LOG1P = 3,
  GWM_MATH_EXPM1 = 4,
  GWM_MATH_RCP = 5,
  GWM_MATH_POW1PM1 = 6
};
int gwm_fp32_math(gwm_ctx* ctx, int fn, const void* dev_x, void* dev_y, long n);

/* ---- grid / metrics ------------------------------------------------------
 * Replaces GridMetrics::make_hybrid_height (grid_metrics.hpp:87-239) and
 * build_rectilinear_domain (src/domain/idealized_domain_builder.cpp:39-64).
 * z_interfaces: NULL for the reference's uniform eta levels, or nz+1 flat
 * interface heights (stretched-level extension).  terrain: NULL (flat) or
 * nx*ny floats, x fastest.  The metric fields are built on the device. */
int gwm_grid_create(gwm_ctx* ctx, const gwm_grid_desc* desc, const double* z_interfaces,
                    const float* terrain, gwm_grid** out);
int gwm_grid_destroy(gwm_grid* grid);
/* Metric download, ids as in gwm_oracle.h: 0 z_interfaces, 1 z_centers,
 * 2 inv_dz_cell, 3 inv_dz_face, 4 terrain, 5 slope_x, 6 slope_y (context
 * precision, global (k*ny+j)*nx+i layout); 7..10 flat FP64 columns. */
int gwm_grid_download_metric(gwm_grid* grid, int which, void* host);
int gwm_grid_terrain_active(const gwm_grid* grid);

/* Terrain builders of the product drivers: the reference's cosine mountain
 * (idealized_domain_builder.cpp:14-35) and the Witch-of-Agnesi ridge
 * h(x) = h0 / (1 + ((x - xc)/a)^2), uniform in y (extension). */
int gwm_terrain_cosine_mountain(int nx, int ny, double cx, double cy, double hx, double hy,
                                double height, float* out);
int gwm_terrain_agnesi_ridge(int nx, int ny, double dx, double xc_m, double half_width_m,
                             double height_m, float* out);
/* Stretched flat interfaces: dz grows smoothly from dz_bottom to dz_top and
 * the column is rescaled to end exactly at z_top (extension). */
int gwm_stretched_levels(int nz, double dz_bottom, double dz_top, double z_top, double* out);

/* ---- state ----------------------------------------------------------------
 * DryState (include/gwm/dycore/dry_core.hpp:14-34) + TracerState
 * (include/gwm/state/tracer_state.hpp:12-119) + WarmRainSurfaceAccumulation
 * (include/gwm/physics/warm_rain.hpp:25-46), all on the device. */
int gwm_state_create(gwm_ctx* ctx, const gwm_grid* grid, int n_tracers, gwm_state** out);
int gwm_state_destroy(gwm_state* state);
/* FP32 perturbation form (extension, no reference counterpart): rho_d and
 * rho_d*theta_m are stored as FP32 departures from per-level bases, the FP64
 * level means of the state they were loaded from; pressure enters through
 * p' = p_b expm1(gamma log1p((rho theta)'/(rho theta)_b)), so pressure
 * differences keep their relative precision and the dry-mass budget is the
 * FP64 base sum plus a compensated sum of departures.  Needs an FP32
 * context; on flat terrain a doubly periodic grid.  On terrain (some
 * |h| > 1e-6 m) the base is 3-D: the FP64 state loaded, cell by cell, with
 * its pressure-gradient and gravity forcing held per face, so that K14
 * acts on the departures without the column reference (whose terms cancel
 * against its recurrence up to a residual that the base forcing carries);
 * open or periodic sides, dry states only
 * (n_tracers 0; gwm_plan_view and the ensemble entry points refuse it).
 * Uploads, downloads and fills of GWM_RHO / GWM_RHO_THETA take and return
 * full values (an upload re-bases that field); every other operator accepts
 * the state. */
int gwm_state_create_perturbation(gwm_ctx* ctx, const gwm_grid* grid, int n_tracers,
                                  gwm_state** out);
int gwm_state_is_perturbation(const gwm_state* state);
/* Element count of a field's storage (halo included). */
long gwm_state_field_size(const gwm_state* state, int field);
int gwm_state_upload(gwm_state* state, int field, const void* host);
int gwm_state_download(gwm_state* state, int field, void* host);
/* Download as FP64 (a perturbation-form departure plus its FP64 base,
 * without rounding to FP32). */
int gwm_state_download_f64(gwm_state* state, int field, double* host);
int gwm_state_fill(gwm_state* state, int field, double value);
int gwm_state_copy(gwm_state* dst, const gwm_state* src);
/* Keep a copy of the current state as the lateral-boundary reference. */
int gwm_state_capture_boundary(gwm_state* state);
/* Replace one field of the lateral-boundary reference (captured before) with
 * host values in the field's storage layout, as gwm_state_upload takes them;
 * on a stage view it writes the owner's reference.  With STRIPS, a boundary
 * callback that refreshes the reference this way is the device form of the
 * prepared-case boundary updaters (PreparedCaseBoundaryUpdater::apply,
 * src/ingest/prepared_case_init.cpp:1032-1047, and the tracer updater,
 * 1064-1083): the stage never leaves the device. */
int gwm_state_set_boundary(gwm_state* state, int field, const void* host);
/* Restart dump.  The reference fixes the artifact contract of a native
 * checkpoint and leaves the data for later (tools/casebuilder/prepare_case.py:376-392,
 * "gwm-next-checkpoint-stub/v1", restart family "gwm-native-checkpoint");
 * this writes that checkpoint: the canonical fields in storage layout, the
 * tracers, the surface bucket, the perturbation form's FP64 bases and the
 * captured boundary reference, each with an FNV-1a checksum, plus a manifest
 * `path`.json that satisfies the stub contract (metadata_json: an optional
 * JSON value stored in it).  Reading into a state of the same grid,
 * precision, form and tracer count restores it bit for bit; a mismatch or a
 * failed checksum is an error that names what differs. */
int gwm_state_checkpoint_write(gwm_state* state, const char* path, const char* metadata_json);
int gwm_state_checkpoint_read(gwm_state* state, const char* path);
/* Tracers the STRIPS boundary overwrites (bit t: tracer t; all by default);
 * the prepared tracer strips cover the tracers a snapshot carries
 * (apply_cached_tracer_boundary_strips, prepared_case_init.cpp:567-638). */
int gwm_state_set_tracer_strip_mask(gwm_state* state, unsigned mask);

/* ---- halo semantics -------------------------------------------------------
 * HaloExchange::exchange_scalar / exchange_face and
 * synchronize_owned_face_interfaces (include/gwm/comm/halo_exchange.hpp:65-91,
 * src/comm/halo_exchange.cpp:148-285,563-620) for the 1x1 layout, in place on
 * one state field (periodic copy or zero across an open side). */
int gwm_halo_fill(gwm_state* state, int field);
int gwm_sync_interfaces(gwm_state* state, int field);
/* Multi-subdomain exchange for per-rank device arrays (virtual ranks):
 * desc13 holds 13 ints per rank (rank, coord_x, coord_y, ranks_x, ranks_y,
 * global_nx, global_ny, nz, halo, i_begin, i_end, j_begin, j_end);
 * orientation 0 cell, 1 x-face, 2 y-face, 3 z-face.  fields[n] is a device
 * pointer of context precision in the reference layout for rank n. */
int gwm_halo_exchange_layout(gwm_ctx* ctx, int n_ranks, const int* desc13, int periodic_x,
                             int periodic_y, int orientation, void* const* device_fields);
/* Device allocation helpers for callers that hold raw per-rank arrays. */
int gwm_device_alloc(gwm_ctx* ctx, size_t bytes, void** out);
int gwm_device_free(gwm_ctx* ctx, void* ptr);
int gwm_memcpy_h2d(gwm_ctx* ctx, void* dst, const void* src, size_t bytes);
int gwm_memcpy_d2h(gwm_ctx* ctx, void* dst, const void* src, size_t bytes);

/* ---- operators ------------------------------------------------------------ */
/* compute_dry_pressure_fields (src/dycore/dry_pressure_gradient.cu:177-202)
 * on the halo-filled state -> GWM_PRESSURE. */
int gwm_pressure(gwm_state* state);
/* compute_slow_tendencies (src/dycore/dry_core.cu:400-539) -> GWM_TEND_*. */
int gwm_slow_tendencies(gwm_state* state, const gwm_step_config* cfg);
/* apply_local_split_explicit_fast_modes (src/dycore/dry_fast_modes.cu:400-656). */
int gwm_fast_modes(gwm_state* state, const gwm_step_config* cfg);
/* advance_dry_state_ssprk3 (src/dycore/dry_core.cu:541-651). */
int gwm_ssprk3_step(gwm_state* state, const gwm_step_config* cfg,
                    const gwm_boundary_config* boundary);
/* advance_dry_state_ssprk3 with a host FastModeIntegrator plug-in
 * (include/gwm/dycore/dry_core.hpp:73-80): fast_fn runs on each stage state
 * instead of the device fast modes (a callback that does nothing reproduces
 * the prepared driver's NullFastModeIntegrator). */
int gwm_ssprk3_step_custom_fast(gwm_state* state, const gwm_step_config* cfg,
                                const gwm_boundary_config* boundary, gwm_boundary_fn fast_fn,
                                void* fast_user);
/* advance_passive_tracers_ssprk3 (src/dycore/dry_transport.cu:440-518). */
int gwm_tracers_ssprk3(gwm_state* state, const gwm_step_config* cfg, int boundary_mode);
/* advance_passive_tracers_ssprk3 with a host TracerBoundaryUpdater plug-in
 * (include/gwm/dycore/passive_tracer.hpp:13-19): fn runs on a stage view
 * whose tracer fields are the stage input of every tracer, at the
 * step-relative times 0, dt and dt/2 (dry_transport.cu:485-508), before the
 * strips of boundary_mode (GWM_BOUNDARY_NONE or GWM_BOUNDARY_STRIPS).  The
 * tracers are independent, so one call per stage for all tracers stands for
 * the reference's one call per stage and tracer. */
int gwm_tracers_ssprk3_callback(gwm_state* state, const gwm_step_config* cfg,
                                int boundary_mode, gwm_boundary_fn fn, void* user);
/* apply_warm_rain_microphysics, both overloads (include/gwm/physics/warm_rain.hpp:57-68):
 * metric_fallout = 0 -> nominal fallout (warm_rain.cu:293-350),
 * metric_fallout = 1 -> layout/metric overload (warm_rain.cu:352-432). */
int gwm_warm_rain(gwm_state* state, const gwm_warm_rain_config* cfg, int metric_fallout);
/* apply_reference_boundaries (boundary_conditions.cpp:111-207) / relaxation. */
int gwm_apply_boundary(gwm_state* state, int mode, double relax_width);
/* One step of gwm_prepared_case_driver's loop (gwm_prepared_case_driver.cpp:427-463):
 * dry SSPRK3, tracer SSPRK3, warm rain (metric fallout). */
int gwm_moist_step(gwm_state* state, const gwm_step_config* cfg,
                   const gwm_warm_rain_config* wr, const gwm_boundary_config* boundary);
/* Diagnostics with compensated FP64 device sums. */
int gwm_budget_sums(gwm_state* state, double* out);
/* Plan-view products of the state at level slice_k (clamped to [0, nz-1]):
 * extract_dry_plan_view / extract_runtime_plan_view
 * (include/gwm/io/plan_view_output.hpp:40-52, src/io/plan_view_output.cpp:601-839),
 * computed on the device in one column pass.  out: double[GWM_PLAN_VIEW_FIELDS][ny][nx]
 * (row-major y, x), fields in the reference's order: terrain_height, z_center,
 * rho_d, theta_m, air_pressure, air_temperature, u_velocity, v_velocity,
 * wind_speed, w_velocity; with tracers specific_humidity, relative_humidity,
 * dewpoint; with cloud and rain water cloud_water_mixing_ratio,
 * rain_water_mixing_ratio, total_condensate, column_cloud_water,
 * column_rain_water, column_total_condensate, column_rain_fraction,
 * synthetic_reflectivity; with with_precip accumulated_surface_precipitation
 * and mean_surface_precipitation_rate (over steps * dt).  *nfields: 10, 13,
 * 21 or 23; only those planes are written. */
#define GWM_PLAN_VIEW_FIELDS 23
int gwm_plan_view(gwm_state* state, int slice_k, int steps, double dt, int with_precip,
                  double* out, int* nfields);
/* gwm_budget_sums without any allocation, for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  The reduction partials live
 * in a static device array of 4096 doubles (32 KiB per GPU, held until the
 * process ends; the budget needs 512 x 5 of them for any grid size), host
 * staging is on the stack, and the precipitation bucket is reduced on the
 * device.  Same slots as gwm_budget_sums: counts, extremes and the
 * non-finite count agree exactly, sums to rounding (the bucket is summed in
 * another order).  Calls that overlap on one GPU share the array and race;
 * each call synchronises the state's stream before it returns. */
int gwm_budget_sums_noalloc(gwm_state* state, double* out);

/* ---- initial states (host-built, uploaded) ---------------------------------
 * make_constant_dry_state (dry_core.cu:348-361), make_hydrostatic_rest_state
 * (dry_core.cu:363-398), make_warm_bubble_state / make_density_current_state
 * (idealized_cases.cpp:105-180), make_mountain_wave_background_state
 * (idealized_cases.cpp:182-223).  Bubble parameters p[12] = rho_background
 * (<= 0: hydrostatic background with rho_surface = -p[0]), theta_background,
 * u, v, w, center x/y/z fractions, radius x/y/z fractions, theta perturbation. */
int gwm_init_constant(gwm_state* state, double rho, double theta, double u, double v, double w);
int gwm_init_hydrostatic_rest(gwm_state* state, double rho_surface, double theta_ref);
int gwm_init_bubble(gwm_state* state, const double* p12, int sign);
int gwm_init_mountain_wave(gwm_state* state, double rho_surface, double theta_ref, double u,
                           double v, double w);
/* Extensions for the benchmark cases: a constant-N sounding
 * theta = theta_s exp(N^2 z / g) over terrain with uniform wind, and the
 * warm-rain convection sounding (theta = 300 + 43 (z/12km)^1.25 below 12 km,
 * isothermal above; RH = 1 - 0.75 (z/12km)^1.25, qv <= 14 g/kg) with seeded
 * uniform theta noise of +/- amplitude below noise_top_m. */
int gwm_init_constant_n(gwm_state* state, double theta_surface, double brunt_vaisala,
                        double p_surface, double u, double v);
int gwm_init_convection(gwm_state* state, double p_surface, unsigned long long seed,
                        double noise_amplitude, double noise_top_m);

/* ---- warm-rain surrogate --------------------------------------------------
 * No counterpart in the reference: an extension that replaces the source
 * terms of apply_warm_rain_microphysics (warm_rain.cu:36-120) by a per-cell
 * MLP, 7 features (T, p, rho_d, qv, qc, qr, dt) -> 64 -> 64 -> 64 -> 4
 * outputs (dqv, dqc, dqr, dtheta), GELU (tanh form).  Parameters are FP32,
 * flat, W[in][out]: W1 b1 W2 b2 W3 b3 W4 b4 (GWM_SURROGATE_PARAMS values).
 * Training runs in FP32 on the device; inference on the matrix cores with
 * BF16 weights and activations and FP32 accumulation.  Cells whose raw
 * features leave the box [box_lo, box_hi] use the plain scheme.  All
 * surrogate calls need an FP32 context. */
enum { GWM_SURROGATE_IN = 7, GWM_SURROGATE_OUT = 4, GWM_SURROGATE_PARAMS = 9092 };
typedef struct gwm_surrogate gwm_surrogate;
typedef struct {
  float x_mean[7], x_std[7]; /* feature normalisation */
  float y_mean[4], y_std[4]; /* target normalisation  */
  float box_lo[7], box_hi[7];/* fallback box, raw feature units */
} gwm_surrogate_norm;
/* Seeded Xavier-uniform weights, zero biases. */
int gwm_surrogate_create(gwm_ctx* ctx, unsigned long long seed, gwm_surrogate** out);
int gwm_surrogate_destroy(gwm_surrogate* m);
int gwm_surrogate_set_norm(gwm_surrogate* m, const gwm_surrogate_norm* n);
int gwm_surrogate_get_norm(gwm_surrogate* m, gwm_surrogate_norm* n);
/* Host copies of the FP32 parameters (set also resets the optimiser). */
int gwm_surrogate_set_params(gwm_surrogate* m, const float* host);
int gwm_surrogate_get_params(gwm_surrogate* m, float* host);
/* n training pairs from the state: cells drawn by a counter hash of (seed,
 * sample); dev_x [n][7] raw features before the adjustment, dev_y [n][4]
 * the plain scheme's change over one call (device FP32 arrays); dev_cell
 * (optional, int64 [n]) the cell index (k*ny + j)*nx + i. */
int gwm_surrogate_sample(gwm_state* state, const gwm_warm_rain_config* cfg,
                         unsigned long long seed, long n, void* dev_x, void* dev_y,
                         void* dev_cell);
/* Adam on the mean squared error of the normalised outputs, minibatches of
 * `batch` consecutive rows, learning rate cosine-decayed from lr to
 * lr_final; epoch_loss[e] = mean squared normalised error of epoch e. */
int gwm_surrogate_train(gwm_surrogate* m, const void* dev_x, const void* dev_y, long n, int batch,
                        int epochs, double lr, double lr_final, double* epoch_loss);
/* Loss and gradient (FP32 kernels, no update) over n rows. */
int gwm_surrogate_loss_grad(gwm_surrogate* m, const void* dev_x, const void* dev_y, int n,
                            double* loss, float* grad_host);
/* BF16 matrix-core inference: raw features [n][7] -> raw outputs [n][4]. */
int gwm_surrogate_infer(gwm_surrogate* m, const void* dev_x, long n, void* dev_y);
/* gwm_warm_rain with the surrogate source terms (fallout unchanged);
 * n_fallback (optional) receives the number of cells outside the box. */
int gwm_warm_rain_surrogate(gwm_state* state, const gwm_warm_rain_config* cfg, gwm_surrogate* m,
                            int metric_fallout, long long* n_fallback);
/* Cells of the last surrogate warm-rain call on this network's context
 * (synchronises): taken by the plain scheme outside the training box, and
 * evaluated by the network (cells without cloud or rain at or below
 * saturation take neither: every source term is zero there).  No
 * counterpart in the reference. */
int gwm_surrogate_last_counts(gwm_surrogate* m, long long* fallback, long long* network);
/* gwm_moist_step with the surrogate warm rain. */
int gwm_moist_step_surrogate(gwm_state* state, const gwm_step_config* cfg,
                             const gwm_warm_rain_config* wr, gwm_surrogate* m,
                             const gwm_boundary_config* boundary);

/* ---- graph-replayed steps and ensembles -----------------------------------
 * No counterpart in the reference (its stepper launches every operator from
 * the host): the moist step (gwm_moist_step without boundaries, or
 * gwm_moist_step_surrogate when `surrogate` is not NULL) is captured once
 * into a HIP graph of two steps -- the buffer rotation has period two --
 * after one eager step, and replayed.  Results equal the eager steps'.  A
 * change of configuration or of the surrogate's normalisation re-captures. */
int gwm_moist_steps_graph(gwm_state* state, const gwm_step_config* cfg,
                          const gwm_warm_rain_config* wr, gwm_surrogate* surrogate, int nsteps);
/* Ensemble of states of any sizes: every member advances nsteps moist
 * steps, as one launch per kernel of the step over all members' work-groups
 * (each member's two steps recorded, the batched sequence captured into a
 * HIP graph and replayed); members of one precision on one device, which
 * may share a context.  Where the steps cannot be batched (mixed precision,
 * boundaries, a configuration whose kernels differ between members), the
 * members run as gwm_ensemble_steps_streams does.  Results equal each
 * member's own gwm_moist_step.  Returns after all members finished.
 * surrogates: NULL, or one per member (entries may be NULL). */
int gwm_ensemble_steps(gwm_state* const* states, int n, const gwm_step_config* cfg,
                       const gwm_warm_rain_config* wr, gwm_surrogate* const* surrogates,
                       int nsteps);
/* The batched ensemble's autotuner: per ensemble signature (member shapes
 * and forms), the time of a two-step pair for each launch-code candidate
 * (the members' own tuned codes; then for all members 2, 4, 8 level chunks
 * with 8-row tiles), or "untried"; the fastest is used once all are timed.
 * One line per signature. */
int gwm_ensemble_tuning_report(char* buf, size_t size);
/* The same with each member on its own context (stream): members run
 * concurrently, their two-step graph replays interleaved. */
int gwm_ensemble_steps_streams(gwm_state* const* states, int n, const gwm_step_config* cfg,
                               const gwm_warm_rain_config* wr,
                               gwm_surrogate* const* surrogates, int nsteps);

#ifdef __cplusplus
}
#endif

#endif
