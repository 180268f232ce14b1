/*
 * gwm_oracle.h -- C API of the CPU oracle for the dry + warm-rain time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP product
 * path is compared against (and the timed CPU baseline).  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product library never links, loads or calls it.
 *
 * The oracle restates the reference's algorithm on the CPU, templated on the
 * working precision (float or double), including the virtual-rank layout and
 * its halo-exchange semantics.  Every function in gwm_oracle.cpp cites the
 * reference file:line it follows.
 */
#ifndef GWM_ORACLE_H
#define GWM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Field identifiers (same numbering as the product C-ABI). */
enum {
  ORC_RHO = 0,
  ORC_RHO_THETA = 1,
  ORC_MOM_U = 2,
  ORC_MOM_V = 3,
  ORC_MOM_W = 4,
  ORC_PRESSURE = 5,
  ORC_TEND_RHO = 6,
  ORC_TEND_RHO_THETA = 7,
  ORC_TEND_U = 8,
  ORC_TEND_V = 9,
  ORC_TEND_W = 10,
  ORC_PRECIP = 11,
  ORC_TRACER0 = 16
};

/* Metric identifiers for orc_get_metric (global arrays, no halo). */
enum {
  ORC_METRIC_Z_INTERFACES = 0, /* nx*ny*(nz+1) */
  ORC_METRIC_Z_CENTERS = 1,    /* nx*ny*nz */
  ORC_METRIC_INV_DZ_CELL = 2,  /* nx*ny*nz */
  ORC_METRIC_INV_DZ_FACE = 3,  /* nx*ny*(nz+1) */
  ORC_METRIC_TERRAIN = 4,      /* nx*ny */
  ORC_METRIC_SLOPE_X = 5,      /* nx*ny */
  ORC_METRIC_SLOPE_Y = 6,      /* nx*ny */
  ORC_METRIC_FLAT_Z_INTERFACES = 7, /* nz+1 (double) */
  ORC_METRIC_FLAT_Z_CENTERS = 8,    /* nz (double) */
  ORC_METRIC_WEIGHT_INTERFACES = 9, /* nz+1 (double) */
  ORC_METRIC_WEIGHT_CENTERS = 10    /* nz (double) */
};

typedef struct {
  double dt;
  int fast_substeps;
  double gravity;
  int fast_update_horizontal_momentum;
  int fast_update_density;
} orc_step_cfg;

typedef struct {
  double dt;
  double condensation_relaxation;
  double cloud_autoconversion_threshold;
  double cloud_autoconversion_rate;
  double rain_evaporation_rate;
  double rain_terminal_velocity;
  double sedimentation_layer_depth_m;
  int enable_latent_heating;
  double accretion_rate; /* extension; 0 reproduces the reference */
} orc_warm_rain_cfg;

/* Boundary modes for the stepper. */
enum { ORC_BOUNDARY_NONE = 0, ORC_BOUNDARY_STRIPS = 1, ORC_BOUNDARY_RELAX = 2 };

/* Summary slots written by orc_summary (double[ORC_SUMMARY_N]). */
enum {
  ORC_S_TOTAL_DRY_MASS = 0,
  ORC_S_TOTAL_RHO_THETA = 1,
  ORC_S_TOTAL_MOM_X = 2,
  ORC_S_TOTAL_MOM_Y = 3,
  ORC_S_TOTAL_MOM_Z = 4,
  ORC_S_RHO_MIN = 5,
  ORC_S_RHO_MAX = 6,
  ORC_S_THETA_MIN = 7,
  ORC_S_THETA_MAX = 8,
  ORC_S_U_MIN = 9,
  ORC_S_U_MAX = 10,
  ORC_S_V_MIN = 11,
  ORC_S_V_MAX = 12,
  ORC_S_W_MIN = 13,
  ORC_S_W_MAX = 14,
  ORC_S_VAPOR_PATH = 15,
  ORC_S_CLOUD_PATH = 16,
  ORC_S_RAIN_PATH = 17,
  ORC_S_PRECIP_SUM = 18,
  ORC_S_PRECIP_MEAN = 19,
  ORC_S_PRECIP_MAX = 20,
  ORC_S_PRECIP_CELLS = 21,
  ORC_S_PRECIPITATING_CELLS = 22,
  ORC_S_TRACER_QMIN0 = 23, /* 3 tracers: min at 23,25,27 ; max at 24,26,28 */
  ORC_SUMMARY_N = 32
};

/* Create a world: the global grid, its virtual-rank layout and the state.
 * precision: 4 (float) or 8 (double).
 * z_interfaces: NULL for the reference's uniform eta levels (z_top/nz), or
 *   nz+1 monotone flat interface heights (stretched levels, extension).
 * terrain: NULL for flat, else nx*ny floats (row-major, x fastest). */
void* orc_world_create(int precision, int nx, int ny, int nz, int halo,
                       int ranks_x, int ranks_y, int periodic_x,
                       int periodic_y, double dx, double dy, double z_top,
                       const double* z_interfaces, const float* terrain,
                       double terrain_taper_eta, int n_tracers);
void orc_world_destroy(void* world);
const char* orc_last_error(void);

int orc_num_ranks(void* world);
/* desc[13] = rank, coord_x, coord_y, ranks_x, ranks_y, global_nx, global_ny,
 * nz, halo, i_begin, i_end, j_begin, j_end */
int orc_rank_desc(void* world, int rank, int* desc);
int orc_neighbors(void* world, int rank, int* wesn);

/* Storage extent of a field on one rank (including halo); returns elements. */
long orc_rank_storage_size(void* world, int rank, int field, int* dims4);
int orc_get_rank_storage(void* world, int rank, int field, void* out);
int orc_set_rank_storage(void* world, int rank, int field, const void* in);

/* Owned values gathered / scattered as global arrays (x fastest, no halo).
 * Face fields carry their extra face (nx+1, ny+1 or nz+1). */
int orc_get_global(void* world, int field, void* out);
int orc_set_global(void* world, int field, const void* in);
/* Fill every storage element (halo included) of a field with a constant. */
int orc_fill(void* world, int field, double value);

int orc_get_metric(void* world, int which, void* out);

/* Reference halo semantics on one field of every rank. */
int orc_exchange(void* world, int field);
int orc_sync_interfaces(void* world, int field);

/* Operators. */
int orc_pressure(void* world);
int orc_slow_tendencies(void* world, const orc_step_cfg* cfg);
int orc_momentum_flux_only(void* world);
int orc_pressure_gradient_only(void* world);
int orc_fast_modes(void* world, const orc_step_cfg* cfg);
int orc_set_boundary_from_current(void* world);
/* Bit t set: tracer t takes the strip boundary (default all). */
int orc_set_tracer_strip_mask(void* world, unsigned mask);
/* apply_reference_boundaries (strips) or the relaxation extension, now. */
int orc_apply_boundary(void* world, int mode, double width);
/* Scalar thermodynamics: fn 0 theta(rho, rho_theta), 1 pressure(rho,
 * rho_theta), 2 rho(p, theta), 3 e_sat(T), 4 q_sat(T, p), 5 T(rho, rho_theta),
 * 6 hydrostatic rho(rho_s, theta, z), 7 c_p, 8 epsilon. */
double orc_thermo(int precision, int fn, double a, double b, double c);
int orc_ssprk3_step(void* world, const orc_step_cfg* cfg, int boundary_mode,
                    double relax_width, double relax_tau);
int orc_tracers_ssprk3(void* world, const orc_step_cfg* cfg, int boundary_mode);
int orc_warm_rain(void* world, const orc_warm_rain_cfg* cfg, int metric_fallout);
int orc_summary(void* world, double* out);

/* Initial states (reference cases + extensions). params: see gwm_oracle.cpp */
int orc_init_constant(void* world, double rho, double theta, double u,
                      double v, double w);
int orc_init_hydrostatic_rest(void* world, double rho_surface,
                              double theta_ref);
int orc_init_bubble(void* world, const double* p12, int sign);
int orc_init_mountain_wave(void* world, double rho_surface, double theta_ref,
                           double u, double v, double w);

/* K11 column hydrostatic reference of one column (dry_fast_modes.cu:108-158)
 * in the given precision: th, zc (nz) in; pref, rref (nz) out. */
int orc_column_reference(int precision, int nz, const double* th, const double* zc,
                         double g, double p_bottom, double* pref, double* rref);

/* Timing helper for the CPU baseline: runs n steps of the moist or dry step
 * and returns wall seconds. */
/* extract_dry_plan_view / extract_runtime_plan_view: [field][j][i] doubles,
 * returns the field count (10, 13, 21 or 23), -1 on error. */
int orc_plan_view(void* w, int slice_k, int steps, double dt, int with_precip, double* out);
double orc_time_steps(void* world, const orc_step_cfg* cfg,
                      const orc_warm_rain_cfg* wr, int moist, int steps,
                      int threads);

#ifdef __cplusplus
}
#endif

#endif
