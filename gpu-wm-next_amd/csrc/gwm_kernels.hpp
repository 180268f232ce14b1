// gwm_kernels.hpp -- kernel argument blocks and launcher declarations.
#pragma once

#include <cstring>
#include <vector>

#include "gwm_internal.hpp"

namespace gwm {

template <class R>
struct SlowArgs {
  Geo g;
  DState<R> s;    // stage input state (rho, rho_theta, faces)
  DState<R> q0;   // SSP base (mode 2)
  DState<R> out;  // output (tendencies in mode 0)
  DMetrics<R> m;
  const R* theta_ref;  // nz
  R gravity;
  R dt;
  R wa, wb;
  int mode;       // 0 tendencies, 1 update, 2 update + blend
  int out_theta;  // write theta (theta' in the perturbation form) instead of rho*theta
  Base<R> base;   // perturbation form (base.rho != nullptr)
};

template <class R>
struct FastArgs {
  Geo g;
  const R* rho;    // current rho
  const R* theta;  // fixed theta of the subcycle
  const R *u, *v, *w;     // face departures m - m_base (read unless first)
  const R *ub, *vb, *wb;  // faces at subcycle entry (read on the last substep)
  int first, last;        // first / last substep of the subcycle
  int down;               // flat whole-column launches: march the columns top down
  R *rho_out, *u_out, *v_out, *w_out;  // face departures, full faces if last
  R* rt_out;  // rho*theta rebuild (last substep), may be null
  const R *rref_lev, *pref_lev;  // flat reference (nz)
  DMetrics<R> m;
  R dtf, gravity;
  int horizontal, density, col_ref;
  Base<R> base;  // perturbation form: rho, rho_out, rt_out are departures, theta is theta'
  int part_theta;  // partials (if emitted) are the level sums of theta of the written
                   // (rho, rho*theta), for theta_ref, instead of rho and p
};

template <class R>
struct TracerArgs {
  Geo g;
  const R* rho;
  const R *u, *v, *w;
  const R* idz_c;
  const R* lev_idz_c;
  const R* in[kMaxTracers];
  const R* q0[kMaxTracers];
  R* out[kMaxTracers];
  int ntr;
  R dt, wa, wb;
  int mode;
  const R* base_rho;  // perturbation form: rho is rho' against base_rho (nz)
};

template <class R>
struct WarmRainArgs {
  Geo g;
  const R* rho;
  R* rt;
  R *qv, *qc, *qr;
  R* precip;
  const R* idz_c;
  const R* lev_idz_c;
  R dt, relax, thr, krate, erate, vt, depth, kacc;
  int heat;
  int fallout;  // 0 none, 1 nominal, 2 metric
  Base<R> base;  // perturbation form
};

// Per-block partials of k_reduce: sum, compensation, min, max and the count
// of non-finite values.
constexpr int kReduceParts = 5;

template <class R>
struct ReduceArgs {
  const R* f;
  const R* rho;
  const R* idz3;
  const R* idz_lev;
  int fx, fy, fz, h;
  int kind;  // 0 value, 1 value*dz, 2 value/rho, 3 count of values > 0
  double* partials;
  // Perturbation form: per-level bases added for min / max (and for rho in
  // kind 2); sums stay over the stored departures.  On terrain (cb.drho set)
  // the cell's base is the level value plus its deviation; rho*theta's is
  // rho_b theta_b of the cell's.
  const R* fbase;
  const R* rbase;
  Base<R> cb;
};

template <class R>
void launch_level_partials(hipStream_t st, const Geo& g, const R* rho, const R* aux, int mode,
                           double* partials, int* nchunk_out, Base<R> base = Base<R>{});
// Partials per level written by launch_level_partials for ny rows.
int level_partials_chunks(int ny);
template <class R>
void launch_level_finalize(hipStream_t st, int nz, int nchunk, const double* partials, int mode,
                           double count, R* o0, R* o1);
template <class R>
void launch_pressure(hipStream_t st, const Geo& g, const R* rho, const R* rt, R* p);
// The step kernels' FP32 helpers elementwise (fn: GWM_MATH_* of gwm_hip.h).
void launch_fp32_math(hipStream_t st, int fn, const float* x, float* y, size_t n);
// Column-marching slow stage (gwm_slow.hip): compute_slow_tendencies fused
// with the stage update / SSP blend; mode 0 writes tendencies, 1 updates,
// 2 updates and blends (SlowArgs).
template <class R>
void launch_slow_column(hipStream_t st, const SlowArgs<R>& a, bool terrain, int nkc);
// Column-marching fast substep (gwm_fast.hip).  With part_out (flat refs),
// also writes the next substep's level partials, one per 64-column row
// segment, [k][row][tile column][2], and returns their count per level in
// *nparts.  code: level chunks per column in the low 4 bits (1 = whole
// columns; the column-reference path keeps 1) and, in bit 4, a tile height
// of 4 rows instead of 8 (flat launches without the column reference).
template <class R>
void launch_fast_column(hipStream_t st, const FastArgs<R>& a, bool terrain, int code,
                        double* part_out, int* nparts);
int fast_column_parts(int nx, int ny);
template <class R>
void launch_theta(hipStream_t st, const Geo& g, const R* rho, const R* rt, R* th,
                  Base<R> base = Base<R>{});
// Column-marching tracer stage (gwm_slow.hip): one SSPRK3 stage of every
// tracer with frozen rho and face momenta (mode 1 update, 2 update + blend).
template <class R>
void launch_tracer_column(hipStream_t st, const TracerArgs<R>& a, bool terrain, int nkc);
template <class R>
void launch_rebuild_rt(hipStream_t st, const Geo& g, const R* rho, const R* th, R* rt,
                       Base<R> base = Base<R>{});
template <class R>
void launch_warm_rain(hipStream_t st, const WarmRainArgs<R>& a, bool terrain);
template <class R>
void launch_halo_fill(hipStream_t st, const Geo& g, int orient, R* f);
template <class R>
void launch_sync_interfaces(hipStream_t st, const Geo& g, int orient, R* f);
template <class R>
void launch_exchange_layout(hipStream_t st, int phase, int orient, int nranks, const int* d,
                            R* const* F, int max_elems);
template <class R>
void launch_boundary(hipStream_t st, const Geo& g, int orient, R* f, const R* ref, int mode,
                     double width);
template <class R>
void launch_reduce(hipStream_t st, const ReduceArgs<R>& a, int nblocks);

// Plan-view products (gwm-next-plan-view/v1, src/io/plan_view_output.cpp),
// in the reference's field order: 10 dry fields, 3 with specific humidity,
// 8 more with cloud and rain water, 2 more with the surface bucket.
enum PlanViewField {
  kPvTerrain = 0, kPvZCenter, kPvRho, kPvTheta, kPvPressure, kPvTemperature, kPvU, kPvV,
  kPvWind, kPvW,
  kPvQ, kPvRh, kPvDewpoint,
  kPvQc, kPvQr, kPvCondensate, kPvColumnCloud, kPvColumnRain, kPvColumnCondensate,
  kPvRainFraction, kPvReflectivity,
  kPvPrecip, kPvPrecipRate,
  kPvCount
};

template <class R>
struct PlanViewArgs {
  Geo g;
  const R *rho, *rt, *u, *v, *w;  // state (perturbation form: departures)
  const R *qv, *qc, *qr;          // tracer masses; null when absent
  const R* precip;                // surface bucket (ny, nx); null when absent
  const R* idz_c;                 // inverse cell depth, metric index
  const R* lev_idz_c;             // the same per level (flat terrain)
  const R* zc;                    // cell-centre heights, metric index
  const R* terrain;               // terrain height (ny, nx)
  Base<R> base;
  int k;                  // slice level
  double elapsed_hours;   // max(steps * dt / 3600, 1e-9)
  double* out;            // [kPvCount][ny][nx]
};

template <class R>
void launch_plan_view(hipStream_t st, const PlanViewArgs<R>& a, bool terrain);

// Allocation-free budget sums: k_reduce block partials live in a static
// device array of kNoAllocScratchDoubles doubles (32 KiB), one per GPU.
constexpr int kNoAllocScratchDoubles = 4096;
constexpr int kNoAllocMaxDevices = 64;
// Resolves (once per device, with `device` current) and returns the device
// address of that array on `device`.
hipError_t noalloc_scratch_setup(int device, double** out);
template <class R>
void launch_metrics(hipStream_t st, const Geo& g, const double* zif, const double* zc,
                    const double* wif, const double* wc, const float* terrain, R* zif_m, R* zc_m,
                    R* idz_c, R* idz_f, R* sx, R* sy, R* terr);
template <class R>
void launch_copy(hipStream_t st, const R* src, R* dst, size_t n);
template <class R>
void launch_fill(hipStream_t st, R* dst, size_t n, R v);

// ---------------------------------------------------------------------------
// Batched launches: the step of several states (ensemble members of any
// sizes) as one launch per kernel of the step.  While a Recorder is set on
// the calling host thread, the step kernels' launchers launch nothing: each
// appends its launch (the batched form of its kernel, block shape, dynamic
// LDS, the member's grid and argument item) instead.  Recording one member's
// steps gives the same sequence of kernels as any other member's with the
// same configuration; run_batched launches each position of the sequence
// once over all members' work-groups.
// ---------------------------------------------------------------------------
struct LaunchRecord {
  const void* fn = nullptr;  // batched kernel: (const Item* items, BatchMap map)
  unsigned bx = 1, by = 1, bz = 1;
  size_t lds = 0;
  int gx = 1, gy = 1;        // the member's grid
  std::vector<unsigned char> item;
};
struct Recorder {
  std::vector<LaunchRecord> ops;
  const char* unsupported = nullptr;  // a launcher without a batched form was called
};
// The calling thread's recorder (null: launchers launch).
Recorder*& recorder();
template <class Item>
inline void record_launch(const void* fn, dim3 block, size_t lds, Item it, int gx, int gy) {
  it.gx = gx;
  it.gy = gy;
  LaunchRecord r;
  r.fn = fn;
  r.bx = block.x;
  r.by = block.y;
  r.bz = block.z;
  r.lds = lds;
  r.gx = gx;
  r.gy = gy;
  r.item.resize(sizeof(Item));
  std::memcpy(r.item.data(), &it, sizeof(Item));
  recorder()->ops.push_back(std::move(r));
}
// Marks the recording unusable (called by launchers with no batched form).
inline bool refuse_record(const char* what) {
  Recorder* r = recorder();
  if (!r) return false;
  if (!r->unsupported) r->unsupported = what;
  return true;
}

template <class R>
struct SlowItem {
  SlowArgs<R> a;
  int tiles_x, tiles, nkc;
  int gx, gy;
};
template <class R>
struct FastItem {
  FastArgs<R> a;
  int tiles_x, tiles, nkc;
  double* part_out;
  int gx, gy;
};
template <class R>
struct TracerItem {
  TracerArgs<R> a;
  int tiles_x, tiles, nkc;
  int gx, gy;
};
template <class R>
struct WarmRainItem {
  WarmRainArgs<R> a;
  int gx, gy;
};
template <class R>
struct PartialsItem {
  Geo g;
  const R* rho;
  const R* aux;
  int mode;
  double* partials;
  Base<R> base;
  int gx, gy;
};
template <class R>
struct FinalizeItem {
  int nz, nchunk;
  const double* partials;
  int mode;
  double count;
  R* out0;
  R* out1;
  int gx, gy;
};

}  // namespace gwm
