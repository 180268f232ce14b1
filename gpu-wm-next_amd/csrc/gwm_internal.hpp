// gwm_internal.hpp -- shared device/host definitions of the gfx950 path.
//
// Storage layout (identical to the reference's Field3D, field3d.hpp:49-54):
//   cell   (nx  , ny  , nz  ) idx = (k*(ny+2h)   + j+h)*(nx+2h)   + i+h
//   x-face (nx+1, ny  , nz  ) idx = (k*(ny+2h)   + j+h)*(nx+1+2h) + i+h
//   y-face (nx  , ny+1, nz  ) idx = (k*(ny+1+2h) + j+h)*(nx+2h)   + i+h
//   z-face (nx  , ny  , nz+1) idx as cell, nz+1 levels
// Kernels never rely on halo storage: a neighbour outside the owned range is
// read through the 1x1 halo map (periodic wrap, or zero across an open side),
// which is the value the reference's exchange would have placed there
// (src/comm/halo_exchange.cpp:32-116, 313-373).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <type_traits>

namespace gwm {

struct Geo {
  int nx, ny, nz, h;
  int per_x, per_y;
  int pxc, pyc;   // cell pitches nx+2h, ny+2h
  int pxu;        // x-face row pitch nx+1+2h
  int pyv;        // y-face plane pitch ny+1+2h
  double dx, dy;

  // Element offsets are formed in 32-bit arithmetic (grids are limited to
  // 2^31 elements per field, checked at grid creation), which keeps address
  // arithmetic in single registers.
  __host__ __device__ size_t cell(int i, int j, int k) const {
    return static_cast<unsigned>((k * pyc + (j + h)) * pxc + (i + h));
  }
  __host__ __device__ size_t xf(int i, int j, int k) const {
    return static_cast<unsigned>((k * pyc + (j + h)) * pxu + (i + h));
  }
  __host__ __device__ size_t yf(int i, int j, int k) const {
    return static_cast<unsigned>((k * pyv + (j + h)) * pxc + (i + h));
  }
  __host__ __device__ size_t zf(int i, int j, int k) const { return cell(i, j, k); }
  __host__ __device__ size_t metric(int i, int j, int k) const {
    return static_cast<unsigned>((k * ny + j) * nx + i);
  }
  __host__ __device__ size_t cell_total() const { return static_cast<size_t>(pxc) * pyc * nz; }
  __host__ __device__ size_t xf_total() const { return static_cast<size_t>(pxu) * pyc * nz; }
  __host__ __device__ size_t yf_total() const { return static_cast<size_t>(pxc) * pyv * nz; }
  __host__ __device__ size_t zf_total() const { return static_cast<size_t>(pxc) * pyc * (nz + 1); }
};

// Halo map for a cell coordinate along an axis of n owned cells: returns the
// owned source index, or -1 when the value is zero (open side).
__host__ __device__ inline int map_cell(int i, int n, int periodic) {
  if (i >= 0 && i < n) return i;
  if (!periodic) return -1;
  int m = i % n;
  return m < 0 ? m + n : m;
}
// Along the face-normal axis of a face field (n+1 owned faces, faces 0 and n
// coincide under periodicity): ghost -1 <- face n-1, ghost n+1 <- face 1.
// Indices further out (work-items past the end of a partial tile) wrap
// modulo n, so every periodic index maps to an owned face.
__host__ __device__ inline int map_face(int i, int n, int periodic) {
  if (i >= 0 && i <= n) return i;
  if (!periodic) return -1;
  const int m = i % n;
  return m < 0 ? m + n : m;
}

// ---------------------------------------------------------------------------
// Math wrappers so one template serves FP32 and FP64.
// ---------------------------------------------------------------------------
__device__ __host__ inline float rmax(float a, float b) { return fmaxf(a, b); }
__device__ __host__ inline double rmax(double a, double b) { return fmax(a, b); }
__device__ __host__ inline float rmin(float a, float b) { return fminf(a, b); }
__device__ __host__ inline double rmin(double a, double b) { return fmin(a, b); }
__device__ inline float rpow(float a, float b) { return powf(a, b); }
__device__ inline double rpow(double a, double b) { return pow(a, b); }
__device__ inline float rexp(float a) { return expf(a); }
__device__ inline double rexp(double a) { return exp(a); }
__device__ inline float rexpm1(float a) { return expm1f(a); }
__device__ inline double rexpm1(double a) { return expm1(a); }
__device__ inline float rlog1p(float a) { return log1pf(a); }
__device__ inline double rlog1p(double a) { return log1p(a); }

// ---------------------------------------------------------------------------
// Arithmetic of the time-step kernels (q*).  FP64 keeps correctly rounded
// division and the library pow/exp/log1p/expm1, so the FP64 path matches the
// FP64 oracle to ~1e-12.  FP32 uses the hardware reciprocal, base-2
// logarithm and base-2 exponential (v_rcp_f32, v_log_f32, v_exp_f32, each
// within 1 ulp): a power costs ~16 instructions instead of ~170 and a
// division 4 instead of ~12, which is what the step's FP32 kernels are bound
// by.  Each
// result stays within a few ulp of the correctly rounded value (bounds
// measured on the device by tests/test_fp32_math.py).  Arguments of qpow and
// qlog are positive normal numbers (the callers floor them).
// ---------------------------------------------------------------------------
// Quotients take the hardware reciprocal r and one correction with the exact
// residual e = a - q b (FMA): q + e r is the correctly rounded a / b except
// in rare near-tie cases, 4 instructions instead of ~12 (normal operands; the
// callers floor the divisors).
__device__ inline float qrcp(float b) {
  const float r = __builtin_amdgcn_rcpf(b);
  return __builtin_fmaf(__builtin_fmaf(-r, b, 1.0f), r, r);
}
__device__ inline double qrcp(double b) { return 1.0 / b; }
__device__ inline float qdiv(float a, float b) {
  const float r = __builtin_amdgcn_rcpf(b);
  const float q = a * r;
  return __builtin_fmaf(__builtin_fmaf(-q, b, a), r, q);
}
__device__ inline double qdiv(double a, double b) { return a / b; }
// 2^t for t = hi + lo (|lo| << |hi|): the integer part of hi goes to the
// exponent exactly (v_ldexp_f32), the hardware exp2 sees only the fraction,
// so the rounding of a large t costs no accuracy.
__device__ inline float qexp2_split(float hi, float lo) {
  const float n = __builtin_rintf(hi);
  const float f = (hi - n) + lo;  // hi - n is exact
  return __builtin_amdgcn_ldexpf(__builtin_amdgcn_exp2f(f), static_cast<int>(n));
}
// x^y = 2^(y log2 x), x > 0 normal: x = m 2^e with m in [sqrt(1/2), sqrt(2)),
// so the hardware log2 works near 1 where its error is small in absolute
// terms; y*e is carried with its exact rounding error (FMA) into the
// fraction.  About 16 instructions, within ~2 ulp.
__device__ inline float qpow(float x, float y) {
  float m = __builtin_amdgcn_frexp_mantf(x);  // [0.5, 1)
  int e = __builtin_amdgcn_frexp_expf(x);
  if (m < 0.70710678118654752f) {
    m *= 2.0f;
    e -= 1;
  }
  const float l = __builtin_amdgcn_logf(m);  // log2 m in [-0.5, 0.5]
  const float fe = static_cast<float>(e);
  const float ye = y * fe;
  const float ye_err = __builtin_fmaf(y, fe, -ye);
  return qexp2_split(ye, __builtin_fmaf(y, l, ye_err));
}
__device__ inline double qpow(double x, double y) { return pow(x, y); }
// e^x = 2^(x log2 e) with the product carried to double length (FMA and
// the low part of log2 e), within ~1 ulp.
__device__ inline float qexp(float x) {
  constexpr float L = 1.44269504088896341f;
  constexpr float L_lo = 1.9259629891e-8f;  // log2(e) - L
  const float t = x * L;
  const float t_err = __builtin_fmaf(x, L, -t);
  return qexp2_split(t, __builtin_fmaf(x, L_lo, t_err));
}
__device__ inline double qexp(double x) { return exp(x); }
// log1p by Goldberg's correction: log(u) * x / (u - 1) with u = 1 + x
// rounded, exact when u == 1 (accurate for |x| small, where 1 + x loses bits).
__device__ inline float qlog1p(float x) {
  const float u = 1.0f + x;
  if (u == 1.0f) return x;
  return (__builtin_amdgcn_logf(u) * 0.693147180559945309f) *
         (x * __builtin_amdgcn_rcpf(u - 1.0f));
}
__device__ inline double qlog1p(double x) { return log1p(x); }
// expm1 by Kahan's correction: (u - 1) * y / log(u) with u = exp(y) rounded.
__device__ inline float qexpm1(float y) {
  const float u = qexp(y);
  if (u == 1.0f) return y;
  const float um1 = u - 1.0f;
  if (um1 == -1.0f) return -1.0f;
  return um1 * (y * __builtin_amdgcn_rcpf(__builtin_amdgcn_logf(u) * 0.693147180559945309f));
}
__device__ inline double qexpm1(double y) { return expm1(y); }
// Division by a uniform grid step d with its reciprocal rd = 1 / d (correctly
// rounded): FP32 multiplies by rd (within an ulp of the quotient, one
// instruction instead of three), FP64 divides.
__device__ inline float qdivc(float x, float d, float rd) {
  (void)d;
  return x * rd;
}
__device__ inline double qdivc(double x, double d, double rd) {
  (void)rd;
  return x / d;
}

// Physical constants in the working precision (dry_thermo.hpp:17-19,
// moist_thermo.hpp:18-21).  FP32 uses the reference's float literals.
template <class R>
struct Phys;
template <>
struct Phys<float> {
  static constexpr float Rd = 287.05f, p0 = 100000.0f, kappa = 0.2854f;
  static constexpr float Rv = 461.5f, Lv = 2.5e6f, eps = 287.05f / 461.5f;
  static constexpr float mom_floor = -1.0e30f;
};
template <>
struct Phys<double> {
  static constexpr double Rd = 287.05, p0 = 100000.0, kappa = 0.2854;
  static constexpr double Rv = 461.5, Lv = 2.5e6, eps = 287.05 / 461.5;
  static constexpr double mom_floor = -1.0e30;
};

template <class R>
__device__ inline R d_theta(R rho, R rt) {  // dry_thermo.hpp:21-24
  return rt / rmax(rho, R(1.0e-6));
}
template <class R>
__device__ inline R d_pressure(R rho, R rt) {  // dry_thermo.hpp:26-33
  const R th = d_theta(rho, rt);
  const R ratio = rmax(rho * Phys<R>::Rd * th / Phys<R>::p0, R(1.0e-6));
  return Phys<R>::p0 * rpow(ratio, R(1) / (R(1) - Phys<R>::kappa));
}
template <class R>
__device__ inline R d_rho_from_p_theta(R p, R th) {  // dry_thermo.hpp:35-40
  const R ex = rpow(rmax(p / Phys<R>::p0, R(1.0e-6)), Phys<R>::kappa);
  return p / rmax(Phys<R>::Rd * th * ex, R(1.0e-6));
}
// The same three functions in the step kernels' arithmetic (q* above); the
// operator gwm_pressure keeps the library forms.
template <class R>
__device__ inline R q_theta(R rho, R rt) {
  return qdiv(rt, rmax(rho, R(1.0e-6)));
}
template <class R>
__device__ inline R q_pressure(R rho, R rt) {
  const R th = q_theta(rho, rt);
  const R ratio = rmax(qdiv(rho * Phys<R>::Rd * th, Phys<R>::p0), R(1.0e-6));
  return Phys<R>::p0 * qpow(ratio, R(1) / (R(1) - Phys<R>::kappa));
}
template <class R>
__device__ inline R q_rho_from_p_theta(R p, R th) {
  const R ex = qpow(rmax(qdiv(p, Phys<R>::p0), R(1.0e-6)), Phys<R>::kappa);
  return qdiv(p, rmax(Phys<R>::Rd * th * ex, R(1.0e-6)));
}
template <class R>
__device__ inline R q_qsat(R T, R p) {
  const R tc = T - R(273.15);
  const R es = R(6.112) * qexp(qdiv(R(17.67) * tc, rmax(tc + R(243.5), R(1.0e-6)))) * R(100);
  const R denom = rmax(p - (R(1) - Phys<R>::eps) * es, R(1000));
  return rmin(rmax(qdiv(Phys<R>::eps * es, denom), R(0)), R(0.05));
}
template <class R>
__device__ inline R d_cp() {
  return Phys<R>::Rd / Phys<R>::kappa;
}
template <class R>
__device__ inline R d_qsat(R T, R p) {  // moist_thermo.hpp:23-36
  const R tc = T - R(273.15);
  const R es = R(6.112) * rexp((R(17.67) * tc) / rmax(tc + R(243.5), R(1.0e-6))) * R(100);
  const R denom = rmax(p - (R(1) - Phys<R>::eps) * es, R(1000));
  return rmin(rmax(Phys<R>::eps * es / denom, R(0)), R(0.05));
}

// ---------------------------------------------------------------------------
// Perturbation form (flat, doubly periodic grids; terrain below).  The state holds
// rho' = rho - rho_b(k) and (rho theta)' = rho theta - (rho theta)_b(k)
// against a per-level base (the FP64 level means of the initial state,
// rounded to working precision here); rho == nullptr means the full form.
//   theta' = theta - theta_b = ((rho theta)' - theta_b rho') / rho,
//   p'     = p - p_b = p_b expm1(gamma log1p((rho theta)' / (rho theta)_b)),
// gamma = 1 / (1 - kappa), which is the EOS of dry_thermo.hpp:26-33 written
// for the departure, so that pressure differences keep their relative
// precision.
// ---------------------------------------------------------------------------
//
// Terrain form: the base varies within a level (terrain-following levels
// cross the base's vertical structure), so it is the per-level value above
// plus a per-cell deviation (drho, dth, dp in the cell layout); every
// departure relation above holds cell by cell with
//   rho_b = rho[k] + drho, theta_b = th[k] + dth, p_b = p[k] + dp,
//   (rho theta)_b = rho_b theta_b,
// and the base's own pressure-gradient and gravity forcing, fixed for the
// run, is held per face in FP32 (fu, fv: the terrain-following horizontal
// pressure gradient of p_b; fw: -dp_b/dz - g rho_b at the z faces plus the
// share of K14 that the column reference's recurrence leaves on the base
// column), so that only departures pass through FP32 pressure arithmetic.  All six are
// nullptr on flat grids.
template <class R>
struct Base {
  const R* rho;  // rho_b (nz)
  const R* rt;   // (rho theta)_b
  const R* th;   // theta_b = (rho theta)_b / rho_b
  const R* p;    // p_b = EOS(rho_b, (rho theta)_b)
  const R* drho = nullptr;
  const R* dth = nullptr;
  const R* dp = nullptr;
  const R* fu = nullptr;
  const R* fv = nullptr;
  const R* fw = nullptr;
};

template <class R>
__device__ inline R p_pert(R rtp, R rtb, R pb) {
  const R gamma = R(1) / (R(1) - Phys<R>::kappa);
  return pb * qexpm1(gamma * qlog1p(qdiv(rtp, rtb)));
}
// FP32: p' = p_b ((1 + x)^gamma - 1), x = (rho theta)' / (rho theta)_b.
// Departures from the level means are small, so for |x| < 1/16 the binomial
// series to x^6 replaces expm1(gamma log1p(x)) (seven fused operations
// instead of ~35 with five transcendentals): the first omitted term,
// C(gamma, 7) x^7, is below 2e-10 of the sum there, far under FP32
// rounding.  Larger departures take the log1p / expm1 form.
template <>
__device__ inline float p_pert<float>(float rtp, float rtb, float pb) {
  constexpr float gf = 1.0f / (1.0f - Phys<float>::kappa);
  constexpr double g = gf;
  constexpr double c1 = g, c2 = c1 * (g - 1) / 2, c3 = c2 * (g - 2) / 3, c4 = c3 * (g - 3) / 4,
                   c5 = c4 * (g - 4) / 5, c6 = c5 * (g - 5) / 6;
  const float x = qdiv(rtp, rtb);
  if (__builtin_fabsf(x) < 0.0625f) {
    float s = static_cast<float>(c6);
    s = __builtin_fmaf(s, x, static_cast<float>(c5));
    s = __builtin_fmaf(s, x, static_cast<float>(c4));
    s = __builtin_fmaf(s, x, static_cast<float>(c3));
    s = __builtin_fmaf(s, x, static_cast<float>(c2));
    s = __builtin_fmaf(s, x, static_cast<float>(c1));
    return pb * (s * x);
  }
  return pb * qexpm1(gf * qlog1p(x));
}
template <class R>
__device__ inline R theta_pert(R rhop, R rtp, R rhob, R thb) {
  return qdiv(rtp - thb * rhop, rmax(rhob + rhop, R(1.0e-6)));
}

// ---------------------------------------------------------------------------
// Device state (pointers into device memory, context precision).
// ---------------------------------------------------------------------------
template <class R>
struct DState {
  R* rho;
  R* rt;
  R* u;
  R* v;
  R* w;
};

constexpr int kMaxTracers = 8;

// Metric arrays on the device (global, no halo, (k*ny+j)*nx+i) plus the
// per-level columns used on flat terrain.
template <class R>
struct DMetrics {
  const R* zc;     // z_centers_metric  nx*ny*nz
  const R* idz_c;  // inv_dz_cell       nx*ny*nz
  const R* idz_f;  // inv_dz_face       nx*ny*(nz+1)
  const R* lev_idz_c;  // nz   (flat columns: identical in every column)
  const R* lev_idz_f;  // nz+1
  const R* lev_zc;     // nz
};

// ---------------------------------------------------------------------------
// Mapped loads (1x1 halo semantics).
// ---------------------------------------------------------------------------
template <class R>
struct Loader {
  Geo g;
  // element of a field in device (global) memory
  __device__ static R at(const R* f, size_t idx) {
    typedef const __attribute__((address_space(1))) R gR;
    return ((gR*)f)[idx];
  }
  __device__ R cell(const R* f, int i, int j, int k) const {
    const int si = map_cell(i, g.nx, g.per_x), sj = map_cell(j, g.ny, g.per_y);
    return (si < 0 || sj < 0) ? R(0) : at(f, g.cell(si, sj, k));
  }
  __device__ R xf(const R* f, int i, int j, int k) const {
    const int si = map_face(i, g.nx, g.per_x), sj = map_cell(j, g.ny, g.per_y);
    return (si < 0 || sj < 0) ? R(0) : at(f, g.xf(si, sj, k));
  }
  __device__ R yf(const R* f, int i, int j, int k) const {
    const int si = map_cell(i, g.nx, g.per_x), sj = map_face(j, g.ny, g.per_y);
    return (si < 0 || sj < 0) ? R(0) : at(f, g.yf(si, sj, k));
  }
  __device__ R zf(const R* f, int i, int j, int k) const {
    const int si = map_cell(i, g.nx, g.per_x), sj = map_cell(j, g.ny, g.per_y);
    return (si < 0 || sj < 0) ? R(0) : at(f, g.zf(si, sj, k));
  }
  // Metric column index with the reference's periodic wrap
  // (dry_fast_modes.cu:24-30).  On an open axis a column beyond the edge
  // (the closed outer faces, whose fluxes are selected away after an
  // unconditional load) is clamped to the edge column, so the load stays
  // inside the metric arrays.
  __device__ int mi(int i) const {
    return g.per_x ? map_cell(i, g.nx, 1) : (i < 0 ? 0 : (i >= g.nx ? g.nx - 1 : i));
  }
  __device__ int mj(int j) const {
    return g.per_y ? map_cell(j, g.ny, 1) : (j < 0 ? 0 : (j >= g.ny ? g.ny - 1 : j));
  }
};

// The work-group index and grid extent a kernel body works with.  A plain
// launch passes the hardware's; a batched launch (several states of an
// ensemble in one grid) passes each member's own, its blocks being a
// contiguous range of the launch that starts at a multiple of 8, so that
// block b of a member runs on XCD b mod 8 as in a launch of its own.
struct VBlock {
  int x, y;    // work-group index
  int gx, gy;  // grid extent
};
__device__ inline VBlock hw_block() {
  return VBlock{static_cast<int>(blockIdx.x), static_cast<int>(blockIdx.y),
                static_cast<int>(gridDim.x), static_cast<int>(gridDim.y)};
}

// Member lookup of a batched launch: entry[b] of work-group b holds its
// member (high 12 bits) and its index within the member's own grid (low 20
// bits), one load; kNoMember marks the padding up to the next member.
constexpr unsigned kNoMember = 0xffffffffu;
constexpr int kBatchBlockBits = 20;
struct BatchMap {
  const unsigned* entry;
};
// Items carry the member's grid extent in fields gx, gy.  Returns null for a
// padding work-group (uniform over it: the whole work-group returns).
template <class Item>
__device__ inline const Item* batch_member(const Item* items, const BatchMap& bm, VBlock& vb) {
  // uniform over the work-group: kept in a scalar register, so the item's
  // fields are scalar loads (the vector-load counter then only tracks the
  // kernel's own loads, which stay in flight across its waits)
  const unsigned e = __builtin_amdgcn_readfirstlane(bm.entry[blockIdx.x]);
  if (e == kNoMember) return nullptr;
  const Item* it = items + (e >> kBatchBlockBits);
  const int b = static_cast<int>(e & ((1u << kBatchBlockBits) - 1u));
  vb.gx = it->gx;
  vb.gy = it->gy;
  vb.x = b % vb.gx;
  vb.y = b / vb.gx;
  return it;
}

// Work-group (bx, by) of a 2-D grid whose work-groups (x fastest) run on XCD
// b mod 8, remapped so that each XCD holds a contiguous band of them.  Rows
// start 4 B past a cache line, so work-groups that are neighbours in x share
// the lines at their ends; on one XCD a shared line is fetched into one L2.
__device__ inline void xcd_block2d(const VBlock& v, int& bx, int& by) {
  bx = v.x;
  by = v.y;
  const int gx = v.gx, nb = gx * v.gy;
  if ((nb & 7) == 0) {
    const int b = bx + gx * by;
    const int r = (b & 7) * (nb >> 3) + (b >> 3);
    bx = r % gx;
    by = r / gx;
  }
}
__device__ inline void xcd_block2d(int& bx, int& by) { xcd_block2d(hw_block(), bx, by); }

// One DPP move of a double (both halves); lanes of rows outside ROWS hold
// unspecified values.
template <int CTRL, int ROWS>
__device__ inline double dpp_f64(double v) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const int lo = __builtin_amdgcn_mov_dpp(static_cast<int>(u & 0xffffffffull), CTRL, ROWS, 0xF,
                                          false);
  const int hi = __builtin_amdgcn_mov_dpp(static_cast<int>(u >> 32), CTRL, ROWS, 0xF, false);
  return __builtin_bit_cast(double, (static_cast<unsigned long long>(static_cast<unsigned>(hi)) << 32) |
                                        static_cast<unsigned>(lo));
}
// Sum over the 64 lanes of a wave (all lanes active), complete in lane 63
// only: lane pairs and quads (quad_perm), half rows (row_half_mirror), rows
// (row_mirror), then rows 0+1 into row 1 and 2+3 into row 3 (row_bcast:15)
// and rows 0-3 into row 3 (row_bcast:31; rows outside a broadcast end with
// values lane 63 never reads).  DPP moves instead of LDS permutes, in a
// fixed order.
__device__ inline double wave_sum_to63(double v) {
  v += dpp_f64<0xB1, 0xF>(v);   // quad_perm [1,0,3,2]
  v += dpp_f64<0x4E, 0xF>(v);   // quad_perm [2,3,0,1]
  v += dpp_f64<0x141, 0xF>(v);  // row_half_mirror
  v += dpp_f64<0x140, 0xF>(v);  // row_mirror
  v += dpp_f64<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_f64<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
  return v;
}
// The same sum of floats (one DPP move per step instead of two, and the
// move folds into the add).
template <int CTRL, int ROWS>
__device__ inline float dpp_f32(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), CTRL, ROWS, 0xF, false));
}
__device__ inline float wave_sum_to63(float v) {
  v += dpp_f32<0xB1, 0xF>(v);
  v += dpp_f32<0x4E, 0xF>(v);
  v += dpp_f32<0x141, 0xF>(v);
  v += dpp_f32<0x140, 0xF>(v);
  v += dpp_f32<0x142, 0xA>(v);
  v += dpp_f32<0x143, 0xC>(v);
  return v;
}

// Element offsets (level 0) of one staged position (ii, jj) of a tile: its
// cell (and z-faces), x-faces ii and ii+1, y-faces jj and jj+1, with the 1x1
// halo map applied once; -1 marks a value that is zero (open side).
struct StagePos {
  int c, u0, u1, v0, v1;
};
__device__ inline StagePos stage_pos(const Geo& g, int ii, int jj) {
  const int ci = map_cell(ii, g.nx, g.per_x), cj = map_cell(jj, g.ny, g.per_y);
  const int fi0 = map_face(ii, g.nx, g.per_x), fi1 = map_face(ii + 1, g.nx, g.per_x);
  const int fj0 = map_face(jj, g.ny, g.per_y), fj1 = map_face(jj + 1, g.ny, g.per_y);
  StagePos p;
  p.c = (ci < 0 || cj < 0) ? -1 : (cj + g.h) * g.pxc + ci + g.h;
  p.u0 = (fi0 < 0 || cj < 0) ? -1 : (cj + g.h) * g.pxu + fi0 + g.h;
  p.u1 = (fi1 < 0 || cj < 0) ? -1 : (cj + g.h) * g.pxu + fi1 + g.h;
  p.v0 = (ci < 0 || fj0 < 0) ? -1 : (fj0 + g.h) * g.pxc + ci + g.h;
  p.v1 = (ci < 0 || fj1 < 0) ? -1 : (fj1 + g.h) * g.pxc + ci + g.h;
  return p;
}
// Element idx of an array addressed by a 32-bit byte offset from its base
// (grids are limited so that every array stays below 4 GiB): the compiler
// then uses the scalar-base + 32-bit vector-offset form of global loads and
// stores instead of 64-bit address arithmetic per access.
template <class T>
__device__ inline T ld32(const T* f, unsigned idx) {
  // Every field lives in device (global) memory: the access is issued as a
  // global one also where the pointer came from memory (the argument items
  // of a batched launch), which the compiler would otherwise issue as flat.
  typedef const __attribute__((address_space(1))) char gchar;
  typedef const __attribute__((address_space(1))) T gT;
  return *(gT*)((gchar*)f + idx * static_cast<unsigned>(sizeof(T)));
}
// The same with a 64-bit element index.
template <class T>
__device__ inline void gstore(T* f, size_t idx, T v) {
  typedef __attribute__((address_space(1))) T gT;
  ((gT*)f)[idx] = v;
}
template <class T>
__device__ inline void st32(T* f, unsigned idx, T v) {
  typedef __attribute__((address_space(1))) char gchar;
  typedef __attribute__((address_space(1))) T gT;
  *(gT*)((gchar*)f + idx * static_cast<unsigned>(sizeof(T))) = v;
}

// Value at level-0 offset off of level k (level stride `plane`), 0 for off < 0.
template <class R>
__device__ inline R ld_at(const R* f, int off, unsigned plane, int k) {
  return off < 0 ? R(0) : f[static_cast<unsigned>(off) + static_cast<unsigned>(k) * plane];
}

// Kernel launch record for optional per-kernel event timing.
struct LaunchTimer;

}  // namespace gwm
