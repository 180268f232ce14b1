// gwm_output.hip -- output products of the step on the device: the plan-view
// bundle (gwm-next-plan-view/v1) of src/io/plan_view_output.cpp.
//
// The reference derives every product on the host from 3-D host copies of
// the state (one pass per derived field, then a gather per field).  Here one
// kernel walks each column once: the slice level's derived fields, the
// column integrals of cloud and rain water, the column maximum of the
// synthetic reflectivity and the surface bucket, written as [field][j][i]
// doubles, so the host downloads nx*ny*23 values instead of the state.
#include "gwm_kernels.hpp"

namespace gwm {

namespace {

__device__ inline float pv_log(float x) { return logf(x); }
__device__ inline double pv_log(double x) { return log(x); }
__device__ inline float pv_log10(float x) { return log10f(x); }
__device__ inline double pv_log10(double x) { return log10(x); }
__device__ inline float pv_sqrt(float x) { return sqrtf(x); }
__device__ inline double pv_sqrt(double x) { return sqrt(x); }

// relative_humidity_from_specific_humidity (moist_thermo.hpp:38-44)
template <class R>
__device__ inline R pv_rh(R q, R T, R p) {
  const R qs = d_qsat(T, p);
  return rmin(rmax(qs > R(0) ? R(100) * q / qs : R(0), R(0)), R(100));
}

// dewpoint_from_specific_humidity (moist_thermo.hpp:46-55)
template <class R>
__device__ inline R pv_dewpoint(R q_in, R p) {
  const R eps = Phys<R>::eps;
  const R q = rmin(rmax(q_in, R(0)), R(1));
  const R e = p * q / rmax(eps + (R(1) - eps) * q, R(1.0e-6));
  const R e_hpa = rmax(e * R(0.01), R(1.0e-6));
  const R lt = pv_log(e_hpa / R(6.112));
  return (R(243.5) * lt) / (R(17.67) - lt) + R(273.15);
}

// synthetic_reflectivity_dbz_from_rain (moist_thermo.hpp:66-74)
template <class R>
__device__ inline R pv_dbz(R rho, R qr_in) {
  const R qr = rmax(qr_in, R(0));
  if (qr <= R(1.0e-8) || rho <= R(1.0e-8)) return R(-20);
  const R z = R(3.63e9) * rpow(rho * qr, R(1.75));
  return R(10) * pv_log10(rmax(z, R(1.0e-6)));
}

template <class R, bool TERRAIN>
__global__ __launch_bounds__(256) void k_plan_view(PlanViewArgs<R> a) {
  const Geo g = a.g;
  int bx, by;
  xcd_block2d(bx, by);
  const int i = bx * blockDim.x + threadIdx.x;
  const int j = by * blockDim.y + threadIdx.y;
  if (i >= g.nx || j >= g.ny) return;
  const size_t plane = static_cast<size_t>(g.nx) * g.ny;
  const size_t o = static_cast<size_t>(j) * g.nx + i;
  double* out = a.out;
  auto put = [&](int f, double v) { out[static_cast<size_t>(f) * plane + o] = v; };
  const R kMinRho = R(1.0e-6);
  const bool pert = a.base.rho != nullptr;
  auto rho_at = [&](int k) {
    const R r = a.rho[g.cell(i, j, k)];
    return pert ? r + a.base.rho[k] : r;
  };

  // slice level (extract_dry_plan_view, plan_view_output.cpp:630-705)
  const int k = a.k;
  const size_t c = g.cell(i, j, k);
  const R rho = rho_at(k);
  const R rt = pert ? a.rt[c] + a.base.rt[k] : a.rt[c];
  const R rr = rmax(rho, kMinRho);
  const R theta = rt / rr;
  const R p = d_pressure(rho, rt);
  const R T = theta * rpow(rmax(p, R(1)) / Phys<R>::p0, Phys<R>::kappa);
  const R u = R(0.5) * (a.u[g.xf(i, j, k)] + a.u[g.xf(i + 1, j, k)]) / rr;
  const R v = R(0.5) * (a.v[g.yf(i, j, k)] + a.v[g.yf(i, j + 1, k)]) / rr;
  const R w = R(0.5) * (a.w[g.zf(i, j, k)] + a.w[g.zf(i, j, k + 1)]) / rr;
  put(kPvTerrain, static_cast<double>(a.terrain[o]));
  put(kPvZCenter, static_cast<double>(a.zc[g.metric(i, j, k)]));
  put(kPvRho, static_cast<double>(rho));
  put(kPvTheta, static_cast<double>(theta));
  put(kPvPressure, static_cast<double>(p));
  put(kPvTemperature, static_cast<double>(T));
  put(kPvU, static_cast<double>(u));
  put(kPvV, static_cast<double>(v));
  put(kPvWind, static_cast<double>(pv_sqrt(u * u + v * v)));
  put(kPvW, static_cast<double>(w));
  if (!a.qv) return;

  // moist slice fields (extract_runtime_plan_view, :708-760)
  const R q = a.qv[c] / rr;
  put(kPvQ, static_cast<double>(q));
  put(kPvRh, static_cast<double>(pv_rh(q, T, p)));
  put(kPvDewpoint, static_cast<double>(pv_dewpoint(q, p)));
  if (!a.qc || !a.qr) return;
  const R qc = a.qc[c] / rr;
  const R qr = a.qr[c] / rr;
  put(kPvQc, static_cast<double>(qc));
  put(kPvQr, static_cast<double>(qr));
  put(kPvCondensate, static_cast<double>(qc + qr));

  // column maps (derive_column_tracer_map, :282-330; reflectivity, :355-397)
  double col_c = 0.0, col_r = 0.0;
  R dbz = R(-20);
  for (int kk = 0; kk < g.nz; ++kk) {
    const size_t ck = g.cell(i, j, kk);
    const R idz = TERRAIN ? a.idz_c[g.metric(i, j, kk)] : a.lev_idz_c[kk];
    const double dz = static_cast<double>(R(1) / idz);
    col_c += static_cast<double>(a.qc[ck]) * dz;
    col_r += static_cast<double>(a.qr[ck]) * dz;
    const R rk = rho_at(kk);
    dbz = rmax(dbz, pv_dbz(rk, a.qr[ck] / rmax(rk, kMinRho)));
  }
  const double total = col_c + col_r;
  put(kPvColumnCloud, col_c);
  put(kPvColumnRain, col_r);
  put(kPvColumnCondensate, total);
  const double tpos = total > 0.0 ? total : 0.0;
  put(kPvRainFraction, tpos <= 0.0 ? 0.0 : fmin(fmax(col_r / tpos, 0.0), 1.0));
  put(kPvReflectivity, static_cast<double>(dbz));
  if (!a.precip) return;

  // surface bucket (derive_accumulated_surface_precipitation_map, :399-435)
  const double acc = static_cast<double>(a.precip[o]);
  put(kPvPrecip, acc);
  put(kPvPrecipRate, acc / a.elapsed_hours);
}

}  // namespace

template <class R>
void launch_plan_view(hipStream_t st, const PlanViewArgs<R>& a, bool terrain) {
  const dim3 b(64, 4, 1);
  const dim3 gr((a.g.nx + 63) / 64, (a.g.ny + 3) / 4);
  if (terrain)
    hipLaunchKernelGGL((k_plan_view<R, true>), gr, b, 0, st, a);
  else
    hipLaunchKernelGGL((k_plan_view<R, false>), gr, b, 0, st, a);
}

template void launch_plan_view<float>(hipStream_t, const PlanViewArgs<float>&, bool);
template void launch_plan_view<double>(hipStream_t, const PlanViewArgs<double>&, bool);

}  // namespace gwm
