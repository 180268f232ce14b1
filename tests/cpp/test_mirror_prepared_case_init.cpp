// Prepared-case start-up and lateral strips through the host mirror, on the
// scenario of the reference's unit test tests/unit/test_prepared_case_init.cpp
// (2 x 2 x 3 HRRR analysis on a terrain ramp, open sides, two boundary
// snapshots an hour apart), with its tolerances:
//  - the start-up balance diagnostics are within the reference's bounds
//    (EOS, hydrostatic and fast-vertical residuals 1e-6, mass divergence
//    1e-6 absolute / 1e-3 relative, w = 0 at the ground and the lid, tracer
//    closure 1e-6, source heights present);
//  - rho matches a column integrated independently here (trapezoid
//    hydrostatic recurrence from the analysis pressure at the ground) and
//    rho*theta / rho the analysis theta;
//  - the analysis w does not enter the face momenta;
//  - the tracer masses are rho_d times the analysis mixing ratio, cloud and
//    rain start at zero;
//  - the boundary updaters at t = 1800 s and 3600 s overwrite the strips
//    with the balanced state of the linearly interpolated snapshot, and leave
//    tracers the snapshots do not carry (cloud, rain) as they are.
#include <algorithm>
#include <cmath>
#include <vector>

#include "gwm/core/dry_thermo.hpp"
#include "gwm/ingest/prepared_case_init.hpp"
#include "check.hpp"

namespace {

using gwm::real;
using namespace gwm::ingest;

size_t at3(int i, int j, int k, int nx, int ny) { return (static_cast<size_t>(k) * ny + j) * nx + i; }

// Analysis: p 90 kPa, T 300 K, q 5 g/kg, (u, v, w) = (6, -3, 8), heights
// 120 + 700 k + 40 i + 25 j, terrain 0 / 180 / 320 / 560 m.
AnalysisStateIR analysis(real dp = 0.0f, real dt = 0.0f, real dq = 0.0f, real du = 0.0f,
                         real dv = 0.0f, real dw = 0.0f, real dz = 0.0f) {
  AnalysisStateIR a;
  a.source = SourceKind::HRRR;
  a.grid.nx = 2;
  a.grid.ny = 2;
  a.grid.nz = 3;
  a.grid.dx = 3000.0;
  a.grid.dy = 3000.0;
  a.grid.z_top = 3000.0;
  a.cycle_time_utc = a.valid_time_utc = "2026-04-03T00:00:00Z";
  a.metadata["status"] = "populated";
  const size_t n3 = a.grid.cell_count_3d(), n2 = a.grid.cell_count_2d();
  auto& at = a.atmosphere.values;
  at["air_pressure"] = std::vector<real>(n3, 90000.0f + dp);
  at["air_temperature"] = std::vector<real>(n3, 300.0f + dt);
  at["specific_humidity"] = std::vector<real>(n3, 0.005f + dq);
  at["u_wind"] = std::vector<real>(n3, 6.0f + du);
  at["v_wind"] = std::vector<real>(n3, -3.0f + dv);
  at["w_wind"] = std::vector<real>(n3, 8.0f + dw);
  auto& z = at["geopotential_height"];
  z.resize(n3);
  for (int k = 0; k < 3; ++k)
    for (int j = 0; j < 2; ++j)
      for (int i = 0; i < 2; ++i)
        z[at3(i, j, k, 2, 2)] = 120.0f + 700.0f * k + 40.0f * i + 25.0f * j + dz;
  auto& s = a.surface.values;
  s["surface_pressure"] = std::vector<real>(n2, 95000.0f);
  s["air_temperature_2m"] = std::vector<real>(n2, 299.0f);
  s["specific_humidity_2m"] = std::vector<real>(n2, 0.006f);
  s["u_wind_10m"] = std::vector<real>(n2, 5.0f);
  s["v_wind_10m"] = std::vector<real>(n2, 1.0f);
  s["skin_temperature"] = std::vector<real>(n2, 301.0f);
  auto& st = a.static_surface.values;
  st["terrain_height"] = {0.0f, 180.0f, 320.0f, 560.0f};
  st["land_mask"] = std::vector<real>(n2, 1.0f);
  st["land_use_index"] = std::vector<real>(n2, 7.0f);
  return a;
}

// The snapshot offsets of the cache at t0 + 1 h.
AnalysisStateIR late() { return analysis(-1800.0f, 2.0f, 0.0025f, 3.0f, -1.5f, 20.0f, 30.0f); }
AnalysisStateIR half() { return analysis(-900.0f, 1.0f, 0.00125f, 1.5f, -0.75f, 10.0f, 15.0f); }

BoundaryCacheIR cache() {
  const auto a = analysis();
  const auto b = late();
  BoundaryCacheIR c;
  c.source = a.source;
  c.grid = a.grid;
  c.cycle_time_utc = a.cycle_time_utc;
  c.metadata["status"] = "populated";
  BoundarySnapshotIR s0{0, a.valid_time_utc, a.atmosphere, a.surface};
  BoundarySnapshotIR s1{3600, "2026-04-03T01:00:00Z", b.atmosphere, b.surface};
  c.snapshots = {s0, s1};
  return c;
}

real theta_of(const AnalysisStateIR& a, int i, int j, int k) {
  const size_t c = at3(i, j, k, a.grid.nx, a.grid.ny);
  return a.atmosphere.values.at("air_temperature")[c] *
         std::pow(gwm::core::kReferencePressure /
                      std::max(a.atmosphere.values.at("air_pressure")[c], 1.0f),
                  gwm::core::kKappa);
}

// Independent hydrostatic column: p(0) from the analysis, then the trapezoid
// recurrence over the metric level heights with four fixed-point passes.
std::vector<real> column_rho(const AnalysisStateIR& a, const gwm::domain::GridMetrics& m, int i,
                             int j) {
  using gwm::core::dry_rho_from_pressure_theta;
  const int nz = a.grid.nz;
  std::vector<real> p(nz), r(nz);
  p[0] = std::max(a.atmosphere.values.at("air_pressure")[at3(i, j, 0, 2, 2)], 1.0f);
  r[0] = dry_rho_from_pressure_theta(p[0], theta_of(a, i, j, 0));
  for (int k = 1; k < nz; ++k) {
    const real th = theta_of(a, i, j, k);
    const real dz = m.z_center(i, j, k) - m.z_center(i, j, k - 1);
    real pn = std::max(1.0f, p[k - 1] - 9.81f * r[k - 1] * dz);
    for (int it = 0; it < 4; ++it)
      pn = std::max(p[k - 1] - 9.81f * 0.5f * (r[k - 1] + dry_rho_from_pressure_theta(pn, th)) * dz,
                    1.0f);
    p[k] = pn;
    r[k] = dry_rho_from_pressure_theta(pn, th);
  }
  return r;
}

void check_balance(const PreparedCaseBalanceDiagnostics& d) {
  CHECK_TRUE(d.max_rel_eos <= 1.0e-6f);
  CHECK_TRUE(d.max_rel_hydrostatic <= 1.0e-6f);
  CHECK_TRUE(d.max_rel_fast_vertical <= 1.0e-6f);
  CHECK_TRUE(d.max_abs_mass_divergence <= 1.0e-6f);
  CHECK_TRUE(d.max_rel_mass_divergence <= 1.0e-3f);
  CHECK_TRUE(d.max_abs_mom_w_bottom == 0.0f);
  CHECK_TRUE(d.max_abs_mom_w_top == 0.0f);
  CHECK_TRUE(d.max_abs_tracer_closure <= 1.0e-6f);
  CHECK_TRUE(d.max_abs_z_src_minus_metric.has_value());
}

}  // namespace

int main() {
  using namespace gwm;
  try {
    const auto a = analysis();
    PreparedCaseInitConfig cfg{};
    const auto layout = make_prepared_case_layout(a, cfg);
    CHECK_TRUE(layout.size() == 1);
    CHECK_TRUE(layout[0].nx_local() == 2 && layout[0].ny_local() == 2);
    const auto m = make_prepared_case_metrics(a, cfg);
    CHECK_NEAR(m.terrain_height(0, 0), 0.0, 1e-6);
    CHECK_NEAR(m.terrain_height(1, 1), 560.0, 1e-6);

    auto dry = make_dry_states_from_analysis(a, m, layout);
    auto tr = make_warm_rain_tracers_from_analysis(a, dry, layout, "wr");
    CHECK_TRUE(dry.size() == 1 && tr.size() == 1 && tr[0].size() == 3);
    check_balance(diagnose_prepared_case_balance(a, m, dry, layout, &tr));

    const auto c00 = column_rho(a, m, 0, 0), c11 = column_rho(a, m, 1, 1);
    CHECK_NEAR(dry[0].rho_d(0, 0, 0), c00[0], 1e-5);
    CHECK_NEAR(dry[0].rho_d(1, 1, 2), c11[2], 1e-5);
    CHECK_NEAR(dry[0].rho_theta_m(0, 0, 0) / dry[0].rho_d(0, 0, 0), theta_of(a, 0, 0, 0), 1e-5);
    CHECK_NEAR(dry[0].rho_theta_m(1, 1, 2) / dry[0].rho_d(1, 1, 2), theta_of(a, 1, 1, 2), 1e-5);

    // The analysis w does not enter the momenta.
    auto aw = a;
    auto& w = aw.atmosphere.values["w_wind"];
    for (int k = 0; k < 3; ++k)
      for (int j = 0; j < 2; ++j)
        for (int i = 0; i < 2; ++i) w[at3(i, j, k, 2, 2)] = 50.0f + 2.0f * i - 3.0f * j + 5.0f * k;
    const auto dry_w = make_dry_states_from_analysis(aw, m, layout);
    for (int k = 0; k < 3; ++k) {
      CHECK_NEAR(dry_w[0].mom_u.storage()(0, 0, k), dry[0].mom_u.storage()(0, 0, k), 1e-6);
      CHECK_NEAR(dry_w[0].mom_v.storage()(0, 0, k), dry[0].mom_v.storage()(0, 0, k), 1e-6);
    }
    for (int fk = 0; fk <= 3; ++fk) {
      CHECK_NEAR(dry_w[0].mom_w.storage()(0, 0, fk), dry[0].mom_w.storage()(0, 0, fk), 1e-6);
      CHECK_NEAR(dry_w[0].mom_w.storage()(1, 1, fk), dry[0].mom_w.storage()(1, 1, fk), 1e-6);
    }
    const auto& q = a.atmosphere.values.at("specific_humidity");
    CHECK_NEAR(tr[0].mass(state::kSpecificHumidityTracerName)(0, 0, 0),
               dry[0].rho_d(0, 0, 0) * q[at3(0, 0, 0, 2, 2)], 1e-5);
    CHECK_NEAR(tr[0].mass(state::kSpecificHumidityTracerName)(1, 1, 2),
               dry[0].rho_d(1, 1, 2) * q[at3(1, 1, 2, 2, 2)], 1e-5);
    CHECK_TRUE(tr[0].mass(state::kCloudWaterTracerName).owned_sum() == 0.0);
    CHECK_TRUE(tr[0].mass(state::kRainWaterTracerName).owned_sum() == 0.0);

    // Dry strips at 1800 s (halfway snapshot) and 3600 s (late snapshot).
    auto bs = make_dry_states_from_analysis(a, m, layout, "seed");
    bs[0].fill_constant(0.5f, 290.0f, 0.0f, 0.0f, 0.0f);
    PreparedCaseBoundaryUpdater up(analysis(), cache(), cfg);
    up.set_step_start_time(1800.0f);
    up.apply(bs, layout, 0.0f);
    const auto hs = make_dry_states_from_analysis(half(), m, layout);
    CHECK_NEAR(bs[0].rho_d(0, 0, 1), hs[0].rho_d(0, 0, 1), 1e-5);
    CHECK_NEAR(bs[0].mom_u.storage()(0, 0, 1), hs[0].mom_u.storage()(0, 0, 1), 1e-5);
    CHECK_NEAR(bs[0].mom_v.storage()(1, 1, 2), hs[0].mom_v.storage()(1, 1, 2), 1e-5);
    for (int fk = 0; fk <= 3; ++fk)
      CHECK_NEAR(bs[0].mom_w.storage()(0, 0, fk), hs[0].mom_w.storage()(0, 0, fk), 1e-5);
    bs[0].fill_constant(0.25f, 285.0f, -1.0f, 2.0f, 3.0f);
    up.apply(bs, layout, 1800.0f);
    const auto ls = make_dry_states_from_analysis(late(), m, layout);
    CHECK_NEAR(bs[0].rho_d(0, 0, 1), ls[0].rho_d(0, 0, 1), 1e-5);
    CHECK_NEAR(bs[0].mom_u.storage()(0, 0, 1), ls[0].mom_u.storage()(0, 0, 1), 1e-5);
    CHECK_NEAR(bs[0].mom_v.storage()(1, 1, 2), ls[0].mom_v.storage()(1, 1, 2), 1e-5);

    // Tracer strips: vapour from the snapshot, cloud and rain kept.
    auto bt = make_warm_rain_tracers_from_analysis(a, dry, layout, "bt");
    auto& qv = bt[0].mass(state::kSpecificHumidityTracerName);
    auto& qc = bt[0].mass(state::kCloudWaterTracerName);
    auto& qr = bt[0].mass(state::kRainWaterTracerName);
    qv(0, 0, 0) = 0.0f;
    qc(0, 0, 0) = 0.020f;
    qr(0, 0, 0) = 0.010f;
    PreparedCaseTracerBoundaryUpdater tup(analysis(), cache(), cfg);
    tup.set_step_start_time(1800.0f);
    tup.apply(bt, layout, 0.0f);
    const auto ht = make_warm_rain_tracers_from_analysis(half(), hs, layout, "ht");
    CHECK_NEAR(qv(0, 0, 0), ht[0].mass(state::kSpecificHumidityTracerName)(0, 0, 0), 1e-5);
    CHECK_NEAR(qv(1, 1, 2), ht[0].mass(state::kSpecificHumidityTracerName)(1, 1, 2), 1e-5);
    CHECK_NEAR(qc(0, 0, 0), 0.020, 1e-6);
    CHECK_NEAR(qr(0, 0, 0), 0.010, 1e-6);
    qv.fill(0.0f);
    qc.fill(0.030f);
    qr.fill(0.040f);
    tup.apply(bt, layout, 1800.0f);
    const auto lt = make_warm_rain_tracers_from_analysis(late(), ls, layout, "lt");
    CHECK_NEAR(qv(0, 0, 0), lt[0].mass(state::kSpecificHumidityTracerName)(0, 0, 0), 1e-5);
    CHECK_NEAR(qv(1, 1, 2), lt[0].mass(state::kSpecificHumidityTracerName)(1, 1, 2), 1e-5);
    CHECK_NEAR(qc(0, 0, 0), 0.030, 1e-6);
    CHECK_NEAR(qr(0, 0, 0), 0.040, 1e-6);
    check_balance(diagnose_prepared_case_balance(late(), m, bs, layout, &bt));
  } catch (const std::exception& e) {
    check_fail(e.what());
  }
  return 0;
}
