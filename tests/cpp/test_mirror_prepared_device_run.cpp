// The prepared run on the device (PreparedCaseDeviceRun) against the host
// plug-in path of the mirror API (advance_dry_state_ssprk3 and
// advance_passive_tracers_ssprk3 with the prepared boundary updaters, which
// download, edit and re-upload every RK stage): three moist steps of a
// 14x10x6 open-sided case whose two boundary snapshots a minute apart make
// every stage's forecast second a different interpolation.  Both paths write
// the same strip values at the same points of the stage sequence, so the
// states must agree bit for bit; the device path refreshes its boundary
// reference only when a stage's forecast second differs from the last one.
#include <cmath>
#include <vector>

#include "gwm/ingest/prepared_case_init.hpp"
#include "check.hpp"

namespace {

using gwm::real;
using namespace gwm::ingest;

constexpr int NX = 14, NY = 10, NZ = 6;

size_t at3(int i, int j, int k) { return (static_cast<size_t>(k) * NY + j) * NX + i; }

AnalysisStateIR analysis(real warm, real wind, real moist) {
  AnalysisStateIR a;
  a.source = SourceKind::HRRR;
  a.grid.nx = NX;
  a.grid.ny = NY;
  a.grid.nz = NZ;
  a.grid.dx = 2000.0;
  a.grid.dy = 2000.0;
  a.grid.z_top = 6000.0;
  a.cycle_time_utc = a.valid_time_utc = "2026-04-03T00:00:00Z";
  a.metadata["status"] = "populated";
  const size_t n3 = a.grid.cell_count_3d(), n2 = a.grid.cell_count_2d();
  auto& at = a.atmosphere.values;
  for (const char* n : {"air_pressure", "air_temperature", "specific_humidity", "u_wind", "v_wind",
                        "w_wind", "geopotential_height"})
    at[n].resize(n3);
  for (int k = 0; k < NZ; ++k)
    for (int j = 0; j < NY; ++j)
      for (int i = 0; i < NX; ++i) {
        const size_t c = at3(i, j, k);
        const real z = 300.0f + 950.0f * k + 12.0f * i + 7.0f * j;
        at["geopotential_height"][c] = z;
        at["air_temperature"][c] = 300.0f + warm - 0.0065f * z + 0.2f * std::sin(0.7f * i + j);
        at["air_pressure"][c] = 100000.0f * std::exp(-z / 8200.0f);
        at["specific_humidity"][c] = (0.009f + moist) * std::exp(-z / 2500.0f);
        at["u_wind"][c] = 7.0f + wind + 0.15f * i - 0.1f * k;
        at["v_wind"][c] = -2.0f + 0.5f * wind + 0.08f * j;
        at["w_wind"][c] = 0.0f;
      }
  auto& s = a.surface.values;
  s["surface_pressure"] = std::vector<real>(n2, 99000.0f);
  s["air_temperature_2m"] = std::vector<real>(n2, 299.0f + warm);
  s["specific_humidity_2m"] = std::vector<real>(n2, 0.009f);
  s["u_wind_10m"] = std::vector<real>(n2, 5.0f);
  s["v_wind_10m"] = std::vector<real>(n2, 1.0f);
  s["skin_temperature"] = std::vector<real>(n2, 301.0f);
  auto& st = a.static_surface.values;
  st["terrain_height"].resize(n2);
  for (int j = 0; j < NY; ++j)
    for (int i = 0; i < NX; ++i) st["terrain_height"][static_cast<size_t>(j) * NX + i] = 8.0f * i + 5.0f * j;
  st["land_mask"] = std::vector<real>(n2, 1.0f);
  st["land_use_index"] = std::vector<real>(n2, 7.0f);
  return a;
}

BoundaryCacheIR cache() {
  BoundaryCacheIR c;
  c.source = SourceKind::HRRR;
  c.grid = analysis(0, 0, 0).grid;
  c.cycle_time_utc = "2026-04-03T00:00:00Z";
  c.boundary_interval_seconds = 60;
  for (int n = 0; n < 2; ++n) {
    const auto a = analysis(1.5f * n, 2.0f * n, 0.001f * n);
    c.snapshots.push_back(BoundarySnapshotIR{60 * n, a.valid_time_utc, a.atmosphere, a.surface});
  }
  return c;
}

bool same(const gwm::state::Field3D<real>& a, const gwm::state::Field3D<real>& b) {
  for (int k = 0; k < a.nz(); ++k)
    for (int j = 0; j < a.ny(); ++j)
      for (int i = 0; i < a.nx(); ++i)
        if (!(a(i, j, k) == b(i, j, k))) return false;
  return true;
}

}  // namespace

int main() {
  using namespace gwm;
  try {
    const auto a = analysis(0, 0, 0);
    const auto c = cache();
    PreparedCaseInitConfig cfg{};
    cfg.ranks_x = 2;
    cfg.ranks_y = 1;
    const auto layout = make_prepared_case_layout(a, cfg);
    const auto metrics = make_prepared_case_metrics(a, cfg);
    auto states_h = make_dry_states_from_analysis(a, metrics, layout, "host");
    auto tracers_h = make_warm_rain_tracers_from_analysis(a, states_h, layout, "host_tr");
    auto states_d = make_dry_states_from_analysis(a, metrics, layout, "dev");
    auto tracers_d = make_warm_rain_tracers_from_analysis(a, states_d, layout, "dev_tr");
    std::vector<physics::WarmRainSurfaceAccumulation> acc_h(layout.size()), acc_d(layout.size());
    for (size_t r = 0; r < layout.size(); ++r) {
      acc_h[r].reset(layout[r].nx_local(), layout[r].ny_local(), metrics.dx * metrics.dy);
      acc_d[r].reset(layout[r].nx_local(), layout[r].ny_local(), metrics.dx * metrics.dy);
    }
    dycore::DryStepperConfig step{};
    step.dt = 2.0f;
    step.fast_substeps = 3;
    physics::WarmRainConfig wr{};
    wr.dt = step.dt;
    wr.rain_terminal_velocity = 12.0f;

    PreparedCaseBoundaryUpdater bh(a, c, cfg), bd(a, c, cfg);
    PreparedCaseTracerBoundaryUpdater th(a, c, cfg), td(a, c, cfg);
    dycore::LocalSplitExplicitFastMode fast;
    {
      PreparedCaseDeviceRun run(metrics, layout, states_d, tracers_d, acc_d, &bd, &td);
      real t0 = 0.0f;
      for (int n = 0; n < 3; ++n) {
        bh.set_step_start_time(t0);
        th.set_step_start_time(t0);
        bd.set_step_start_time(t0);
        td.set_step_start_time(t0);
        dycore::advance_dry_state_ssprk3(states_h, layout, metrics, step, bh, fast);
        dycore::advance_passive_tracers_ssprk3(tracers_h, states_h, layout, metrics, step, th);
        physics::apply_warm_rain_microphysics(states_h, tracers_h, layout, metrics, wr, &acc_h);
        run.dry_step(step, true);
        run.tracer_step(step);
        run.warm_rain(wr);
        t0 += step.dt;
      }
      run.download(states_d, tracers_d, acc_d);
      // stage seconds 0, 2, 1 | 2, 4, 3 | 4, 6, 5: the reference changes
      // seconds at every stage, so the reference is refreshed nine times
      // (seven distinct seconds are balanced on the host; the updaters keep
      // the recent ones)
      CHECK_TRUE(run.dry_refreshes() == 9 && run.tracer_refreshes() == 9);
    }
    for (size_t r = 0; r < layout.size(); ++r) {
      for (int f = 0; f < 5; ++f) CHECK_TRUE(same(states_h[r].field(f), states_d[r].field(f)));
      for (int q = 0; q < 3; ++q) CHECK_TRUE(same(tracers_h[r].mass(q), tracers_d[r].mass(q)));
      CHECK_TRUE(same(acc_h[r].liquid_precipitation_kg_m2, acc_d[r].liquid_precipitation_kg_m2));
    }
    // the strips moved with the boundary: the west strip differs from the
    // start-up state, and the run stayed finite
    const auto s0 = make_dry_states_from_analysis(a, metrics, layout, "s0");
    CHECK_TRUE(!(s0[0].rho_theta_m(0, 3, 2) == states_d[0].rho_theta_m(0, 3, 2)));
    for (const auto& st : states_d) CHECK_TRUE(std::isfinite(st.rho_d(2, 2, 2)));
    return 0;
  } catch (const std::exception& e) {
    check_fail(e.what());
  }
}
