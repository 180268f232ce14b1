// Closed-box warm rain through the nominal overload: total water is kept,
// condensation warms rho*theta and makes rain (checks of
// tests/regression/test_warm_rain_closed_box.cpp:93-100 in the reference);
// then the metric overload on a layout with a bucket conserves water plus
// surface precipitation.
#include "gwm/domain/idealized_domain_builder.hpp"
#include "gwm/dycore/dry_core.hpp"
#include "gwm/physics/warm_rain.hpp"
#include "check.hpp"

namespace {
double water(const gwm::state::TracerState& t) {
  return t.total_mass(0) + t.total_mass(1) + t.total_mass(2);
}
}  // namespace

int main() {
  using namespace gwm;
  try {
    std::vector<dycore::DryState> st;
    st.emplace_back(4, 3, 2, 1, "box");
    st.front().fill_constant(1.0f, 300.0f, 0.0f, 0.0f, 0.0f);
    std::vector<state::TracerState> tr;
    tr.emplace_back(state::make_warm_rain_registry(), 4, 3, 2, 1, "box");
    tr.front().fill_zero();
    for (int k = 0; k < 2; ++k)
      for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 4; ++i)
          tr.front().mass(state::kSpecificHumidityTracerName)(i, j, k) = j == 1 ? 0.030f : 0.020f;
    const double w0 = water(tr.front());
    const double th0 = st.front().total_rho_theta_m();
    physics::WarmRainConfig c{};
    c.dt = 2.0f;
    c.cloud_autoconversion_rate = 0.20f;
    c.rain_evaporation_rate = 0.0f;
    physics::apply_warm_rain_microphysics(st, tr, c);
    physics::apply_warm_rain_microphysics(st, tr, c);
    CHECK_NEAR(water(tr.front()), w0, 1.0e-4);
    CHECK_TRUE(st.front().total_rho_theta_m() > th0);
    CHECK_TRUE(tr.front().total_mass(1) + tr.front().total_mass(2) > 0.0);
    CHECK_TRUE(tr.front().total_mass(2) > 0.0);

    // Metric overload with fallout into the bucket, 2x1 layout.
    domain::RectilinearDomainConfig dc{};
    dc.nx = 8; dc.ny = 4; dc.nz = 5; dc.ranks_x = 2; dc.ranks_y = 1;
    dc.dx = 1000.0; dc.dy = 1000.0; dc.z_top = 5000.0;
    const auto dom = domain::build_rectilinear_domain(dc);
    auto dry = dycore::make_hydrostatic_rest_state(dom.layout, dom.metrics, 1.2f, 300.0f, 8000.0f,
                                                   "wr");
    const size_t n3 = 8 * 4 * 5;
    std::vector<real> qv(n3, 0.005f), qc(n3, 0.0f), qr(n3, 0.002f);
    auto tracers = dycore::make_warm_rain_tracers_from_global_fields(dry, dom.layout, 8, 4, qv, &qc,
                                                                     &qr);
    std::vector<physics::WarmRainSurfaceAccumulation> acc(2);
    physics::WarmRainConfig m{};
    m.dt = 5.0f;
    m.rain_terminal_velocity = 12.0f;
    m.rain_evaporation_rate = 0.0f;
    const auto b0 = physics::summarize_warm_rain(tracers, dom.layout, dom.metrics, &acc);
    physics::apply_warm_rain_microphysics(dry, tracers, dom.layout, dom.metrics, m, &acc);
    const auto b1 = physics::summarize_warm_rain(tracers, dom.layout, dom.metrics, &acc);
    CHECK_TRUE(acc[0].matches(4, 4) && acc[1].matches(4, 4));
    CHECK_TRUE(b1.accumulated_surface_precipitation_sum_mm > 0.0);
    const double before = b0.vapor_water_path_sum_kg_m2 + b0.cloud_water_path_sum_kg_m2 +
                          b0.rain_water_path_sum_kg_m2;
    const double after = b1.vapor_water_path_sum_kg_m2 + b1.cloud_water_path_sum_kg_m2 +
                         b1.rain_water_path_sum_kg_m2 + b1.accumulated_surface_precipitation_sum_mm;
    CHECK_NEAR(after / before, 1.0, 1.0e-5);
    return 0;
  } catch (const std::exception& e) {
    check_fail(e.what());
  }
}
