// A constant state stays constant through advance_dry_state_ssprk3 on a 2x2
// layout (tolerances of tests/regression/test_dry_constant_state_bundle.cpp:34-35
// in the reference), and summarize_dry_states reads the closed forms.
#include "gwm/domain/idealized_domain_builder.hpp"
#include "gwm/dycore/dry_core.hpp"
#include "gwm/dycore/dry_diagnostics.hpp"
#include "check.hpp"

int main() {
  using namespace gwm;
  try {
    domain::RectilinearDomainConfig c{};
    c.nx = 16; c.ny = 12; c.nz = 6; c.ranks_x = 2; c.ranks_y = 2;
    c.dx = 1000.0; c.dy = 1000.0; c.z_top = 6000.0;
    const auto dom = domain::build_rectilinear_domain(c);
    auto states = dycore::make_constant_dry_state(dom.layout, 1.1f, 301.0f, 0.0f, 0.0f, 0.0f,
                                                  "const");
    const auto s0 = dycore::summarize_dry_states(states);
    CHECK_NEAR(s0.total_dry_mass, 1.1 * 16 * 12 * 6, 1.0e-3);
    CHECK_NEAR(s0.theta_m.min, 301.0, 1.0e-3);
    dycore::DryStepperConfig s{};
    s.dt = 2.0f;
    s.fast_substeps = 4;
    dycore::NullBoundaryUpdater b;
    dycore::LocalSplitExplicitFastMode f;
    for (int n = 0; n < 3; ++n)
      dycore::advance_dry_state_ssprk3(states, dom.layout, dom.metrics, s, b, f);
    for (const auto& st : states)
      for (int k = 0; k < st.rho_d.nz(); ++k)
        for (int j = 0; j < st.rho_d.ny(); ++j)
          for (int i = 0; i < st.rho_d.nx(); ++i) {
            CHECK_NEAR(st.rho_d(i, j, k), 1.1f, 1.0e-5f);
            CHECK_NEAR(st.rho_theta_m(i, j, k), 1.1f * 301.0f, 1.0e-3f);
          }
    std::vector<dycore::DrySlowTendencies> tend;
    for (const auto& d : dom.layout) tend.emplace_back(d.nx_local(), d.ny_local(), d.nz, d.halo, "t");
    dycore::compute_slow_tendencies(states, dom.layout, dom.metrics, s, tend);
    for (const auto& t : tend) CHECK_NEAR(t.rho_d(2, 2, 2), 0.0, 1.0e-6);
    return 0;
  } catch (const std::exception& e) {
    check_fail(e.what());
  }
}
