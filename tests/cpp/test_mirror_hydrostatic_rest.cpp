// Hydrostatic rest on a 2x2 virtual-rank layout through the host mirror:
// one SSPRK3 step keeps mass, rho*theta and horizontal momentum (tolerances
// of tests/regression/test_hydrostatic_rest.cpp:50-53 in the reference).
#include "gwm/domain/idealized_domain_builder.hpp"
#include "gwm/dycore/dry_core.hpp"
#include "gwm/dycore/dry_diagnostics.hpp"
#include "check.hpp"

int main() {
  using namespace gwm;
  try {
    domain::RectilinearDomainConfig c{};
    c.nx = 18; c.ny = 18; c.nz = 12; c.ranks_x = 2; c.ranks_y = 2;
    c.dx = 3000.0; c.dy = 3000.0; c.z_top = 18000.0;
    const auto dom = domain::build_rectilinear_domain(c);
    auto states = dycore::make_hydrostatic_rest_state(dom.layout, dom.metrics, 1.22f, 300.0f,
                                                      8200.0f, "hydro");
    CHECK_TRUE(states.size() == 4);
    const auto before = dycore::summarize_dry_states(states);
    dycore::DryStepperConfig s{};
    s.dt = 1.0f;
    s.fast_substeps = 3;
    dycore::NullBoundaryUpdater b;
    dycore::LocalSplitExplicitFastMode f;
    dycore::advance_dry_state_ssprk3(states, dom.layout, dom.metrics, s, b, f);
    const auto after = dycore::summarize_dry_states(states);
    CHECK_NEAR(before.total_dry_mass, after.total_dry_mass, 1.0e-3);
    CHECK_NEAR(before.total_rho_theta_m, after.total_rho_theta_m, 3.0e-2);
    CHECK_NEAR(after.total_horizontal_momentum_x, 0.0, 1.0e-4);
    CHECK_NEAR(after.total_horizontal_momentum_y, 0.0, 1.0e-4);
    // Halos are exchange-filled after the call: the west ghost of rank 1
    // equals the east-most owned cell of rank 0.
    CHECK_TRUE(states[1].rho_d(-1, 3, 4) == states[0].rho_d(8, 3, 4));
    return 0;
  } catch (const std::exception& e) {
    check_fail(e.what());
  }
}
