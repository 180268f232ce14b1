// Plug-in classes through the host mirror: a custom BoundaryUpdater is called
// once per RK stage at t0, t0 + dt and t0 + dt/2 (src/dycore/dry_core.cu:568,
// 591, 620 in the reference); a custom FastModeIntegrator that forwards to
// apply_local_split_explicit_fast_modes reproduces LocalSplitExplicitFastMode
// up to FP32 rounding (the plug-in receives rho*theta rebuilt from rho and
// theta, whose last bit can differ, and pressure differences amplify it in
// the momenta); a custom integrator that does nothing leaves the fast terms
// out.
#include <iostream>
#include <vector>

#include "gwm/domain/idealized_domain_builder.hpp"
#include "gwm/dycore/dry_core.hpp"
#include "gwm/dycore/dry_fast_modes.hpp"
#include "gwm/dycore/idealized_cases.hpp"
#include "check.hpp"

namespace {
struct Recorder final : gwm::dycore::BoundaryUpdater {
  std::vector<float> times;
  void apply(std::vector<gwm::dycore::DryState>& s,
             const std::vector<gwm::domain::SubdomainDescriptor>& layout, float t) override {
    CHECK_TRUE(s.size() == layout.size());
    times.push_back(t);
  }
};
struct Forward final : gwm::dycore::FastModeIntegrator {
  int calls = 0;
  void apply(std::vector<gwm::dycore::DryState>& s,
             const std::vector<gwm::domain::SubdomainDescriptor>& layout,
             const gwm::domain::GridMetrics& m, const gwm::dycore::DryStepperConfig& c) override {
    ++calls;
    gwm::dycore::apply_local_split_explicit_fast_modes(s, layout, m, c);
  }
};
struct Nothing final : gwm::dycore::FastModeIntegrator {
  void apply(std::vector<gwm::dycore::DryState>&,
             const std::vector<gwm::domain::SubdomainDescriptor>&, const gwm::domain::GridMetrics&,
             const gwm::dycore::DryStepperConfig&) override {}
};

double max_rel(const std::vector<gwm::dycore::DryState>& a,
               const std::vector<gwm::dycore::DryState>& b, int f) {
  double num = 0.0, den = 0.0;
  for (size_t n = 0; n < a.size(); ++n) {
    const auto& x = a[n].field(f);
    const auto& y = b[n].field(f);
    for (int k = 0; k < x.nz(); ++k)
      for (int j = 0; j < x.ny(); ++j)
        for (int i = 0; i < x.nx(); ++i) {
          num = std::max(num, std::fabs(double(x(i, j, k)) - double(y(i, j, k))));
          den = std::max(den, std::fabs(double(y(i, j, k))));
        }
  }
  return num / std::max(den, 1e-30);
}
}  // namespace

int main() {
  using namespace gwm;
  try {
    domain::RectilinearDomainConfig c{};
    c.nx = 20; c.ny = 16; c.nz = 10; c.ranks_x = 2; c.ranks_y = 1;
    c.dx = 500.0; c.dy = 500.0; c.z_top = 10000.0;
    const auto dom = domain::build_rectilinear_domain(c);
    dycore::ThermoBubbleConfig bc{};
    bc.rho_background = 1.16f;
    bc.center_z_fraction = 0.2f;
    bc.radius_x_fraction = 0.2f;
    bc.radius_y_fraction = 0.2f;
    bc.radius_z_fraction = 0.2f;
    const auto init = dycore::make_warm_bubble_state(dom.layout, dom.metrics, bc, "bubble");
    auto copy = [&] {
      std::vector<dycore::DryState> v;
      for (const auto& s : init) {
        v.push_back(s.clone_empty_like("c"));
        v.back().copy_all_from(s);
      }
      return v;
    };
    dycore::DryStepperConfig s{};
    s.dt = 2.0f;
    s.fast_substeps = 3;

    auto ref = copy();
    dycore::NullBoundaryUpdater null_b;
    dycore::LocalSplitExplicitFastMode lsf;
    dycore::advance_dry_state_ssprk3(ref, dom.layout, dom.metrics, s, null_b, lsf);

    auto fwd = copy();
    Recorder rec;
    Forward forward;
    dycore::advance_dry_state_ssprk3(fwd, dom.layout, dom.metrics, s, rec, forward);
    CHECK_TRUE(rec.times.size() == 3);
    CHECK_NEAR(rec.times[0], 0.0f, 0.0f);
    CHECK_NEAR(rec.times[1], 2.0f, 0.0f);
    CHECK_NEAR(rec.times[2], 1.0f, 0.0f);
    CHECK_TRUE(forward.calls == 3);
    const double tol[5] = {1.0e-5, 1.0e-5, 2.0e-3, 2.0e-3, 2.0e-3};
    for (int f = 0; f < 5; ++f) {
      const double e = max_rel(fwd, ref, f);
      std::cout << "forwarded fast modes, field " << f << ": max rel " << e << "\n";
      CHECK_TRUE(e < tol[f]);
    }

    auto none = copy();
    Nothing nothing;
    dycore::advance_dry_state_ssprk3(none, dom.layout, dom.metrics, s, null_b, nothing);
    const double e = max_rel(none, ref, GWM_MOM_W);
    std::cout << "no fast modes, w: max rel " << e << "\n";
    CHECK_TRUE(e > 1.0e-2);
    return 0;
  } catch (const std::exception& e) {
    check_fail(e.what());
  }
}
