// gwm_idealized_driver -- flag-compatible replacement of the reference's
// idealized driver (src/core/gwm_idealized_driver.cpp:233-276): build the
// rectilinear domain, initialise one of the four idealized cases, advance
// `--steps` SSPRK3 steps with the null boundary and the split-explicit fast
// modes, and print the initial/final summary JSON.
//
// Unlike the reference loop this one stays on the device between steps: the
// state is created, stepped and summarised through the C-ABI with no host
// copy of the fields.  The virtual-rank split (--ranks-x/--ranks-y) is tiling
// only and the device holds the whole domain.
//
// --plan-view-json writes the reference's plan-view bundle, computed on the
// device (gwm_plan_view) at --plan-view-level.
//
// Extensions: --precision fp32|fp64, --device N, --stretched-levels
// DZ_BOTTOM,DZ_TOP (flat interfaces stretched to --z-top).
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "gwm_hip.h"
#include "gwm/gwm_io.hpp"

namespace {

struct Options {
  std::string case_kind = "density_current";
  // RectilinearDomainConfig (include/gwm/domain/idealized_domain_builder.hpp)
  int nx = 0, ny = 0, nz = 0, halo = 1, ranks_x = 1, ranks_y = 1;
  double dx = 1000.0, dy = 1000.0, z_top = 16000.0;
  bool periodic_x = true, periodic_y = true;
  std::string terrain_kind = "flat";
  double tcx = 0.5, tcy = 0.5, thx = 0.15, thy = 0.15, theight = 0.0;
  // ThermoBubbleConfig / MountainWaveBackgroundConfig defaults.  Values are
  // kept in double and rounded to the context precision where used, so an
  // FP32 run sees the reference's float values and an FP64 run the decimal
  // ones.
  double rho_background = 1.0, theta_background = 300.0;
  double u = 0.0, v = 0.0, w = 0.0;
  bool u_set = false;
  double theta_perturbation = 2.0;
  double cx = 0.5, cy = 0.5, cz = 0.5, rx = 0.1, ry = 0.1, rz = 0.1;
  double rho_surface = 1.18, theta_ref = 300.0, density_scale_height = 8200.0;
  // DryStepperConfig
  double dt = 2.0, gravity = 9.81;
  int fast_substeps = 1;
  int steps = 10;
  std::string summary_json_path, plan_view_json_path;
  int plan_view_level = 0;
  int precision = GWM_F32, device = 0;
  double dz_bottom = 0.0, dz_top = 0.0;
};

void check(int rc) {
  if (rc != GWM_OK) throw std::runtime_error(gwm_last_error());
}

bool parse_bool(const std::string& s) {
  if (s == "true" || s == "1") return true;
  if (s == "false" || s == "0") return false;
  throw std::runtime_error("Invalid boolean value: " + s);
}

void parse(int argc, char** argv, Options& o) {
  for (int n = 1; n < argc; ++n) {
    const std::string a = argv[n];
    auto val = [&]() -> std::string {
      if (n + 1 >= argc) throw std::runtime_error("Missing value for " + a);
      return argv[++n];
    };
    auto i32 = [&] { return std::stoi(val()); };
    auto f64 = [&] { return std::stod(val()); };
    if (a == "--case") o.case_kind = val();
    else if (a == "--nx") o.nx = i32();
    else if (a == "--ny") o.ny = i32();
    else if (a == "--nz") o.nz = i32();
    else if (a == "--halo") o.halo = i32();
    else if (a == "--ranks-x") o.ranks_x = i32();
    else if (a == "--ranks-y") o.ranks_y = i32();
    else if (a == "--dx") o.dx = f64();
    else if (a == "--dy") o.dy = f64();
    else if (a == "--z-top") o.z_top = f64();
    else if (a == "--periodic-x") o.periodic_x = parse_bool(val());
    else if (a == "--periodic-y") o.periodic_y = parse_bool(val());
    else if (a == "--terrain-kind") {
      o.terrain_kind = val();
      if (o.terrain_kind != "flat" && o.terrain_kind != "cosine_mountain")
        throw std::runtime_error("Unsupported terrain kind: " + o.terrain_kind);
    } else if (a == "--terrain-center-x") o.tcx = f64();
    else if (a == "--terrain-center-y") o.tcy = f64();
    else if (a == "--terrain-half-width-x") o.thx = f64();
    else if (a == "--terrain-half-width-y") o.thy = f64();
    else if (a == "--terrain-height") o.theight = f64();
    else if (a == "--rho-background") o.rho_background = f64();
    else if (a == "--theta-background") o.theta_background = f64();
    else if (a == "--u-background") { o.u = f64(); o.u_set = true; }
    else if (a == "--v-background") o.v = f64();
    else if (a == "--w-background") o.w = f64();
    else if (a == "--theta-perturbation") o.theta_perturbation = f64();
    else if (a == "--center-x-fraction") o.cx = f64();
    else if (a == "--center-y-fraction") o.cy = f64();
    else if (a == "--center-z-fraction") o.cz = f64();
    else if (a == "--radius-x-fraction") o.rx = f64();
    else if (a == "--radius-y-fraction") o.ry = f64();
    else if (a == "--radius-z-fraction") o.rz = f64();
    else if (a == "--rho-surface") o.rho_surface = f64();
    else if (a == "--theta-ref") o.theta_ref = f64();
    else if (a == "--density-scale-height") o.density_scale_height = f64();
    else if (a == "--steps") o.steps = i32();
    else if (a == "--dt") o.dt = f64();
    else if (a == "--fast-substeps") o.fast_substeps = i32();
    else if (a == "--gravity") o.gravity = f64();
    else if (a == "--summary-json") o.summary_json_path = val();
    else if (a == "--plan-view-json") o.plan_view_json_path = val();
    else if (a == "--plan-view-level") o.plan_view_level = i32();
    else if (a == "--precision") {
      const auto p = val();
      if (p == "fp32") o.precision = GWM_F32;
      else if (p == "fp64") o.precision = GWM_F64;
      else throw std::runtime_error("Unsupported precision: " + p);
    } else if (a == "--device") o.device = i32();
    else if (a == "--stretched-levels") {
      const auto s = val();
      const auto c = s.find(',');
      if (c == std::string::npos) throw std::runtime_error("--stretched-levels wants DZB,DZT");
      o.dz_bottom = std::stod(s.substr(0, c));
      o.dz_top = std::stod(s.substr(c + 1));
    } else throw std::runtime_error("Unknown argument: " + a);
  }
}

// summary_to_json layout (src/dycore/dry_diagnostics.cpp:98-123).
std::string summary_json(const double* s, const std::string& ind) {
  std::ostringstream o;
  auto range = [&](const char* k, int lo, bool last) {
    o << ind << "\"" << k << "\": {\"min\": " << s[lo] << ", \"max\": " << s[lo + 1] << "}"
      << (last ? "\n" : ",\n");
  };
  o << "{\n";
  o << ind << "\"total_dry_mass\": " << s[GWM_S_TOTAL_DRY_MASS] << ",\n";
  o << ind << "\"total_rho_theta_m\": " << s[GWM_S_TOTAL_RHO_THETA] << ",\n";
  o << ind << "\"total_horizontal_momentum_x\": " << s[GWM_S_TOTAL_MOM_X] << ",\n";
  o << ind << "\"total_horizontal_momentum_y\": " << s[GWM_S_TOTAL_MOM_Y] << ",\n";
  o << ind << "\"total_vertical_momentum\": " << s[GWM_S_TOTAL_MOM_Z] << ",\n";
  range("rho_d", GWM_S_RHO_MIN, false);
  range("theta_m", GWM_S_THETA_MIN, false);
  range("u_face", GWM_S_U_MIN, false);
  range("v_face", GWM_S_V_MIN, false);
  range("w_face", GWM_S_W_MIN, true);
  o << "}";
  return o.str();
}

int run(const Options& o) {
  if (o.nx <= 0 || o.ny <= 0 || o.nz <= 0)
    throw std::runtime_error("RectilinearDomainConfig dimensions must be positive");
  if (o.dx <= 0.0 || o.dy <= 0.0 || o.z_top <= 0.0)
    throw std::runtime_error("RectilinearDomainConfig spacing/top must be positive");
  if (o.ranks_x <= 0 || o.ranks_y <= 0)
    throw std::runtime_error("RectilinearDomainConfig rank counts must be positive");

  gwm_ctx* ctx = nullptr;
  check(gwm_ctx_create(o.device, o.precision, &ctx));
  gwm_grid* grid = nullptr;
  gwm_state* st = nullptr;
  int rc = EXIT_SUCCESS;
  try {
    std::vector<float> terrain(static_cast<size_t>(o.nx) * o.ny, 0.0f);
    const bool cosine = o.terrain_kind == "cosine_mountain";
    if (cosine)
      check(gwm_terrain_cosine_mountain(o.nx, o.ny, o.tcx, o.tcy, o.thx, o.thy, o.theight,
                                        terrain.data()));
    std::vector<double> zi;
    if (o.dz_bottom > 0.0) {
      zi.resize(static_cast<size_t>(o.nz) + 1);
      check(gwm_stretched_levels(o.nz, o.dz_bottom, o.dz_top, o.z_top, zi.data()));
    }
    gwm_grid_desc d{o.nx, o.ny, o.nz, o.halo, o.periodic_x ? 1 : 0, o.periodic_y ? 1 : 0,
                    o.dx, o.dy, o.z_top, 0.25};
    check(gwm_grid_create(ctx, &d, zi.empty() ? nullptr : zi.data(),
                          cosine ? terrain.data() : nullptr, &grid));
    check(gwm_state_create(ctx, grid, 0, &st));

    // make_case_state (gwm_idealized_driver.cpp:192-220)
    const double bubble[12] = {o.rho_background, o.theta_background, o.u, o.v, o.w, o.cx, o.cy,
                               o.cz, o.rx, o.ry, o.rz, o.theta_perturbation};
    if (o.case_kind == "warm_bubble") check(gwm_init_bubble(st, bubble, +1));
    else if (o.case_kind == "density_current") check(gwm_init_bubble(st, bubble, -1));
    else if (o.case_kind == "hydrostatic_rest")
      check(gwm_init_hydrostatic_rest(st, o.rho_surface, o.theta_ref));
    else if (o.case_kind == "mountain_wave_background")
      check(gwm_init_mountain_wave(st, o.rho_surface, o.theta_ref, o.u_set ? o.u : 15.0, o.v,
                                   o.w));
    else
      throw std::runtime_error("Unsupported case kind: " + o.case_kind);

    double initial[GWM_SUMMARY_N], final_[GWM_SUMMARY_N];
    check(gwm_budget_sums(st, initial));
    const gwm_step_config cfg{o.dt, o.fast_substeps, o.gravity, 1, 1};
    const gwm_boundary_config none{GWM_BOUNDARY_NONE, 0.0, nullptr, nullptr};
    for (int n = 0; n < o.steps; ++n) check(gwm_ssprk3_step(st, &cfg, &none));
    check(gwm_budget_sums(st, final_));

    // driver_summary_json (gwm_idealized_driver.cpp:222-231)
    std::ostringstream js;
    js << "{\n";
    js << "  \"case\": \"" << o.case_kind << "\",\n";
    js << "  \"steps\": " << o.steps << ",\n";
    js << "  \"dt\": " << o.dt << ",\n";
    js << "  \"initial\": " << summary_json(initial, "    ") << ",\n";
    js << "  \"final\": " << summary_json(final_, "    ") << "\n";
    js << "}";
    std::cout << js.str() << "\n";
    if (!o.summary_json_path.empty()) {
      std::ofstream out(o.summary_json_path, std::ios::binary);
      if (!out) throw std::runtime_error("Failed to open summary path: " + o.summary_json_path);
      out << js.str() << "\n";
    }
    if (!o.plan_view_json_path.empty()) {
      // extract_dry_plan_view (gwm_idealized_driver.cpp:265-270)
      const auto pv = gwm::io::detail::from_device(st, o.nx, o.ny, o.nz, o.dx, o.dy, o.case_kind,
                                                   o.steps, static_cast<gwm::real>(o.dt),
                                                   o.plan_view_level, false);
      gwm::io::write_plan_view_bundle_json(pv, o.plan_view_json_path);
    }
  } catch (...) {
    if (st) gwm_state_destroy(st);
    if (grid) gwm_grid_destroy(grid);
    gwm_ctx_destroy(ctx);
    throw;
  }
  gwm_state_destroy(st);
  gwm_grid_destroy(grid);
  gwm_ctx_destroy(ctx);
  return rc;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    Options o;
    parse(argc, argv, o);
    return run(o);
  } catch (const std::exception& ex) {
    std::cerr << "gwm_idealized_driver error: " << ex.what() << "\n";
    return EXIT_FAILURE;
  }
}
