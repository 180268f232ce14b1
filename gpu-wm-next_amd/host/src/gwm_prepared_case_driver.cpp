// gwm_prepared_case_driver -- flag-compatible replacement of the reference's
// prepared-case driver (src/core/gwm_prepared_case_driver.cpp:354-509): load
// the analysis state and boundary cache, build the balanced dry state and the
// warm-rain tracers, report the start-up balance, and advance `--steps`
// steps of {dry SSPRK3, tracer SSPRK3, warm rain} with the prepared lateral
// strips, printing the summary JSON with the reference's keys.
//
// The start-up and the balanced boundary fields of each forecast second are
// host work (gwm_ingest.hpp), as in the reference; the steps run on one
// device domain that stays on the device for the whole run, and each RK
// stage's lateral strips are applied there from a boundary reference the
// updaters refresh when the forecast second changes (PreparedCaseDeviceRun).
// GWM_TRACE_DRIVER and GWM_PHASE_SUMMARY_DIR / _STEP / _STEP_MIN / _STEP_MAX
// behave as in the reference, and so does --plan-view-json (the products
// computed on the device at --plan-view-level).
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <optional>
#include <sstream>
#include <stdexcept>
#include <string>

#include "gwm/gwm_io.hpp"

namespace {

using gwm::real;

struct Options {
  std::string analysis_state_path, boundary_cache_path;
  gwm::ingest::PreparedCaseInitConfig init{};
  gwm::dycore::DryStepperConfig step{};
  gwm::physics::WarmRainConfig warm_rain = [] {
    gwm::physics::WarmRainConfig c{};
    c.rain_terminal_velocity = 12.0f;  // prepared-driver default
    return c;
  }();
  int steps = 1;
  bool enable_boundaries = true, enable_fast_modes = true, enable_tracer_transport = true,
       enable_warm_rain = true;
  std::string summary_json_path, plan_view_json_path;
  int plan_view_level = 0;
};

bool env_flag(const char* name) {
  const char* v = std::getenv(name);
  return v != nullptr && std::string(v) != "0" && std::string(v) != "false";
}

std::optional<int> env_int(const char* name) {
  const char* v = std::getenv(name);
  if (v == nullptr || *v == '\0') return std::nullopt;
  return std::stoi(v);
}

void trace(const std::string& msg) {
  if (env_flag("GWM_TRACE_DRIVER")) std::cerr << "[gwm_prepared_case_driver] " << msg << std::endl;
}

bool phase_step_selected(int step) {
  if (const auto s = env_int("GWM_PHASE_SUMMARY_STEP")) return step == *s;
  const auto lo = env_int("GWM_PHASE_SUMMARY_STEP_MIN"), hi = env_int("GWM_PHASE_SUMMARY_STEP_MAX");
  return !(lo && step < *lo) && !(hi && step > *hi);
}

bool phase_wanted(int step) {
  const char* dir = std::getenv("GWM_PHASE_SUMMARY_DIR");
  return dir != nullptr && *dir != '\0' && phase_step_selected(step);
}

void phase_summary(int step, real t0, const std::string& phase,
                   const std::vector<gwm::dycore::DryState>& states,
                   const std::vector<gwm::state::TracerState>& tracers,
                   const std::vector<gwm::domain::SubdomainDescriptor>& layout,
                   const gwm::domain::GridMetrics& metrics,
                   const std::vector<gwm::physics::WarmRainSurfaceAccumulation>& acc) {
  const char* dir = std::getenv("GWM_PHASE_SUMMARY_DIR");
  if (dir == nullptr || *dir == '\0' || !phase_step_selected(step)) return;
  std::filesystem::create_directories(dir);
  const auto s = gwm::core::summarize_runtime_state(states, tracers, layout, metrics, &acc);
  std::ostringstream j;
  j << "{\n  \"step\": " << step << ",\n  \"step_start_seconds\": " << t0
    << ",\n  \"phase\": \"" << phase << "\",\n  \"summary\": "
    << gwm::core::runtime_state_summary_to_json(s, "    ") << "\n}\n";
  const auto path =
      std::filesystem::path(dir) / ("step_" + std::to_string(step) + "_" + phase + ".json");
  std::ofstream out(path, std::ios::binary);
  if (!out) throw std::runtime_error("Failed to open phase summary path: " + path.string());
  out << j.str();
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
    auto f32 = [&] { return std::stof(val()); };
    auto i32 = [&] { return std::stoi(val()); };
    auto b = [&] { return parse_bool(val()); };
    if (a == "--analysis-state") o.analysis_state_path = val();
    else if (a == "--boundary-cache") o.boundary_cache_path = val();
    else if (a == "--steps") o.steps = i32();
    else if (a == "--enable-boundaries") o.enable_boundaries = b();
    else if (a == "--enable-fast-modes") o.enable_fast_modes = b();
    else if (a == "--enable-tracer-transport") o.enable_tracer_transport = b();
    else if (a == "--enable-warm-rain") o.enable_warm_rain = b();
    else if (a == "--warm-rain-condensation-relaxation") o.warm_rain.condensation_relaxation = f32();
    else if (a == "--warm-rain-cloud-autoconversion-threshold")
      o.warm_rain.cloud_autoconversion_threshold = f32();
    else if (a == "--warm-rain-cloud-autoconversion-rate")
      o.warm_rain.cloud_autoconversion_rate = f32();
    else if (a == "--warm-rain-rain-evaporation-rate") o.warm_rain.rain_evaporation_rate = f32();
    else if (a == "--warm-rain-terminal-velocity") o.warm_rain.rain_terminal_velocity = f32();
    else if (a == "--warm-rain-enable-latent-heating") o.warm_rain.enable_latent_heating = b();
    else if (a == "--dt") o.step.dt = f32();
    else if (a == "--fast-substeps") o.step.fast_substeps = i32();
    else if (a == "--gravity") o.step.gravity = f32();
    else if (a == "--fast-update-horizontal-momentum") o.step.fast_update_horizontal_momentum = b();
    else if (a == "--fast-update-density") o.step.fast_update_density = b();
    else if (a == "--halo") o.init.halo = i32();
    else if (a == "--ranks-x") o.init.ranks_x = i32();
    else if (a == "--ranks-y") o.init.ranks_y = i32();
    else if (a == "--terrain-taper-eta") o.init.terrain_taper_eta = std::stod(val());
    else if (a == "--periodic-x") o.init.periodic_x = b();
    else if (a == "--periodic-y") o.init.periodic_y = b();
    else if (a == "--summary-json") o.summary_json_path = val();
    else if (a == "--plan-view-json") o.plan_view_json_path = val();
    else if (a == "--plan-view-level") o.plan_view_level = i32();
    else throw std::runtime_error("Unknown argument: " + a);
  }
  if (o.analysis_state_path.empty()) throw std::runtime_error("--analysis-state is required");
  if (o.boundary_cache_path.empty()) throw std::runtime_error("--boundary-cache is required");
  if (o.steps < 0) throw std::runtime_error("--steps must be nonnegative");
}

// The surface runtime the reference builds and then discards: this build
// checks the same inputs and reports its shape (one tile, four soil layers).
void check_surface_inputs(const gwm::ingest::AnalysisStateIR& a) {
  const size_t n2 = a.grid.cell_count_2d();
  auto need = [n2](const gwm::ingest::FieldBundle2D& b, const char* name) {
    const auto it = b.values.find(name);
    gwm::require(it != b.values.end(), std::string("Missing surface field: ") + name);
    gwm::require(it->second.size() == n2, std::string("Surface field has wrong shape: ") + name);
  };
  need(a.surface, "skin_temperature");
  need(a.static_surface, "terrain_height");
  need(a.static_surface, "land_mask");
  need(a.static_surface, "land_use_index");
}

std::string summary_json(const Options& o, const gwm::ingest::PreparedRuntimeCase& rc,
                         const gwm::physics::WarmRainConfig& wr,
                         const gwm::ingest::PreparedCaseBalanceDiagnostics& bal,
                         const gwm::core::RuntimeStateSummary& initial,
                         const gwm::core::RuntimeStateSummary& final_) {
  const double elapsed = static_cast<double>(o.steps) * static_cast<double>(o.step.dt);
  auto tf = [](bool v) { return v ? "true" : "false"; };
  std::ostringstream j;
  j << std::setprecision(17);
  j << "{\n";
  j << "  \"case\": \"prepared_case\",\n";
  j << "  \"source\": \"" << gwm::ingest::to_string(rc.analysis.source) << "\",\n";
  j << "  \"cycle_time_utc\": \"" << rc.analysis.cycle_time_utc << "\",\n";
  j << "  \"analysis_valid_time_utc\": \"" << rc.analysis.valid_time_utc << "\",\n";
  j << "  \"steps\": " << o.steps << ",\n";
  j << "  \"dt\": " << o.step.dt << ",\n";
  j << "  \"fast_substeps\": " << o.step.fast_substeps << ",\n";
  j << "  \"enable_boundaries\": " << tf(o.enable_boundaries) << ",\n";
  j << "  \"enable_fast_modes\": " << tf(o.enable_fast_modes) << ",\n";
  j << "  \"enable_tracer_transport\": " << tf(o.enable_tracer_transport) << ",\n";
  j << "  \"enable_warm_rain\": " << tf(o.enable_warm_rain) << ",\n";
  j << "  \"fast_update_horizontal_momentum\": " << tf(o.step.fast_update_horizontal_momentum)
    << ",\n";
  j << "  \"fast_update_density\": " << tf(o.step.fast_update_density) << ",\n";
  j << "  \"warm_rain\": {\n";
  j << "    \"condensation_relaxation\": " << wr.condensation_relaxation << ",\n";
  j << "    \"cloud_autoconversion_threshold\": " << wr.cloud_autoconversion_threshold << ",\n";
  j << "    \"cloud_autoconversion_rate\": " << wr.cloud_autoconversion_rate << ",\n";
  j << "    \"rain_evaporation_rate\": " << wr.rain_evaporation_rate << ",\n";
  j << "    \"rain_terminal_velocity\": " << wr.rain_terminal_velocity << ",\n";
  j << "    \"sedimentation_layer_depth_m\": " << wr.sedimentation_layer_depth_m << ",\n";
  j << "    \"enable_latent_heating\": " << tf(wr.enable_latent_heating) << "\n";
  j << "  },\n";
  j << "  \"elapsed_seconds\": " << elapsed << ",\n";
  j << "  \"elapsed_hours\": " << elapsed / 3600.0 << ",\n";
  const auto& g = rc.analysis.grid;
  j << "  \"grid\": {\n";
  j << "    \"nx\": " << g.nx << ",\n";
  j << "    \"ny\": " << g.ny << ",\n";
  j << "    \"nz\": " << g.nz << ",\n";
  j << "    \"dx\": " << g.dx << ",\n";
  j << "    \"dy\": " << g.dy << ",\n";
  j << "    \"z_top\": " << g.z_top << "\n";
  j << "  },\n";
  j << "  \"surface_runtime\": {\n";
  j << "    \"ntile\": 1,\n";
  j << "    \"nsoil\": 4\n";
  j << "  },\n";
  j << "  \"startup_balance\": {\n";
  j << "    \"max_rel_eos\": " << bal.max_rel_eos << ",\n";
  j << "    \"max_rel_hydrostatic\": " << bal.max_rel_hydrostatic << ",\n";
  j << "    \"max_rel_fast_vertical\": " << bal.max_rel_fast_vertical << ",\n";
  j << "    \"max_abs_mass_divergence\": " << bal.max_abs_mass_divergence << ",\n";
  j << "    \"max_rel_mass_divergence\": " << bal.max_rel_mass_divergence << ",\n";
  j << "    \"max_abs_mom_w_bottom\": " << bal.max_abs_mom_w_bottom << ",\n";
  j << "    \"max_abs_mom_w_top\": " << bal.max_abs_mom_w_top << ",\n";
  j << "    \"max_abs_tracer_closure\": " << bal.max_abs_tracer_closure;
  if (bal.max_abs_z_src_minus_metric)
    j << ",\n    \"max_abs_z_src_minus_metric\": " << *bal.max_abs_z_src_minus_metric << "\n";
  else
    j << "\n";
  j << "  },\n";
  j << "  \"initial\": " << gwm::core::runtime_state_summary_to_json(initial, "    ") << ",\n";
  j << "  \"final\": " << gwm::core::runtime_state_summary_to_json(final_, "    ") << "\n";
  j << "}";
  return j.str();
}

}  // namespace

int main(int argc, char** argv) {
  try {
    namespace ing = gwm::ingest;
    trace("startup");
    Options o;
    parse(argc, argv, o);
    trace("args_parsed");
    const auto rc = ing::load_prepared_runtime_case(o.analysis_state_path, o.boundary_cache_path);
    trace("runtime_case_loaded");
    const auto layout = ing::make_prepared_case_layout(rc.analysis, o.init);
    trace("layout_built");
    const auto metrics = ing::make_prepared_case_metrics(rc.analysis, o.init);
    trace("metrics_built");
    auto states = ing::make_dry_states_from_analysis(rc.analysis, metrics, layout, "prepared_case");
    trace("dry_states_initialized");
    auto tracers = ing::make_warm_rain_tracers_from_analysis(rc.analysis, states, layout,
                                                             "prepared_case_tracer");
    trace("warm_rain_tracers_initialized");
    const auto balance =
        ing::diagnose_prepared_case_balance(rc.analysis, metrics, states, layout, &tracers);
    trace("startup_balance_done");
    check_surface_inputs(rc.analysis);
    trace("surface_runtime_initialized");
    std::vector<gwm::physics::WarmRainSurfaceAccumulation> acc(states.size());
    for (size_t r = 0; r < states.size(); ++r)
      acc[r].reset(states[r].rho_d.nx(), states[r].rho_d.ny(), metrics.dx * metrics.dy);
    const auto initial = gwm::core::summarize_runtime_state(states, tracers, layout, metrics, &acc);
    trace("initial_summary_done");

    ing::PreparedCaseBoundaryUpdater dry_bc(rc.analysis, rc.boundary_cache, o.init);
    ing::PreparedCaseTracerBoundaryUpdater tr_bc(rc.analysis, rc.boundary_cache, o.init);
    auto wr = o.warm_rain;
    wr.dt = o.step.dt;

    // The step loop of the reference (gwm_prepared_case_driver.cpp:427-463)
    // on one device domain: the state stays on the device from the first
    // step to the last, the boundary updaters refresh the device boundary
    // reference, and host copies are made only for phase summaries.
    ing::PreparedCaseDeviceRun run(metrics, layout, states, tracers, acc,
                                   o.enable_boundaries ? &dry_bc : nullptr,
                                   o.enable_boundaries ? &tr_bc : nullptr);
    trace("device_state_uploaded");
    auto phase = [&](int step, real t0, const char* name) {
      if (!phase_wanted(step)) return;
      run.download(states, tracers, acc);
      phase_summary(step, t0, name, states, tracers, layout, metrics, acc);
    };
    real sim_time = 0.0f;
    for (int step = 0; step < o.steps; ++step) {
      const std::string tag = "step_" + std::to_string(step);
      trace(tag + "_begin");
      if (o.enable_boundaries) {
        dry_bc.set_step_start_time(sim_time);
        tr_bc.set_step_start_time(sim_time);
        trace(tag + "_boundaries_configured");
      } else {
        trace(tag + "_boundaries_disabled");
      }
      run.dry_step(o.step, o.enable_fast_modes);
      trace(tag + "_dry_done");
      phase(step, sim_time, "dry");
      if (o.enable_tracer_transport) {
        run.tracer_step(o.step);
        trace(tag + "_tracers_done");
        phase(step, sim_time, "tracers");
      } else {
        trace(tag + "_tracers_skipped");
      }
      if (o.enable_warm_rain) {
        run.warm_rain(wr);
        trace(tag + "_warm_rain_done");
        phase(step, sim_time, "warm_rain");
      } else {
        trace(tag + "_warm_rain_skipped");
      }
      sim_time += o.step.dt;
    }
    run.download(states, tracers, acc);
    trace("post_step_sync_done");
    const auto final_ = gwm::core::summarize_runtime_state(states, tracers, layout, metrics, &acc);
    trace("final_summary_done");
    const std::string json = summary_json(o, rc, wr, balance, initial, final_);
    std::cout << json << "\n";
    if (!o.summary_json_path.empty()) {
      std::ofstream out(o.summary_json_path, std::ios::binary);
      if (!out) throw std::runtime_error("Failed to open summary path: " + o.summary_json_path);
      out << json << "\n";
    }
    if (!o.plan_view_json_path.empty()) {
      const auto pv = gwm::io::extract_runtime_plan_view(states, tracers, layout, metrics,
                                                         "prepared_case", o.steps, o.step.dt,
                                                         o.plan_view_level, &acc);
      gwm::io::write_plan_view_bundle_json(pv, o.plan_view_json_path);
    }
    return EXIT_SUCCESS;
  } catch (const std::exception& ex) {
    std::cerr << "gwm_prepared_case_driver error: " << ex.what() << "\n";
    return EXIT_FAILURE;
  }
}
