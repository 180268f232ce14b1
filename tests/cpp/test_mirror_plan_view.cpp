// Plan-view output through the mirror API with the reference's known answers
// (tests/unit/test_plan_view_output.cpp in the reference): a constant state
// on a 6x4x3 grid over two virtual ranks, 9 km top, slice level 1, 4 steps
// of 2 s; the dry bundle (10 fields), the vapour-only runtime bundle (13),
// the warm-rain bundle with a surface bucket (23), the JSON layout, and the
// enrichment from a companion analysis state (18 fields).
#include <filesystem>
#include <fstream>
#include <set>
#include <sstream>

#include "gwm/comm/virtual_rank_layout.hpp"
#include "gwm/dycore/passive_tracer.hpp"
#include "gwm/io/plan_view_output.hpp"
#include "check.hpp"

namespace {

const gwm::io::PlanViewField& field(const gwm::io::PlanViewBundle& b, const std::string& name) {
  for (const auto& f : b.fields)
    if (f.name == name) return f;
  check_fail("missing field " + name);
  std::abort();
}

int occurrences(const std::string& text, const std::string& what) {
  int n = 0;
  for (size_t p = text.find(what); p != std::string::npos; p = text.find(what, p + what.size())) ++n;
  return n;
}

// An analysis-state document for the 6x4x3 grid: level values of q,
// pressure and height, constant temperature, calm winds, flat land.
std::string analysis_json() {
  const int n2 = 24, n3 = 72;
  auto list = [](int n, auto value_of) {
    std::ostringstream os;
    os << "[";
    for (int c = 0; c < n; ++c) os << (c ? ", " : "") << value_of(c);
    os << "]";
    return os.str();
  };
  auto level = [&](double a, double b, double c) {
    return list(n3, [=](int i) { return i < 24 ? a : (i < 48 ? b : c); });
  };
  auto constant = [&](int n, double v) { return list(n, [=](int) { return v; }); };
  std::ostringstream os;
  os << "{\"schema_version\": \"gwm-next-analysis-state/v1\", \"source\": \"HRRR\", "
     << "\"grid\": {\"nx\": 6, \"ny\": 4, \"nz\": 3, \"dx\": 1000.0, \"dy\": 1000.0, "
     << "\"z_top\": 9000.0}, \"cycle_time_utc\": \"2026-04-04T00:00:00Z\", "
     << "\"valid_time_utc\": \"2026-04-04T00:00:00Z\", \"forecast_offset_seconds\": 0, "
     << "\"metadata\": {\"status\": \"populated\"}, \"atmosphere\": {"
     << "\"u_wind\": " << constant(n3, 0) << ", \"v_wind\": " << constant(n3, 0)
     << ", \"w_wind\": " << constant(n3, 0) << ", \"air_temperature\": " << constant(n3, 300)
     << ", \"specific_humidity\": " << level(0.004, 0.009, 0.014)
     << ", \"air_pressure\": " << level(100000, 90000, 80000)
     << ", \"geopotential_height\": " << level(0, 4500, 9000) << "}, \"surface\": {"
     << "\"surface_pressure\": " << constant(n2, 100000) << ", \"air_temperature_2m\": "
     << constant(n2, 300) << ", \"specific_humidity_2m\": " << constant(n2, 0.004)
     << ", \"u_wind_10m\": " << constant(n2, 0) << ", \"v_wind_10m\": " << constant(n2, 0)
     << ", \"skin_temperature\": " << constant(n2, 300) << "}, \"static_surface\": {"
     << "\"terrain_height\": " << constant(n2, 0) << ", \"land_mask\": " << constant(n2, 1)
     << ", \"land_use_index\": " << constant(n2, 7) << "}}";
  return os.str();
}

}  // namespace

int main() {
  using namespace gwm;
  try {
    const auto layout = comm::VirtualRankLayout::build(6, 4, 3, 1, 2, 1, true, true);
    auto states = dycore::make_constant_dry_state(layout, 1.2f, 300.0f, 10.0f, -4.0f, 0.5f, "pv");
    const auto metrics = domain::GridMetrics::make_hybrid_height(6, 4, 3, 1000.0, 1000.0, 9000.0,
                                                                 std::vector<real>(24, 0.0f));
    const auto b = io::extract_dry_plan_view(states, layout, metrics, "constant", 4, 2.0f, 1);
    CHECK_TRUE(b.schema_version == "gwm-next-plan-view/v1");
    CHECK_TRUE(b.case_kind == "constant" && b.steps == 4 && b.slice_k == 1);
    CHECK_TRUE(b.nx == 6 && b.ny == 4);
    CHECK_NEAR(b.dt, 2.0, 1e-6);
    CHECK_NEAR(b.dx, 1000.0, 1e-9);
    CHECK_NEAR(b.slice_mean_height_m, 4500.0, 1e-6);
    CHECK_TRUE(b.fields.size() == 10);
    std::set<std::string> names;
    for (const auto& f : b.fields) {
      CHECK_TRUE(names.insert(f.name).second);
      CHECK_TRUE(f.location == "cell_center" && !f.units.empty());
      CHECK_TRUE(f.nx == 6 && f.ny == 4 && f.values.size() == 24);
    }
    CHECK_NEAR(field(b, "rho_d").values.front(), 1.2, 1e-6);
    CHECK_NEAR(field(b, "theta_m").values.front(), 300.0, 1e-4);
    CHECK_NEAR(field(b, "u_velocity").values.front(), 10.0, 1e-5);
    CHECK_NEAR(field(b, "v_velocity").values.front(), -4.0, 1e-5);
    CHECK_NEAR(field(b, "w_velocity").values.front(), 0.5, 1e-6);
    CHECK_NEAR(field(b, "wind_speed").values.front(), std::sqrt(116.0), 1e-5);
    CHECK_NEAR(field(b, "terrain_height").values.front(), 0.0, 1e-9);
    CHECK_NEAR(field(b, "z_center").values.front(), 4500.0, 1e-6);
    const real p = core::dry_pressure_from_rho_theta_m(1.2f, 1.2f * 300.0f);
    const real T = 300.0f * std::pow(p / core::kReferencePressure, core::kKappa);
    CHECK_NEAR(field(b, "air_pressure").values.front(), p, 1e-2);
    CHECK_NEAR(field(b, "air_temperature").values.front(), T, 1e-4);

    const auto json = io::plan_view_bundle_to_json(b);
    CHECK_TRUE(json.find("\"schema_version\": \"gwm-next-plan-view/v1\"") != std::string::npos);
    CHECK_TRUE(json.find("\"case\": \"constant\"") != std::string::npos);
    CHECK_TRUE(json.find("\"slice_k\": 1") != std::string::npos);
    CHECK_TRUE(json.find("\"grid\": {") != std::string::npos);
    CHECK_TRUE(json.find("\"slice_mean_height_m\": 4500") != std::string::npos);
    CHECK_TRUE(json.find("\"name\": \"theta_m\"") != std::string::npos);
    CHECK_TRUE(occurrences(json, "\"storage\": \"row_major_yx\"") == 10);
    CHECK_TRUE(occurrences(json, "\"location\": \"cell_center\"") == 10);

    const std::vector<real> qv(72, 0.004f);
    const auto vapour = dycore::make_specific_humidity_tracers_from_global_field(states, layout, 6,
                                                                                 4, qv, "qv");
    const auto r = io::extract_runtime_plan_view(states, vapour, layout, metrics, "rt", 4, 2.0f, 1);
    CHECK_TRUE(r.fields.size() == 13);
    CHECK_NEAR(field(r, "specific_humidity").values.front(), 0.004, 1e-6);
    CHECK_TRUE(field(r, "relative_humidity").values.front() > 0.0);
    CHECK_TRUE(field(r, "relative_humidity").values.front() <= 100.0);
    CHECK_TRUE(field(r, "dewpoint").values.front() < field(b, "air_temperature").values.front());

    const std::vector<real> qc(72, 0.001f), qr(72, 0.0005f);
    const auto warm = dycore::make_warm_rain_tracers_from_global_fields(states, layout, 6, 4, qv,
                                                                        &qc, &qr, "wr");
    std::vector<physics::WarmRainSurfaceAccumulation> acc(layout.size());
    for (const auto& d : layout) {
      acc[static_cast<size_t>(d.rank)].reset(d.nx_local(), d.ny_local());
      acc[static_cast<size_t>(d.rank)].liquid_precipitation_kg_m2(0, 0, 0) = 0.25f;
    }
    const auto w = io::extract_runtime_plan_view(states, warm, layout, metrics, "wr", 4, 2.0f, 1,
                                                 &acc);
    CHECK_TRUE(w.fields.size() == 23);
    CHECK_NEAR(field(w, state::kCloudWaterTracerName).values.front(), 0.001, 1e-6);
    CHECK_NEAR(field(w, state::kRainWaterTracerName).values.front(), 0.0005, 1e-6);
    CHECK_NEAR(field(w, "total_condensate").values.front(), 0.0015, 1e-6);
    CHECK_NEAR(field(w, "column_cloud_water").values.front(), 10.8, 1e-5);
    CHECK_NEAR(field(w, "column_rain_water").values.front(), 5.4, 1e-5);
    CHECK_NEAR(field(w, "column_total_condensate").values.front(), 16.2, 1e-5);
    CHECK_NEAR(field(w, "column_rain_fraction").values.front(), 1.0 / 3.0, 1e-6);
    CHECK_TRUE(field(w, "synthetic_reflectivity").values.front() <= 75.0);
    CHECK_NEAR(field(w, "accumulated_surface_precipitation").values.front(), 0.25, 1e-6);
    CHECK_NEAR(field(w, "mean_surface_precipitation_rate").values.front(), 112.5, 1e-4);
    // the second rank's bucket starts at global column 3
    CHECK_NEAR(field(w, "accumulated_surface_precipitation").values[3], 0.25, 1e-6);
    CHECK_NEAR(field(w, "accumulated_surface_precipitation").values[1], 0.0, 1e-12);

    const auto dir = std::filesystem::temp_directory_path() / "gwm_mirror_plan_view";
    std::filesystem::create_directories(dir);
    {
      std::ofstream f(dir / "analysis_state.json", std::ios::binary);
      f << analysis_json();
    }
    auto enriched = b;
    io::enrich_plan_view_bundle_from_prepared_case(enriched, (dir / "analysis_state.json").string());
    CHECK_TRUE(enriched.fields.size() == 18);
    CHECK_NEAR(field(enriched, "specific_humidity").values.front(), 0.009, 1e-6);
    CHECK_TRUE(field(enriched, "relative_humidity").values.front() > 0.0);
    CHECK_TRUE(field(enriched, "relative_humidity").values.front() <= 100.0);
    CHECK_TRUE(field(enriched, "dewpoint").values.front() <
               field(b, "air_temperature").values.front());
    for (const char* n : {"surface_pressure", "air_temperature_2m", "specific_humidity_2m",
                          "relative_humidity_2m", "dewpoint_2m"})
      CHECK_TRUE(field(enriched, n).values.size() == 24);
    const auto out = dir / "plan_view.json";
    io::write_plan_view_bundle_json(b, out.string());
    std::ifstream in(out, std::ios::binary);
    const std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    for (const char* n : {"\"name\": \"specific_humidity\"", "\"name\": \"relative_humidity\"",
                          "\"name\": \"dewpoint\"", "\"name\": \"specific_humidity_2m\""})
      CHECK_TRUE(text.find(n) != std::string::npos);
    in.close();
    std::filesystem::remove_all(dir);
    return 0;
  } catch (const std::exception& e) {
    check_fail(e.what());
  }
}
