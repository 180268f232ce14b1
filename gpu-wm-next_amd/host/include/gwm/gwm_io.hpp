// gwm_io.hpp -- plan-view output of the C++ host mirror
// ("gwm-next-plan-view/v1").
//
// Reference declarations mirrored here (FahrenheitResearch/gpu-wm-next):
//   include/gwm/io/plan_view_output.hpp:16-61 (bundle types and functions),
//   include/gwm/core/moist_thermo.hpp:23-74 (the moist diagnostics).
//
// The reference derives every product on the host from 3-D host copies of
// the state, one pass per field, then gathers each field across ranks.  The
// mirror gathers the state once into a device domain and calls
// gwm_plan_view, which computes all products in one pass over the columns;
// only the [field][j][i] planes come back.  The bundle, its field order and
// names, units, the JSON layout and the enrichment from a companion
// analysis state follow the reference, so its readers (tools/verify) work
// unchanged.
#pragma once

#include <filesystem>
#include <fstream>
#include <sstream>

#include "gwm/gwm_ingest.hpp"

namespace gwm::core {

// moist_thermo.hpp:18-21
constexpr real kWaterVaporGasConstant = 461.5f;
constexpr real kLatentHeatVaporization = 2.5e6f;
constexpr real kEpsilon = kDryGasConstant / kWaterVaporGasConstant;

// Magnus form of the saturation vapour pressure over water, in Pa.
inline real saturation_vapor_pressure_pa(real temperature_k) {
  const real tc = temperature_k - 273.15f;
  const real es_hpa = 6.112f * std::exp((17.67f * tc) / std::max(tc + 243.5f, 1.0e-6f));
  return es_hpa * 100.0f;
}

inline real saturation_specific_humidity(real temperature_k, real pressure_pa) {
  const real es = saturation_vapor_pressure_pa(temperature_k);
  const real q = kEpsilon * es / std::max(pressure_pa - (1.0f - kEpsilon) * es, 1000.0f);
  return std::min(std::max(q, 0.0f), 0.05f);
}

// Percent, clipped to [0, 100].
inline real relative_humidity_from_specific_humidity(real q, real temperature_k,
                                                     real pressure_pa) {
  const real qs = saturation_specific_humidity(temperature_k, pressure_pa);
  const real rh = qs > 0.0f ? 100.0f * q / qs : 0.0f;
  return std::min(std::max(rh, 0.0f), 100.0f);
}

// Inverse of the Magnus form at the vapour pressure of q, in K.
inline real dewpoint_from_specific_humidity(real q_in, real pressure_pa) {
  const real q = std::min(std::max(q_in, 0.0f), 1.0f);
  const real e = pressure_pa * q / std::max(kEpsilon + (1.0f - kEpsilon) * q, 1.0e-6f);
  const real lt = std::log(std::max(0.01f * e, 1.0e-6f) / 6.112f);
  return (243.5f * lt) / (17.67f - lt) + 273.15f;
}

// Z = 3.63e9 (rho qr)^1.75 in dBZ, -20 dBZ without rain.
inline real synthetic_reflectivity_dbz_from_rain(real rho_d, real rain_mixing_ratio) {
  const real qr = std::max(rain_mixing_ratio, 0.0f);
  if (qr <= 1.0e-8f || rho_d <= 1.0e-8f) return -20.0f;
  return 10.0f * std::log10(std::max(3.63e9f * std::pow(rho_d * qr, 1.75f), 1.0e-6f));
}

}  // namespace gwm::core

namespace gwm::io {

struct PlanViewField {
  std::string name;
  std::string units;
  std::string location;
  int nx = 0;
  int ny = 0;
  std::vector<double> values;
};

struct PlanViewBundle {
  std::string schema_version = "gwm-next-plan-view/v1";
  std::string case_kind;
  int steps = 0;
  real dt = 0.0f;
  int slice_k = 0;
  int nx = 0;
  int ny = 0;
  double dx = 0.0;
  double dy = 0.0;
  double slice_mean_height_m = 0.0;
  std::vector<PlanViewField> fields;
};

namespace detail {

using gwm::detail::check;

// The device planes of gwm_plan_view, in order: name and units.
struct Plane {
  const char* name;
  const char* units;
};
inline const std::array<Plane, GWM_PLAN_VIEW_FIELDS>& planes() {
  static const std::array<Plane, GWM_PLAN_VIEW_FIELDS> p = {{
      {"terrain_height", "m"},
      {"z_center", "m"},
      {"rho_d", "kg m^-3"},
      {"theta_m", "K"},
      {"air_pressure", "Pa"},
      {"air_temperature", "K"},
      {"u_velocity", "m s^-1"},
      {"v_velocity", "m s^-1"},
      {"wind_speed", "m s^-1"},
      {"w_velocity", "m s^-1"},
      {"specific_humidity", "kg kg^-1"},
      {"relative_humidity", "%"},
      {"dewpoint", "K"},
      {state::kCloudWaterTracerName, "kg kg^-1"},
      {state::kRainWaterTracerName, "kg kg^-1"},
      {"total_condensate", "kg kg^-1"},
      {"column_cloud_water", "kg m^-2"},
      {"column_rain_water", "kg m^-2"},
      {"column_total_condensate", "kg m^-2"},
      {"column_rain_fraction", "1"},
      {"synthetic_reflectivity", "dBZ"},
      {"accumulated_surface_precipitation", "mm"},
      {"mean_surface_precipitation_rate", "mm h^-1"},
  }};
  return p;
}

inline PlanViewField flat_field(const std::string& name, const std::string& units, int nx, int ny,
                                std::vector<double> values) {
  require(static_cast<long>(values.size()) == static_cast<long>(nx) * ny,
          "plan view: field " + name + " does not have nx*ny values");
  return PlanViewField{name, units, "cell_center", nx, ny, std::move(values)};
}

inline void set_or_append(PlanViewBundle& b, PlanViewField f) {
  for (auto& x : b.fields)
    if (x.name == f.name) {
      x = std::move(f);
      return;
    }
  b.fields.push_back(std::move(f));
}

inline const PlanViewField* find(const PlanViewBundle& b, const std::string& name) {
  for (const auto& x : b.fields)
    if (x.name == name) return &x;
  return nullptr;
}

// Index of a named tracer in every rank's registry, or -1 if any lacks it.
inline int tracer_slot(const std::vector<state::TracerState>& tr, const std::string& name) {
  int slot = -1;
  for (const auto& t : tr) {
    const auto idx = t.find(name);
    if (!idx) return -1;
    if (slot < 0) slot = *idx;
    if (*idx != slot) return -1;
  }
  return slot;
}

// The bundle of a device state (gwm_plan_view): header, the planes the
// state's tracers allow, and the mean height of the slice.
inline PlanViewBundle from_device(gwm_state* s, int nx, int ny, int nz, double dx, double dy,
                                  const std::string& case_kind, int steps, real dt, int slice_k,
                                  bool with_precip) {
  const size_t plane = static_cast<size_t>(nx) * ny;
  std::vector<double> out(plane * GWM_PLAN_VIEW_FIELDS);
  int nf = 0;
  check(gwm_plan_view(s, slice_k, steps, static_cast<double>(dt), with_precip ? 1 : 0, out.data(),
                      &nf));
  PlanViewBundle b;
  b.case_kind = case_kind;
  b.steps = steps;
  b.dt = dt;
  b.slice_k = std::clamp(slice_k, 0, nz - 1);
  b.nx = nx;
  b.ny = ny;
  b.dx = dx;
  b.dy = dy;
  for (int f = 0; f < nf; ++f) {
    std::vector<double> v(out.begin() + static_cast<long>(f * plane),
                          out.begin() + static_cast<long>((f + 1) * plane));
    b.fields.push_back(flat_field(planes()[f].name, planes()[f].units, nx, ny, std::move(v)));
  }
  double zsum = 0.0;
  for (double z : b.fields[1].values) zsum += z;
  b.slice_mean_height_m = zsum / static_cast<double>(plane);
  return b;
}

inline PlanViewBundle extract(const std::vector<dycore::DryState>& states,
                              const std::vector<state::TracerState>* tracers,
                              const std::vector<domain::SubdomainDescriptor>& layout,
                              const domain::GridMetrics& metrics, const std::string& case_kind,
                              int steps, real dt, int slice_k,
                              const std::vector<physics::WarmRainSurfaceAccumulation>* acc) {
  require(!states.empty(), "plan view: no states");
  require(states.size() == layout.size(), "plan view: one state per layout rank expected");
  const auto G = gwm::detail::global_of(layout, metrics);
  // device tracer slots: 0 specific humidity, 1 cloud water, 2 rain water
  std::vector<int> slots;
  if (tracers != nullptr && !tracers->empty()) {
    require(tracers->size() == states.size(), "plan view: one tracer state per rank expected");
    const int qv = tracer_slot(*tracers, state::kSpecificHumidityTracerName);
    require(qv >= 0, "plan view: the runtime tracers carry no specific_humidity");
    slots.push_back(qv);
    const int qc = tracer_slot(*tracers, state::kCloudWaterTracerName);
    const int qr = tracer_slot(*tracers, state::kRainWaterTracerName);
    if (qc >= 0 && qr >= 0) {
      slots.push_back(qc);
      slots.push_back(qr);
    }
  }
  const bool with_precip = acc != nullptr && slots.size() == 3;
  if (with_precip) require(acc->size() == layout.size(), "plan view: one bucket per rank expected");
  gwm::detail::StateLease s(metrics.device.get(), static_cast<int>(slots.size()));
  dycore::detail::upload_dry(s, states, layout, G);
  std::vector<real> buf;
  for (size_t t = 0; t < slots.size(); ++t) {
    std::vector<const state::Field3D<real>*> loc;
    for (const auto& x : *tracers) loc.push_back(&x.mass(slots[t]));
    gwm::detail::gather(loc, layout, G, 0, 0, 0, buf);
    check(gwm_state_upload(s, GWM_TRACER0 + static_cast<int>(t), buf.data()));
  }
  if (with_precip) {
    std::vector<real> bucket(static_cast<size_t>(G.nx) * G.ny, 0.0f);
    for (size_t n = 0; n < layout.size(); ++n) {
      const auto& d = layout[n];
      require((*acc)[n].matches(d.nx_local(), d.ny_local()),
              "plan view: surface bucket shape differs from its rank");
      for (int j = 0; j < d.ny_local(); ++j)
        for (int i = 0; i < d.nx_local(); ++i)
          bucket[static_cast<size_t>(d.j_begin + j) * G.nx + d.i_begin + i] =
              (*acc)[n].liquid_precipitation_kg_m2(i, j, 0);
    }
    check(gwm_state_upload(s, GWM_PRECIP, bucket.data()));
  }
  return from_device(s, metrics.nx, metrics.ny, metrics.nz, metrics.dx, metrics.dy, case_kind,
                     steps, dt, slice_k, with_precip);
}

inline void json_string(std::ostringstream& os, const std::string& s) {
  os << '"';
  for (char c : s) {
    if (c == '"' || c == '\\') os << '\\';
    os << c;
  }
  os << '"';
}

}  // namespace detail

// extract_dry_plan_view (src/io/plan_view_output.cpp:601-705)
inline PlanViewBundle extract_dry_plan_view(const std::vector<dycore::DryState>& states,
                                            const std::vector<domain::SubdomainDescriptor>& layout,
                                            const domain::GridMetrics& metrics,
                                            const std::string& case_kind, int steps, real dt,
                                            int slice_k) {
  return detail::extract(states, nullptr, layout, metrics, case_kind, steps, dt, slice_k, nullptr);
}

// extract_runtime_plan_view (src/io/plan_view_output.cpp:707-839)
inline PlanViewBundle extract_runtime_plan_view(
    const std::vector<dycore::DryState>& states, const std::vector<state::TracerState>& tracers,
    const std::vector<domain::SubdomainDescriptor>& layout, const domain::GridMetrics& metrics,
    const std::string& case_kind, int steps, real dt, int slice_k,
    const std::vector<physics::WarmRainSurfaceAccumulation>* accumulations = nullptr) {
  return detail::extract(states, &tracers, layout, metrics, case_kind, steps, dt, slice_k,
                         accumulations);
}

// enrich_plan_view_bundle_from_prepared_case (src/io/plan_view_output.cpp:841-950):
// moisture at the slice from the analysis when the run carried none, and
// the analysis surface fields with their 2 m humidity diagnostics.
inline void enrich_plan_view_bundle_from_prepared_case(PlanViewBundle& bundle,
                                                       const std::string& analysis_state_path) {
  const auto a = ingest::load_analysis_state_json(analysis_state_path);
  require(a.grid.nx == bundle.nx && a.grid.ny == bundle.ny,
          "plan view: the analysis grid differs from the bundle's");
  require(bundle.slice_k >= 0 && bundle.slice_k < a.grid.nz,
          "plan view: slice level outside the analysis grid");
  const auto* p = detail::find(bundle, "air_pressure");
  const auto* T = detail::find(bundle, "air_temperature");
  require(p != nullptr && T != nullptr, "plan view: enrichment needs air_pressure and air_temperature");
  const size_t plane = static_cast<size_t>(bundle.nx) * bundle.ny;
  const real kMinP = 1.0f;
  if (detail::find(bundle, "specific_humidity") == nullptr) {
    auto it = a.atmosphere.values.find("specific_humidity");
    require(it != a.atmosphere.values.end(), "plan view: the analysis has no specific_humidity");
    std::vector<double> q(plane), rh(plane), td(plane);
    for (size_t c = 0; c < plane; ++c) {
      const real qq = std::max(it->second[static_cast<size_t>(bundle.slice_k) * plane + c], 0.0f);
      const real pp = std::max(static_cast<real>(p->values[c]), kMinP);
      q[c] = qq;
      rh[c] = core::relative_humidity_from_specific_humidity(qq, static_cast<real>(T->values[c]), pp);
      td[c] = core::dewpoint_from_specific_humidity(qq, pp);
    }
    detail::set_or_append(bundle, detail::flat_field("specific_humidity", "kg kg^-1", bundle.nx,
                                                     bundle.ny, std::move(q)));
    detail::set_or_append(bundle, detail::flat_field("relative_humidity", "%", bundle.nx, bundle.ny,
                                                     std::move(rh)));
    detail::set_or_append(bundle,
                          detail::flat_field("dewpoint", "K", bundle.nx, bundle.ny, std::move(td)));
  }
  const auto& sf = a.surface.values;
  auto copy2d = [&](const char* name, const char* units) {
    auto it = sf.find(name);
    if (it == sf.end()) return;
    detail::set_or_append(bundle, detail::flat_field(name, units, bundle.nx, bundle.ny,
                                                     std::vector<double>(it->second.begin(),
                                                                         it->second.end())));
  };
  copy2d("surface_pressure", "Pa");
  copy2d("air_temperature_2m", "K");
  copy2d("specific_humidity_2m", "kg kg^-1");
  const auto q2 = sf.find("specific_humidity_2m");
  const auto ps = sf.find("surface_pressure");
  if (q2 != sf.end() && ps != sf.end()) {
    const auto t2 = sf.find("air_temperature_2m");
    require(t2 != sf.end(), "plan view: the analysis has no air_temperature_2m");
    std::vector<double> rh(q2->second.size()), td(q2->second.size());
    for (size_t c = 0; c < q2->second.size(); ++c) {
      const real qq = std::max(q2->second[c], 0.0f);
      const real pp = std::max(ps->second[c], kMinP);
      rh[c] = core::relative_humidity_from_specific_humidity(qq, t2->second[c], pp);
      td[c] = core::dewpoint_from_specific_humidity(qq, pp);
    }
    detail::set_or_append(bundle, detail::flat_field("relative_humidity_2m", "%", bundle.nx,
                                                     bundle.ny, std::move(rh)));
    detail::set_or_append(bundle, detail::flat_field("dewpoint_2m", "K", bundle.nx, bundle.ny,
                                                     std::move(td)));
  }
}

// plan_view_bundle_to_json (src/io/plan_view_output.cpp:952-1000): ten
// significant digits, fields as row-major (y, x) value lists.
inline std::string plan_view_bundle_to_json(const PlanViewBundle& b) {
  std::ostringstream os;
  os << std::setprecision(10);
  os << "{\n  \"schema_version\": ";
  detail::json_string(os, b.schema_version);
  os << ",\n  \"case\": ";
  detail::json_string(os, b.case_kind);
  os << ",\n  \"steps\": " << b.steps << ",\n  \"dt\": " << b.dt << ",\n  \"slice_k\": "
     << b.slice_k << ",\n  \"grid\": {\n    \"nx\": " << b.nx << ",\n    \"ny\": " << b.ny
     << ",\n    \"dx\": " << b.dx << ",\n    \"dy\": " << b.dy
     << ",\n    \"slice_mean_height_m\": " << b.slice_mean_height_m << "\n  },\n  \"fields\": [\n";
  for (size_t n = 0; n < b.fields.size(); ++n) {
    const auto& f = b.fields[n];
    os << "    {\n      \"name\": ";
    detail::json_string(os, f.name);
    os << ",\n      \"units\": ";
    detail::json_string(os, f.units);
    os << ",\n      \"location\": ";
    detail::json_string(os, f.location);
    os << ",\n      \"nx\": " << f.nx << ",\n      \"ny\": " << f.ny
       << ",\n      \"storage\": \"row_major_yx\",\n      \"values\": [";
    for (size_t c = 0; c < f.values.size(); ++c) os << (c ? ", " : "") << f.values[c];
    os << "]\n    }" << (n + 1 < b.fields.size() ? "," : "") << "\n";
  }
  os << "  ]\n}";
  return os.str();
}

// write_plan_view_bundle_json (src/io/plan_view_output.cpp:1002-1012): an
// analysis_state.json beside the output enriches the written bundle.
inline void write_plan_view_bundle_json(const PlanViewBundle& bundle, const std::string& path) {
  PlanViewBundle b = bundle;
  const auto dir = std::filesystem::path(path).parent_path();
  if (!dir.empty() && std::filesystem::exists(dir / "analysis_state.json"))
    enrich_plan_view_bundle_from_prepared_case(b, (dir / "analysis_state.json").string());
  std::ofstream out(path, std::ios::binary);
  if (!out) throw std::runtime_error("plan view: cannot open " + path + " for writing");
  out << plan_view_bundle_to_json(b) << "\n";
}

}  // namespace gwm::io
