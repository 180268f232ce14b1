// gwm_ingest.hpp -- prepared-case input path of the C++ host mirror.
//
// The reference's gwm_prepared_case_driver starts from two JSON files (an
// analysis state and a lateral-boundary cache, schemas
// "gwm-next-analysis-state/v1" and "gwm-next-boundary-cache/v1"), builds a
// hydrostatically balanced dry state with projected face momenta and the
// warm-rain tracers, and during the run overwrites the lateral strips from
// the boundary cache interpolated in time.  This header provides that path
// with the reference's names and behaviour:
//
//   include/gwm/ingest/canonical_ir.hpp      SourceKind, grid / field IR,
//                                            HRRR and RRFS adapters
//   include/gwm/ingest/analysis_state.hpp    AnalysisStateIR
//   include/gwm/ingest/boundary_cache.hpp    BoundarySnapshotIR, BoundaryCacheIR
//   include/gwm/ingest/source_catalog.hpp    to_string, make_adapter, ...
//   include/gwm/ingest/runtime_case.hpp      loaders, validation, interpolation
//                                            (src/ingest/runtime_case.cpp)
//   include/gwm/ingest/prepared_case_init.hpp  layout, metrics, balanced
//                                            states, tracers, balance
//                                            diagnostics, boundary updaters
//                                            (src/ingest/prepared_case_init.cpp)
//   include/gwm/core/dry_thermo.hpp          host thermodynamics
//   include/gwm/core/runtime_summary.hpp     summarize_runtime_state
//                                            (src/core/runtime_summary.cpp)
//
// Everything here runs on the host at start-up or when the boundary
// snapshot changes; the time step itself runs on the device through the
// mirror's operators.  Arithmetic is the reference's FP32 (`real`).
#pragma once

#include <algorithm>
#include <cctype>
#include <cmath>
#include <iomanip>
#include <limits>
#include <memory>
#include <optional>
#include <sstream>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "gwm/gwm_dycore.hpp"
#include "gwm/gwm_json.hpp"

// ---------------------------------------------------------------------------
// core: host thermodynamics (include/gwm/core/dry_thermo.hpp:15-44)
// ---------------------------------------------------------------------------
namespace gwm::core {

constexpr real kDryGasConstant = 287.05f;
constexpr real kReferencePressure = 100000.0f;
constexpr real kKappa = 0.2854f;

inline real dry_theta_from_conserved(real rho_d, real rho_theta_m) {
  return rho_theta_m / std::max(rho_d, 1.0e-6f);
}
inline real dry_pressure_from_rho_theta_m(real rho_d, real rho_theta_m) {
  const real theta = dry_theta_from_conserved(rho_d, rho_theta_m);
  const real ratio = std::max(rho_d * kDryGasConstant * theta / kReferencePressure, 1.0e-6f);
  return kReferencePressure * std::pow(ratio, 1.0f / (1.0f - kKappa));
}
inline real dry_rho_from_pressure_theta(real pressure, real theta) {
  const real exner = std::pow(std::max(pressure / kReferencePressure, 1.0e-6f), kKappa);
  return pressure / std::max(kDryGasConstant * theta * exner, 1.0e-6f);
}
inline real dry_cp() { return kDryGasConstant / kKappa; }

}  // namespace gwm::core

namespace gwm::ingest {

// ---------------------------------------------------------------------------
// Canonical IR (include/gwm/ingest/canonical_ir.hpp:13-184)
// ---------------------------------------------------------------------------
enum class SourceKind { HRRR, RRFS, RAP, GFS, ERA5, ECMWF };

struct HorizontalGridIR {
  int nx = 0, ny = 0, nz = 0;
  double dx = 0.0, dy = 0.0, z_top = 0.0;
  double ref_lat = 0.0, ref_lon = 0.0, truelat1 = 0.0, truelat2 = 0.0, stand_lon = 0.0;
  std::size_t cell_count_2d() const {
    return static_cast<std::size_t>(nx) * static_cast<std::size_t>(ny);
  }
  std::size_t cell_count_3d() const { return cell_count_2d() * static_cast<std::size_t>(nz); }
};

struct FieldBundle3D {
  std::unordered_map<std::string, std::vector<real>> values;
  bool has(const std::string& n) const { return values.count(n) != 0; }
  std::size_t field_count() const { return values.size(); }
};
struct FieldBundle2D {
  std::unordered_map<std::string, std::vector<real>> values;
  bool has(const std::string& n) const { return values.count(n) != 0; }
  std::size_t field_count() const { return values.size(); }
};

struct SourceFieldBinding {
  std::string source_name;
  std::string canonical_name;
};

struct CanonicalAnalysisIR {
  SourceKind source = SourceKind::HRRR;
  HorizontalGridIR grid;
  std::string cycle_time_utc, valid_time_utc;
  int forecast_offset_seconds = 0;
  std::unordered_map<std::string, std::string> metadata;
  FieldBundle3D atmosphere;
  FieldBundle2D surface, static_surface;
  bool empty() const {
    return atmosphere.values.empty() && surface.values.empty() && static_surface.values.empty();
  }
};

// A source's field bindings (source GRIB name -> canonical name).  The
// canonical names are the schema the runtime reads; the adapters differ only
// in the source names.
class SourceAdapter {
 public:
  virtual ~SourceAdapter() = default;
  virtual SourceKind kind() const = 0;
  virtual std::string name() const = 0;
  virtual std::vector<SourceFieldBinding> atmosphere_field_bindings() const = 0;
  virtual std::vector<SourceFieldBinding> surface_field_bindings() const = 0;
  virtual std::vector<SourceFieldBinding> static_field_bindings() const = 0;
  virtual int boundary_interval_seconds() const = 0;
  virtual std::vector<std::string> required_fields() const {
    std::vector<std::string> out;
    auto add =[&out](const std::vector<SourceFieldBinding>& b) {
      for (const auto& x : b)
        if (std::find(out.begin(), out.end(), x.source_name) == out.end())
          out.push_back(x.source_name);
    };
    add(atmosphere_field_bindings());
    add(surface_field_bindings());
    add(static_field_bindings());
    return out;
  }
};

namespace detail {
// Canonical names of the three groups, in the adapters' order.
inline const std::vector<std::string>& canonical_atmosphere() {
  static const std::vector<std::string> v = {"u_wind", "v_wind", "w_wind", "air_temperature",
                                             "specific_humidity", "air_pressure",
                                             "geopotential_height"};
  return v;
}
inline const std::vector<std::string>& canonical_surface() {
  static const std::vector<std::string> v = {"surface_pressure", "air_temperature_2m",
                                             "specific_humidity_2m", "u_wind_10m", "v_wind_10m",
                                             "skin_temperature"};
  return v;
}
inline const std::vector<std::string>& canonical_static() {
  static const std::vector<std::string> v = {"terrain_height", "land_mask", "land_use_index"};
  return v;
}
inline std::vector<SourceFieldBinding> bind(const std::vector<std::string>& source,
                                            const std::vector<std::string>& canonical) {
  std::vector<SourceFieldBinding> out;
  for (size_t n = 0; n < source.size(); ++n) out.push_back({source[n], canonical[n]});
  return out;
}
}  // namespace detail

class HrrrAdapter final : public SourceAdapter {
 public:
  SourceKind kind() const override { return SourceKind::HRRR; }
  std::string name() const override { return "HRRR"; }
  std::vector<SourceFieldBinding> atmosphere_field_bindings() const override {
    return detail::bind({"UGRD", "VGRD", "VVEL", "T", "SPFH", "PRES", "HGT"},
                        detail::canonical_atmosphere());
  }
  std::vector<SourceFieldBinding> surface_field_bindings() const override {
    return detail::bind({"PSFC", "T2", "Q2", "U10", "V10", "TSK"}, detail::canonical_surface());
  }
  std::vector<SourceFieldBinding> static_field_bindings() const override {
    return detail::bind({"HGT", "LANDMASK", "LU_INDEX"}, detail::canonical_static());
  }
  int boundary_interval_seconds() const override { return 3600; }
};

class RrfsAdapter final : public SourceAdapter {
 public:
  SourceKind kind() const override { return SourceKind::RRFS; }
  std::string name() const override { return "RRFS"; }
  std::vector<SourceFieldBinding> atmosphere_field_bindings() const override {
    return detail::bind({"UGRD", "VGRD", "DZDT", "TMP", "SPFH", "PRES", "HGT"},
                        detail::canonical_atmosphere());
  }
  std::vector<SourceFieldBinding> surface_field_bindings() const override {
    return detail::bind({"PRES", "TMP_2M", "SPFH_2M", "UGRD_10M", "VGRD_10M", "TMP_SFC"},
                        detail::canonical_surface());
  }
  std::vector<SourceFieldBinding> static_field_bindings() const override {
    return detail::bind({"HGT_SFC", "LANDMASK", "LANDUSE"}, detail::canonical_static());
  }
  int boundary_interval_seconds() const override { return 3600; }
};

// ---------------------------------------------------------------------------
// Analysis state and boundary cache (analysis_state.hpp:13-69,
// boundary_cache.hpp:12-79)
// ---------------------------------------------------------------------------
namespace detail {
template <class Bundle>
bool has_all(const Bundle& b, const std::vector<SourceFieldBinding>& bindings) {
  for (const auto& x : bindings)
    if (!b.has(x.canonical_name)) return false;
  return true;
}
}  // namespace detail

struct AnalysisStateIR {
  SourceKind source = SourceKind::HRRR;
  HorizontalGridIR grid;
  std::string cycle_time_utc, valid_time_utc;
  int forecast_offset_seconds = 0;
  std::unordered_map<std::string, std::string> metadata;
  FieldBundle3D atmosphere;
  FieldBundle2D surface, static_surface;

  bool has_required_fields(const SourceAdapter& a) const {
    return detail::has_all(atmosphere, a.atmosphere_field_bindings()) &&
           detail::has_all(surface, a.surface_field_bindings()) &&
           detail::has_all(static_surface, a.static_field_bindings());
  }
  CanonicalAnalysisIR to_canonical() const {
    CanonicalAnalysisIR ir;
    ir.source = source;
    ir.grid = grid;
    ir.cycle_time_utc = cycle_time_utc;
    ir.valid_time_utc = valid_time_utc;
    ir.forecast_offset_seconds = forecast_offset_seconds;
    ir.metadata = metadata;
    ir.atmosphere = atmosphere;
    ir.surface = surface;
    ir.static_surface = static_surface;
    return ir;
  }
  static AnalysisStateIR from_canonical(const CanonicalAnalysisIR& ir) {
    AnalysisStateIR s;
    s.source = ir.source;
    s.grid = ir.grid;
    s.cycle_time_utc = ir.cycle_time_utc;
    s.valid_time_utc = ir.valid_time_utc;
    s.forecast_offset_seconds = ir.forecast_offset_seconds;
    s.metadata = ir.metadata;
    s.atmosphere = ir.atmosphere;
    s.surface = ir.surface;
    s.static_surface = ir.static_surface;
    return s;
  }
};

struct BoundarySnapshotIR {
  int forecast_offset_seconds = 0;
  std::string valid_time_utc;
  FieldBundle3D atmosphere;
  FieldBundle2D surface;
  bool has_required_fields(const SourceAdapter& a) const {
    return detail::has_all(atmosphere, a.atmosphere_field_bindings()) &&
           detail::has_all(surface, a.surface_field_bindings());
  }
};

struct BoundaryCacheIR {
  SourceKind source = SourceKind::HRRR;
  HorizontalGridIR grid;
  std::string cycle_time_utc;
  int boundary_interval_seconds = 3600;
  std::unordered_map<std::string, std::string> metadata;
  std::vector<BoundarySnapshotIR> snapshots;

  bool empty() const { return snapshots.empty(); }
  bool has_monotonic_offsets() const {
    for (size_t n = 1; n < snapshots.size(); ++n)
      if (snapshots[n - 1].forecast_offset_seconds >= snapshots[n].forecast_offset_seconds)
        return false;
    return true;
  }
  // The last snapshot at or before `t` and the first at or after it, each
  // falling back to the first / last snapshot outside the cache's range.
  std::pair<const BoundarySnapshotIR*, const BoundarySnapshotIR*> bracket(int t) const {
    if (snapshots.empty()) return {nullptr, nullptr};
    const BoundarySnapshotIR *lo = nullptr, *hi = nullptr;
    for (const auto& s : snapshots) {
      if (s.forecast_offset_seconds <= t) lo = &s;
      if (s.forecast_offset_seconds >= t) {
        hi = &s;
        break;
      }
    }
    return {lo ? lo : &snapshots.front(), hi ? hi : &snapshots.back()};
  }
};

// ---------------------------------------------------------------------------
// Source catalog (source_catalog.hpp:9-25, src/ingest/source_catalog.cpp)
// ---------------------------------------------------------------------------
struct ExternalToolDescriptor {
  std::string name, local_repo_hint, role;
  bool allowed_in_runtime_core = false;
};

inline std::string to_string(SourceKind k) {
  switch (k) {
    case SourceKind::HRRR: return "HRRR";
    case SourceKind::RRFS: return "RRFS";
    case SourceKind::RAP: return "RAP";
    case SourceKind::GFS: return "GFS";
    case SourceKind::ERA5: return "ERA5";
    case SourceKind::ECMWF: return "ECMWF";
  }
  throw std::runtime_error("Unknown SourceKind");
}
inline SourceKind source_kind_from_string(const std::string& value) {
  std::string u = value;
  for (auto& c : u) c = static_cast<char>(std::toupper(static_cast<unsigned char>(c)));
  for (SourceKind k : {SourceKind::HRRR, SourceKind::RRFS, SourceKind::RAP, SourceKind::GFS,
                       SourceKind::ERA5, SourceKind::ECMWF})
    if (u == to_string(k)) return k;
  throw std::runtime_error("Unknown source string: " + value);
}
inline bool is_first_class_source(SourceKind k) {
  return k == SourceKind::HRRR || k == SourceKind::RRFS;
}
inline std::unique_ptr<SourceAdapter> make_adapter(SourceKind k) {
  if (k == SourceKind::HRRR) return std::make_unique<HrrrAdapter>();
  if (k == SourceKind::RRFS) return std::make_unique<RrfsAdapter>();
  return nullptr;  // RAP, GFS, ERA5, ECMWF: no runtime adapter
}
inline std::vector<std::unique_ptr<SourceAdapter>> make_default_adapters() {
  std::vector<std::unique_ptr<SourceAdapter>> v;
  v.push_back(std::make_unique<HrrrAdapter>());
  v.push_back(std::make_unique<RrfsAdapter>());
  return v;
}
// The reference lists its authors' external (non-runtime) tools here; none
// of them is part of this build.
inline std::vector<ExternalToolDescriptor> recommended_external_tools() { return {}; }

// ---------------------------------------------------------------------------
// Runtime case: loaders, validation, interpolation (runtime_case.hpp:8-27,
// src/ingest/runtime_case.cpp)
// ---------------------------------------------------------------------------
struct PreparedRuntimeCase {
  AnalysisStateIR analysis;
  BoundaryCacheIR boundary_cache;
};

namespace detail {

inline SourceKind parse_source_exact(const std::string& s) {
  for (SourceKind k : {SourceKind::HRRR, SourceKind::RRFS, SourceKind::RAP, SourceKind::GFS,
                       SourceKind::ERA5, SourceKind::ECMWF})
    if (s == to_string(k)) return k;
  throw std::runtime_error("Unsupported source string: " + s);
}

inline HorizontalGridIR parse_grid(const json::Value& v) {
  HorizontalGridIR g;
  g.nx = v.at("nx").as_int();
  g.ny = v.at("ny").as_int();
  g.nz = v.at("nz").as_int();
  g.dx = v.at("dx").as_number();
  g.dy = v.at("dy").as_number();
  auto opt = [&v](const char* key, double& out) {
    if (v.has(key)) out = v.at(key).as_number();
  };
  opt("z_top", g.z_top);
  opt("ref_lat", g.ref_lat);
  opt("ref_lon", g.ref_lon);
  opt("truelat1", g.truelat1);
  opt("truelat2", g.truelat2);
  opt("stand_lon", g.stand_lon);
  return g;
}

template <class Bundle>
Bundle parse_bundle(const json::Value& v) {
  Bundle b;
  for (const auto& [name, entry] : v.as_object()) b.values.emplace(name, json::to_real_vector(entry));
  return b;
}

template <class Bundle>
void require_sizes(const Bundle& b, std::size_t n, const char* label) {
  for (const auto& [name, values] : b.values)
    require(values.size() == n, std::string(label) + " field has wrong size: " + name);
}

template <class Bundle>
Bundle blend(const Bundle& lo, const Bundle& hi, real alpha, const char* what) {
  require(lo.values.size() == hi.values.size(),
          std::string("Boundary ") + what + " bundle count mismatch");
  Bundle out;
  for (const auto& [name, a] : lo.values) {
    const auto it = hi.values.find(name);
    require(it != hi.values.end(),
            std::string("Upper boundary ") + what + " is missing field: " + name);
    const auto& b = it->second;
    require(a.size() == b.size(), std::string("Boundary ") + what + " field size mismatch: " + name);
    auto& d = out.values[name];
    d.resize(a.size());
    for (size_t n = 0; n < a.size(); ++n) d[n] = (1.0f - alpha) * a[n] + alpha * b[n];
  }
  return out;
}

}  // namespace detail

inline void validate_runtime_analysis(const AnalysisStateIR& a) {
  require(a.grid.nx > 0 && a.grid.ny > 0 && a.grid.nz > 0,
          "Analysis grid dimensions must be positive");
  require(a.grid.dx > 0.0 && a.grid.dy > 0.0 && a.grid.z_top > 0.0,
          "Analysis grid spacing and z_top must be positive");
  const auto adapter = make_adapter(a.source);
  require(adapter != nullptr, "Analysis source is unsupported by runtime adapter");
  require(a.has_required_fields(*adapter), "Analysis state is missing required canonical fields");
  detail::require_sizes(a.atmosphere, a.grid.cell_count_3d(), "Analysis atmosphere");
  detail::require_sizes(a.surface, a.grid.cell_count_2d(), "Analysis surface");
  detail::require_sizes(a.static_surface, a.grid.cell_count_2d(), "Analysis static_surface");
}

inline void validate_runtime_boundary_cache(const BoundaryCacheIR& c) {
  require(c.grid.nx > 0 && c.grid.ny > 0 && c.grid.nz > 0,
          "Boundary-cache grid dimensions must be positive");
  require(c.boundary_interval_seconds > 0, "Boundary-cache interval must be positive");
  const auto adapter = make_adapter(c.source);
  require(adapter != nullptr, "Boundary-cache source is unsupported by runtime adapter");
  require(!c.empty(), "Boundary cache must contain snapshots");
  require(c.has_monotonic_offsets(), "Boundary cache offsets must be strictly monotonic");
  for (const auto& s : c.snapshots) {
    require(s.has_required_fields(*adapter),
            "Boundary snapshot is missing required canonical fields");
    detail::require_sizes(s.atmosphere, c.grid.cell_count_3d(), "Boundary atmosphere");
    detail::require_sizes(s.surface, c.grid.cell_count_2d(), "Boundary surface");
  }
}

inline AnalysisStateIR load_analysis_state_json(const std::string& path) {
  const auto root = json::parse_file(path);
  require(root.at("schema_version").as_string() == "gwm-next-analysis-state/v1",
          "Unsupported analysis-state schema");
  AnalysisStateIR a;
  a.source = detail::parse_source_exact(root.at("source").as_string());
  a.grid = detail::parse_grid(root.at("grid"));
  a.cycle_time_utc = root.at("cycle_time_utc").as_string();
  a.valid_time_utc = root.at("valid_time_utc").as_string();
  a.forecast_offset_seconds = root.at("forecast_offset_seconds").as_int();
  a.metadata = json::to_string_map(root.at("metadata"));
  a.atmosphere = detail::parse_bundle<FieldBundle3D>(root.at("atmosphere"));
  a.surface = detail::parse_bundle<FieldBundle2D>(root.at("surface"));
  a.static_surface = detail::parse_bundle<FieldBundle2D>(root.at("static_surface"));
  validate_runtime_analysis(a);
  return a;
}

inline BoundaryCacheIR load_boundary_cache_json(const std::string& path) {
  const auto root = json::parse_file(path);
  require(root.at("schema_version").as_string() == "gwm-next-boundary-cache/v1",
          "Unsupported boundary-cache schema");
  BoundaryCacheIR c;
  c.source = detail::parse_source_exact(root.at("source").as_string());
  c.grid = detail::parse_grid(root.at("grid"));
  c.cycle_time_utc = root.at("cycle_time_utc").as_string();
  c.boundary_interval_seconds = root.at("boundary_interval_seconds").as_int();
  c.metadata = json::to_string_map(root.at("metadata"));
  for (const auto& sv : root.at("snapshots").as_array()) {
    BoundarySnapshotIR s;
    s.forecast_offset_seconds = sv.at("forecast_offset_seconds").as_int();
    s.valid_time_utc = sv.at("valid_time_utc").as_string();
    s.atmosphere = detail::parse_bundle<FieldBundle3D>(sv.at("atmosphere"));
    s.surface = detail::parse_bundle<FieldBundle2D>(sv.at("surface"));
    c.snapshots.push_back(std::move(s));
  }
  validate_runtime_boundary_cache(c);
  return c;
}

inline PreparedRuntimeCase load_prepared_runtime_case(const std::string& analysis_path,
                                                      const std::string& boundary_path) {
  PreparedRuntimeCase rc;
  rc.analysis = load_analysis_state_json(analysis_path);
  rc.boundary_cache = load_boundary_cache_json(boundary_path);
  require(rc.analysis.source == rc.boundary_cache.source,
          "Analysis and boundary cache source mismatch");
  require(rc.analysis.grid.nx == rc.boundary_cache.grid.nx &&
              rc.analysis.grid.ny == rc.boundary_cache.grid.ny &&
              rc.analysis.grid.nz == rc.boundary_cache.grid.nz,
          "Analysis and boundary cache grid mismatch");
  return rc;
}

// Linear interpolation in time between the bracketing snapshots.
inline BoundarySnapshotIR interpolate_boundary_snapshot(const BoundaryCacheIR& c, int t) {
  validate_runtime_boundary_cache(c);
  const auto [lo, hi] = c.bracket(t);
  require(lo != nullptr && hi != nullptr, "Boundary cache bracket returned null snapshot");
  if (lo == hi || lo->forecast_offset_seconds == hi->forecast_offset_seconds) return *lo;
  const real alpha = static_cast<real>(t - lo->forecast_offset_seconds) /
                     static_cast<real>(hi->forecast_offset_seconds - lo->forecast_offset_seconds);
  BoundarySnapshotIR out;
  out.forecast_offset_seconds = t;
  out.valid_time_utc = lo->valid_time_utc + "->" + hi->valid_time_utc;
  out.atmosphere = detail::blend(lo->atmosphere, hi->atmosphere, alpha, "atmosphere");
  out.surface = detail::blend(lo->surface, hi->surface, alpha, "surface");
  return out;
}

}  // namespace gwm::ingest

// ---------------------------------------------------------------------------
// Prepared-case initialisation (prepared_case_init.hpp:16-134,
// src/ingest/prepared_case_init.cpp)
// ---------------------------------------------------------------------------
namespace gwm::ingest {

struct PreparedCaseInitConfig {
  int halo = 1;
  int ranks_x = 1;
  int ranks_y = 1;
  double terrain_taper_eta = 0.25;
  bool periodic_x = false;
  bool periodic_y = false;
};

struct PreparedCaseBalanceDiagnostics {
  real max_rel_eos = 0.0f;
  real max_rel_hydrostatic = 0.0f;
  real max_rel_fast_vertical = 0.0f;
  real max_abs_mass_divergence = 0.0f;
  real max_rel_mass_divergence = 0.0f;
  real max_abs_mom_w_bottom = 0.0f;
  real max_abs_mom_w_top = 0.0f;
  real max_abs_tracer_closure = 0.0f;
  std::optional<real> max_abs_z_src_minus_metric{};
};

namespace detail {

// Global (no halo) index maps of the analysis fields: x fastest.
struct GlobalIndex {
  int nx, ny;
  size_t c(int i, int j, int k) const {
    return (static_cast<size_t>(k) * ny + j) * nx + i;
  }
  size_t xf(int i, int j, int k) const {
    return (static_cast<size_t>(k) * ny + j) * (nx + 1) + i;
  }
  size_t yf(int i, int j, int k) const {
    return (static_cast<size_t>(k) * (ny + 1) + j) * nx + i;
  }
  size_t zf(int i, int j, int k) const {
    return (static_cast<size_t>(k) * ny + j) * nx + i;
  }
  size_t c2(int i, int j) const { return static_cast<size_t>(j) * nx + i; }
};

inline constexpr real kPreparedGravity = 9.81f;

inline int wrap_or_clamp(int i, int n, bool periodic) {
  if (!periodic) return std::clamp(i, 0, n - 1);
  const int m = i % n;
  return m < 0 ? m + n : m;
}

inline void require_metrics_compatible(const AnalysisStateIR& a, const domain::GridMetrics& m) {
  require(m.nx == a.grid.nx && m.ny == a.grid.ny && m.nz == a.grid.nz,
          "Prepared-case metrics must match analysis grid dimensions");
}

// rho*theta that reproduces pressure p through the EOS at rho = 1: a power
// law start and two corrections (prepared_case_init.cpp:38-50).
inline real rho_theta_m_from_pressure(real p) {
  using namespace gwm::core;
  real rt = (kReferencePressure / kDryGasConstant) *
            std::pow(std::max(p / kReferencePressure, 1.0e-6f), 1.0f - kKappa);
  for (int it = 0; it < 2; ++it) {
    const real pe = dry_pressure_from_rho_theta_m(1.0f, rt);
    const real ratio = std::max(p / std::max(pe, 1.0f), 1.0e-6f);
    rt *= std::pow(ratio, 1.0f - kKappa);
  }
  return rt;
}

struct BalancedColumns {
  std::vector<real> pressure, theta, rho_d, rho_theta_m;
};

// Hydrostatic columns (prepared_case_init.cpp:320-384): theta from the
// analysis T and p; the bottom level keeps the analysis pressure, each level
// above integrates dp = -g rho dz with a trapezoid, four fixed-point passes.
inline BalancedColumns balanced_columns(const AnalysisStateIR& a, const domain::GridMetrics& m) {
  using namespace gwm::core;
  require_metrics_compatible(a, m);
  const int nx = a.grid.nx, ny = a.grid.ny, nz = a.grid.nz;
  const GlobalIndex G{nx, ny};
  const auto& ps = a.atmosphere.values.at("air_pressure");
  const auto& ts = a.atmosphere.values.at("air_temperature");
  BalancedColumns b;
  const size_t n = a.grid.cell_count_3d();
  b.pressure.assign(n, 0.0f);
  b.theta.assign(n, 0.0f);
  b.rho_d.assign(n, 0.0f);
  b.rho_theta_m.assign(n, 0.0f);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const size_t c0 = G.c(i, j, 0);
      const real p0 = std::max(ps[c0], 1.0f);
      const real th0 = ts[c0] * std::pow(kReferencePressure / p0, kKappa);
      b.theta[c0] = th0;
      b.pressure[c0] = p0;
      b.rho_d[c0] = dry_rho_from_pressure_theta(p0, th0);
      b.rho_theta_m[c0] = rho_theta_m_from_pressure(p0);
      for (int k = 1; k < nz; ++k) {
        const size_t c = G.c(i, j, k), cb = G.c(i, j, k - 1);
        const real psrc = std::max(ps[c], 1.0f);
        const real thk = ts[c] * std::pow(kReferencePressure / psrc, kKappa);
        b.theta[c] = thk;
        const real dz = m.z_center(i, j, k) - m.z_center(i, j, k - 1);
        real pn = std::max(1.0f, b.pressure[cb] - kPreparedGravity * b.rho_d[cb] * dz);
        for (int it = 0; it < 4; ++it) {
          const real rn = dry_rho_from_pressure_theta(pn, thk);
          pn = b.pressure[cb] - kPreparedGravity * 0.5f * (b.rho_d[cb] + rn) * dz;
          pn = std::max(pn, 1.0f);
        }
        b.pressure[c] = pn;
        b.rho_d[c] = dry_rho_from_pressure_theta(pn, thk);
        b.rho_theta_m[c] = rho_theta_m_from_pressure(pn);
      }
    }
  return b;
}

struct FaceMomenta {
  std::vector<real> mom_u, mom_v, mom_w;
};

// Column mean of the horizontal mass-flux divergence weighted by dz.
inline std::vector<real> column_mean_divergence(const std::vector<real>& mu,
                                                const std::vector<real>& mv,
                                                const AnalysisStateIR& a,
                                                const domain::GridMetrics& m) {
  const int nx = a.grid.nx, ny = a.grid.ny, nz = a.grid.nz;
  const GlobalIndex G{nx, ny};
  const real dx = static_cast<real>(a.grid.dx), dy = static_cast<real>(a.grid.dy);
  std::vector<real> out(static_cast<size_t>(nx) * ny, 0.0f);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      real bias = 0.0f, height = 0.0f;
      for (int k = 0; k < nz; ++k) {
        const real hdiv = (mu[G.xf(i + 1, j, k)] - mu[G.xf(i, j, k)]) / dx +
                          (mv[G.yf(i, j + 1, k)] - mv[G.yf(i, j, k)]) / dy;
        const real dz = 1.0f / std::max(m.inv_dz_cell(i, j, k), 1.0e-6f);
        bias += hdiv * dz;
        height += dz;
      }
      out[G.c2(i, j)] = bias / std::max(height, 1.0e-6f);
    }
  return out;
}

// Potential phi with lap(phi) = rhs, phi = 0 outside the domain, by
// successive over-relaxation (omega 1.85, at most 2000 sweeps, stop when no
// update exceeds 1e-10) (prepared_case_init.cpp:126-170).
inline std::vector<real> barotropic_potential(const std::vector<real>& rhs, int nx, int ny,
                                              real dx, real dy) {
  std::vector<real> phi(static_cast<size_t>(nx) * ny, 0.0f);
  if (nx == 0 || ny == 0) return phi;
  const real ix2 = 1.0f / std::max(dx * dx, 1.0e-6f), iy2 = 1.0f / std::max(dy * dy, 1.0e-6f);
  const real denom = 2.0f * (ix2 + iy2);
  if (denom <= 0.0f) return phi;
  const GlobalIndex G{nx, ny};
  for (int it = 0; it < 2000; ++it) {
    real biggest = 0.0f;
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const real w = i > 0 ? phi[G.c2(i - 1, j)] : 0.0f;
        const real e = i + 1 < nx ? phi[G.c2(i + 1, j)] : 0.0f;
        const real s = j > 0 ? phi[G.c2(i, j - 1)] : 0.0f;
        const real n = j + 1 < ny ? phi[G.c2(i, j + 1)] : 0.0f;
        const size_t c = G.c2(i, j);
        const real star = ((w + e) * ix2 + (s + n) * iy2 - rhs[c]) / denom;
        const real upd = phi[c] + 1.85f * (star - phi[c]);
        biggest = std::max(biggest, std::fabs(upd - phi[c]));
        phi[c] = upd;
      }
    if (biggest <= 1.0e-10f) break;
  }
  return phi;
}

// Face momenta from the balanced density and the analysis winds
// (prepared_case_init.cpp:172-318): rho and the wind averaged to each face
// (neighbours wrapped on periodic axes, clamped on open ones); on fully open
// domains up to three barotropic projections remove the column-mean
// horizontal divergence; w from the column continuity equation with the
// column-mean divergence removed, zero at the ground and the lid.
inline FaceMomenta projected_face_momenta(const AnalysisStateIR& a, const domain::GridMetrics& m,
                                          const BalancedColumns& b) {
  const int nx = a.grid.nx, ny = a.grid.ny, nz = a.grid.nz;
  const GlobalIndex G{nx, ny};
  const auto& uw = a.atmosphere.values.at("u_wind");
  const auto& vw = a.atmosphere.values.at("v_wind");
  const real dx = static_cast<real>(a.grid.dx), dy = static_cast<real>(a.grid.dy);
  FaceMomenta f;
  f.mom_u.assign(static_cast<size_t>(nx + 1) * ny * nz, 0.0f);
  f.mom_v.assign(static_cast<size_t>(nx) * (ny + 1) * nz, 0.0f);
  f.mom_w.assign(static_cast<size_t>(nx) * ny * (nz + 1), 0.0f);
  for (int j = 0; j < ny; ++j)
    for (int fi = 0; fi <= nx; ++fi) {
      const int ir = wrap_or_clamp(fi, nx, m.periodic_x), il = wrap_or_clamp(fi - 1, nx, m.periodic_x);
      for (int k = 0; k < nz; ++k) {
        const size_t l = G.c(il, j, k), r = G.c(ir, j, k);
        f.mom_u[G.xf(fi, j, k)] = 0.5f * (b.rho_d[l] + b.rho_d[r]) * (0.5f * (uw[l] + uw[r]));
      }
    }
  for (int fj = 0; fj <= ny; ++fj) {
    const int jn = wrap_or_clamp(fj, ny, m.periodic_y), js = wrap_or_clamp(fj - 1, ny, m.periodic_y);
    for (int i = 0; i < nx; ++i)
      for (int k = 0; k < nz; ++k) {
        const size_t s = G.c(i, js, k), n = G.c(i, jn, k);
        f.mom_v[G.yf(i, fj, k)] = 0.5f * (b.rho_d[s] + b.rho_d[n]) * (0.5f * (vw[s] + vw[n]));
      }
  }
  if (!m.periodic_x && !m.periodic_y) {
    for (int pass = 0; pass < 3; ++pass) {
      const auto mean = column_mean_divergence(f.mom_u, f.mom_v, a, m);
      real worst = 0.0f;
      for (real v : mean) worst = std::max(worst, std::fabs(v));
      if (worst <= 1.0e-8f) break;
      const auto phi = barotropic_potential(mean, nx, ny, dx, dy);
      for (int j = 0; j < ny; ++j)
        for (int fi = 0; fi <= nx; ++fi) {
          const real pl = fi > 0 ? phi[G.c2(fi - 1, j)] : 0.0f;
          const real pr = fi < nx ? phi[G.c2(fi, j)] : 0.0f;
          const real corr = (pr - pl) / dx;
          for (int k = 0; k < nz; ++k) f.mom_u[G.xf(fi, j, k)] -= corr;
        }
      for (int fj = 0; fj <= ny; ++fj)
        for (int i = 0; i < nx; ++i) {
          const real ps = fj > 0 ? phi[G.c2(i, fj - 1)] : 0.0f;
          const real pn = fj < ny ? phi[G.c2(i, fj)] : 0.0f;
          const real corr = (pn - ps) / dy;
          for (int k = 0; k < nz; ++k) f.mom_v[G.yf(i, fj, k)] -= corr;
        }
    }
  }
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      auto hdiv = [&](int k) {
        return (f.mom_u[G.xf(i + 1, j, k)] - f.mom_u[G.xf(i, j, k)]) / dx +
               (f.mom_v[G.yf(i, j + 1, k)] - f.mom_v[G.yf(i, j, k)]) / dy;
      };
      auto dz_of = [&](int k) { return 1.0f / std::max(m.inv_dz_cell(i, j, k), 1.0e-6f); };
      f.mom_w[G.zf(i, j, 0)] = 0.0f;
      real bias = 0.0f, height = 0.0f;
      for (int k = 0; k < nz; ++k) {
        const real dz = dz_of(k);
        bias += hdiv(k) * dz;
        height += dz;
      }
      const real mean = bias / std::max(height, 1.0e-6f);
      for (int k = 0; k < nz; ++k)
        f.mom_w[G.zf(i, j, k + 1)] = f.mom_w[G.zf(i, j, k)] - (hdiv(k) - mean) * dz_of(k);
      f.mom_w[G.zf(i, j, nz)] = 0.0f;
    }
  return f;
}

inline AnalysisStateIR boundary_analysis(const AnalysisStateIR& a, const BoundarySnapshotIR& s) {
  AnalysisStateIR out = a;
  out.valid_time_utc = s.valid_time_utc;
  out.forecast_offset_seconds = s.forecast_offset_seconds;
  out.atmosphere = s.atmosphere;
  out.surface = s.surface;
  return out;
}

// Balanced fields of the boundary snapshot at the current forecast offset,
// rebuilt only when the offset (rounded to a second) changes
// (prepared_case_init.cpp:396-422).
struct BoundaryProjectionCache {
  int offset = std::numeric_limits<int>::min();
  BoundarySnapshotIR snapshot;
  BalancedColumns columns;
  FaceMomenta faces;

  void refresh(const AnalysisStateIR& a, const BoundaryCacheIR& c, const domain::GridMetrics& m,
               real step_start, real sim_time) {
    const int t = static_cast<int>(std::lround(step_start + sim_time));
    if (t == offset) return;
    snapshot = interpolate_boundary_snapshot(c, t);
    const auto ba = boundary_analysis(a, snapshot);
    columns = balanced_columns(ba, m);
    faces = projected_face_momenta(ba, m, columns);
    offset = t;
  }
};

// Host metrics (terrain-following heights and layer thicknesses) without a
// device grid, for the boundary projections of the updaters.
inline domain::GridMetrics host_metrics_like(const domain::GridMetrics& m) {
  domain::GridMetrics h = m;
  h.device.reset();
  return h;
}

}  // namespace detail

inline std::vector<domain::SubdomainDescriptor> make_prepared_case_layout(
    const AnalysisStateIR& a, const PreparedCaseInitConfig& c) {
  validate_runtime_analysis(a);
  require(c.halo > 0, "Prepared-case halo must be positive");
  require(c.ranks_x > 0 && c.ranks_y > 0, "Prepared-case rank counts must be positive");
  return comm::VirtualRankLayout::build(a.grid.nx, a.grid.ny, a.grid.nz, c.halo, c.ranks_x,
                                        c.ranks_y, c.periodic_x, c.periodic_y);
}

inline domain::GridMetrics make_prepared_case_metrics(const AnalysisStateIR& a,
                                                      const PreparedCaseInitConfig& c) {
  validate_runtime_analysis(a);
  const auto it = a.static_surface.values.find("terrain_height");
  require(it != a.static_surface.values.end(), "Prepared analysis state requires terrain_height");
  return domain::GridMetrics::make_hybrid_height(a.grid.nx, a.grid.ny, a.grid.nz, a.grid.dx,
                                                 a.grid.dy, a.grid.z_top, it->second,
                                                 c.terrain_taper_eta, c.periodic_x, c.periodic_y,
                                                 {}, c.halo);
}

inline std::vector<dycore::DryState> make_dry_states_from_analysis(
    const AnalysisStateIR& a, const domain::GridMetrics& m,
    const std::vector<domain::SubdomainDescriptor>& layout,
    const std::string& label_prefix = "prepared_case") {
  validate_runtime_analysis(a);
  require(!layout.empty(), "Prepared-case layout must not be empty");
  detail::require_metrics_compatible(a, m);
  const auto b = detail::balanced_columns(a, m);
  const auto f = detail::projected_face_momenta(a, m, b);
  const detail::GlobalIndex G{a.grid.nx, a.grid.ny};
  std::vector<dycore::DryState> states;
  states.reserve(layout.size());
  for (const auto& d : layout) {
    dycore::DryState s(d.nx_local(), d.ny_local(), d.nz, d.halo,
                       label_prefix + "_rank_" + std::to_string(d.rank));
    s.fill_constant(0.0f, 300.0f, 0.0f, 0.0f, 0.0f);
    for (int k = 0; k < d.nz; ++k)
      for (int j = 0; j < d.ny_local(); ++j)
        for (int i = 0; i < d.nx_local(); ++i) {
          const size_t c = G.c(d.i_begin + i, d.j_begin + j, k);
          s.rho_d(i, j, k) = b.rho_d[c];
          s.rho_theta_m(i, j, k) = b.rho_theta_m[c];
        }
    for (int k = 0; k < d.nz; ++k)
      for (int j = 0; j < d.ny_local(); ++j)
        for (int fi = 0; fi <= d.nx_local(); ++fi)
          s.mom_u.storage()(fi, j, k) = f.mom_u[G.xf(d.i_begin + fi, d.j_begin + j, k)];
    for (int k = 0; k < d.nz; ++k)
      for (int fj = 0; fj <= d.ny_local(); ++fj)
        for (int i = 0; i < d.nx_local(); ++i)
          s.mom_v.storage()(i, fj, k) = f.mom_v[G.yf(d.i_begin + i, d.j_begin + fj, k)];
    for (int fk = 0; fk <= d.nz; ++fk)
      for (int j = 0; j < d.ny_local(); ++j)
        for (int i = 0; i < d.nx_local(); ++i)
          s.mom_w.storage()(i, j, fk) = f.mom_w[G.zf(d.i_begin + i, d.j_begin + j, fk)];
    states.push_back(std::move(s));
  }
  return states;
}

// Tracer masses rho_d * q from the analysis mixing ratios of the registry's
// tracers (specific humidity required; other tracers when present)
// (prepared_case_init.cpp:923-1007).
inline std::vector<state::TracerState> make_tracers_from_analysis(
    const AnalysisStateIR& a, const std::vector<dycore::DryState>& dry,
    const std::vector<domain::SubdomainDescriptor>& layout, const state::TracerRegistry& registry,
    const std::string& label_prefix = "prepared_case_tracer",
    const std::vector<state::TracerState>* baseline = nullptr) {
  validate_runtime_analysis(a);
  require(dry.size() == layout.size(), "Dry-state/layout size mismatch in make_tracers_from_analysis");
  if (baseline)
    require(baseline->size() == layout.size(),
            "Baseline tracer/layout size mismatch in make_tracers_from_analysis");
  const detail::GlobalIndex G{a.grid.nx, a.grid.ny};
  std::vector<state::TracerState> out;
  out.reserve(layout.size());
  for (size_t r = 0; r < layout.size(); ++r) {
    const auto& d = layout[r];
    state::TracerState t(registry, d.nx_local(), d.ny_local(), d.nz, d.halo,
                         label_prefix + "_rank_" + std::to_string(d.rank));
    t.fill_zero();
    if (baseline)
      for (size_t q = 0; q < t.size(); ++q)
        t.mass(static_cast<int>(q)).copy_all_from((*baseline)[r].mass(static_cast<int>(q)));
    out.push_back(std::move(t));
  }
  for (size_t q = 0; q < registry.size(); ++q) {
    const auto& name = registry.at(static_cast<int>(q)).name;
    const std::vector<real>* ratio = nullptr;
    if (name == state::kSpecificHumidityTracerName) {
      ratio = &a.atmosphere.values.at(state::kSpecificHumidityTracerName);
      require(ratio->size() == a.grid.cell_count_3d(),
              "specific_humidity analysis field size mismatch");
    } else {
      const auto it = a.atmosphere.values.find(name);
      if (it == a.atmosphere.values.end()) continue;
      ratio = &it->second;
      require(ratio->size() == a.grid.cell_count_3d(),
              "Tracer analysis field size mismatch for " + name);
    }
    for (size_t r = 0; r < layout.size(); ++r) {
      const auto& d = layout[r];
      auto& rq = out[r].mass(name);
      for (int k = 0; k < d.nz; ++k)
        for (int j = 0; j < d.ny_local(); ++j)
          for (int i = 0; i < d.nx_local(); ++i)
            rq(i, j, k) = dry[r].rho_d(i, j, k) * (*ratio)[G.c(d.i_begin + i, d.j_begin + j, k)];
    }
  }
  return out;
}

inline std::vector<state::TracerState> make_specific_humidity_tracers_from_analysis(
    const AnalysisStateIR& a, const std::vector<dycore::DryState>& dry,
    const std::vector<domain::SubdomainDescriptor>& layout,
    const std::string& label_prefix = "prepared_case_qv") {
  return make_tracers_from_analysis(a, dry, layout, state::make_specific_humidity_registry(),
                                    label_prefix);
}

inline std::vector<state::TracerState> make_warm_rain_tracers_from_analysis(
    const AnalysisStateIR& a, const std::vector<dycore::DryState>& dry,
    const std::vector<domain::SubdomainDescriptor>& layout,
    const std::string& label_prefix = "prepared_case_warm_rain") {
  return make_tracers_from_analysis(a, dry, layout, state::make_warm_rain_registry(),
                                    label_prefix);
}

// Start-up balance of a state against the balanced columns of its analysis
// (prepared_case_init.cpp:746-908): EOS and hydrostatic residuals, the
// fast-mode vertical perturbations, the discrete mass divergence, w at the
// ground and lid, the tracer closure, and the source heights against the
// metric heights when the analysis carries geopotential_height.
inline PreparedCaseBalanceDiagnostics diagnose_prepared_case_balance(
    const AnalysisStateIR& a, const domain::GridMetrics& m,
    const std::vector<dycore::DryState>& dry,
    const std::vector<domain::SubdomainDescriptor>& layout,
    const std::vector<state::TracerState>* tracers = nullptr) {
  using namespace gwm::core;
  validate_runtime_analysis(a);
  detail::require_metrics_compatible(a, m);
  require(dry.size() == layout.size(),
          "Dry-state/layout size mismatch in diagnose_prepared_case_balance");
  if (tracers)
    require(tracers->size() == layout.size(),
            "Tracer/layout size mismatch in diagnose_prepared_case_balance");
  const auto b = detail::balanced_columns(a, m);
  const detail::GlobalIndex G{a.grid.nx, a.grid.ny};
  PreparedCaseBalanceDiagnostics out;
  const size_t n = a.grid.cell_count_3d();
  std::vector<real> rho(n, 0.0f), theta(n, 0.0f), p(n, 0.0f);
  const real dx = static_cast<real>(a.grid.dx), dy = static_cast<real>(a.grid.dy);
  for (size_t r = 0; r < layout.size(); ++r) {
    const auto& d = layout[r];
    const auto& s = dry[r];
    for (int j = 0; j < d.ny_local(); ++j)
      for (int i = 0; i < d.nx_local(); ++i) {
        const int gi = d.i_begin + i, gj = d.j_begin + j;
        for (int k = 0; k < d.nz; ++k) {
          const size_t c = G.c(gi, gj, k);
          rho[c] = s.rho_d(i, j, k);
          theta[c] = dry_theta_from_conserved(s.rho_d(i, j, k), s.rho_theta_m(i, j, k));
          p[c] = dry_pressure_from_rho_theta_m(s.rho_d(i, j, k), s.rho_theta_m(i, j, k));
          const real ddx = (s.mom_u.storage()(i + 1, j, k) - s.mom_u.storage()(i, j, k)) / dx;
          const real ddy = (s.mom_v.storage()(i, j + 1, k) - s.mom_v.storage()(i, j, k)) / dy;
          const real ddz = (s.mom_w.storage()(i, j, k + 1) - s.mom_w.storage()(i, j, k)) *
                           m.inv_dz_cell(gi, gj, k);
          const real div = ddx + ddy + ddz;
          const real scale =
              std::max({std::fabs(ddx), std::fabs(ddy), std::fabs(ddz), 1.0e-12f});
          out.max_abs_mass_divergence = std::max(out.max_abs_mass_divergence, std::fabs(div));
          out.max_rel_mass_divergence =
              std::max(out.max_rel_mass_divergence, std::fabs(div) / scale);
        }
        out.max_abs_mom_w_bottom =
            std::max(out.max_abs_mom_w_bottom, std::fabs(s.mom_w.storage()(i, j, 0)));
        out.max_abs_mom_w_top =
            std::max(out.max_abs_mom_w_top, std::fabs(s.mom_w.storage()(i, j, d.nz)));
      }
  }
  const auto zsrc = a.atmosphere.values.find("geopotential_height");
  for (int j = 0; j < a.grid.ny; ++j)
    for (int i = 0; i < a.grid.nx; ++i) {
      for (int k = 0; k < a.grid.nz; ++k) {
        const size_t c = G.c(i, j, k);
        const real rexp = dry_rho_from_pressure_theta(b.pressure[c], theta[c]);
        out.max_rel_eos =
            std::max(out.max_rel_eos, std::fabs(rho[c] - rexp) / std::max(std::fabs(rexp), 1.0f));
        out.max_rel_hydrostatic =
            std::max(out.max_rel_hydrostatic,
                     std::fabs(p[c] - b.pressure[c]) / std::max(std::fabs(b.pressure[c]), 1.0f));
        if (zsrc != a.atmosphere.values.end()) {
          const real mis = std::fabs(zsrc->second[c] - m.z_center(i, j, k));
          out.max_abs_z_src_minus_metric =
              out.max_abs_z_src_minus_metric ? std::max(*out.max_abs_z_src_minus_metric, mis) : mis;
        }
      }
      for (int fk = 1; fk < a.grid.nz; ++fk) {
        const size_t lo = G.c(i, j, fk - 1), hi = G.c(i, j, fk);
        const real ppl = p[lo] - b.pressure[lo], ppu = p[hi] - b.pressure[hi];
        const real rpf = 0.5f * ((rho[lo] - b.rho_d[lo]) + (rho[hi] - b.rho_d[hi]));
        const real rrf = 0.5f * (b.rho_d[lo] + b.rho_d[hi]);
        const real scale = std::max({std::fabs(ppl) / std::max(std::fabs(b.pressure[lo]), 1.0f),
                                     std::fabs(ppu) / std::max(std::fabs(b.pressure[hi]), 1.0f),
                                     std::fabs(rpf) / std::max(std::fabs(rrf), 1.0f), 0.0f});
        out.max_rel_fast_vertical = std::max(out.max_rel_fast_vertical, scale);
      }
    }
  if (tracers) {
    for (size_t r = 0; r < layout.size(); ++r) {
      const auto& d = layout[r];
      const auto& t = (*tracers)[r];
      for (size_t q = 0; q < t.size(); ++q) {
        const auto it = a.atmosphere.values.find(t.registry().at(static_cast<int>(q)).name);
        if (it == a.atmosphere.values.end()) continue;
        const auto& rq = t.mass(static_cast<int>(q));
        for (int j = 0; j < d.ny_local(); ++j)
          for (int i = 0; i < d.nx_local(); ++i)
            for (int k = 0; k < d.nz; ++k) {
              const real expect =
                  dry[r].rho_d(i, j, k) * it->second[G.c(d.i_begin + i, d.j_begin + j, k)];
              out.max_abs_tracer_closure =
                  std::max(out.max_abs_tracer_closure, std::fabs(rq(i, j, k) - expect));
            }
      }
    }
  }
  return out;
}

// Lateral-strip overwrite from the balanced boundary snapshot: on each open
// side touched by a rank, the outermost owned cells (rho, rho*theta), the
// outermost normal faces and every tangential and vertical face of that
// strip (prepared_case_init.cpp:424-565).
inline void apply_prepared_dry_strips(std::vector<dycore::DryState>& states,
                                      const std::vector<domain::SubdomainDescriptor>& layout,
                                      int gnx, int gny, const detail::BalancedColumns& b,
                                      const detail::FaceMomenta& f) {
  require(states.size() == layout.size(),
          "State/layout size mismatch in apply_cached_dry_boundary_strips");
  const detail::GlobalIndex G{gnx, gny};
  for (size_t r = 0; r < layout.size(); ++r) {
    auto& s = states[r];
    const auto& d = layout[r];
    // Column strip at local i (global gi): cells, the x-face xi, y-faces, z-faces.
    auto column_strip = [&](int i, int xi) {
      const int gi = d.i_begin + i;
      for (int j = 0; j < d.ny_local(); ++j)
        for (int k = 0; k < d.nz; ++k) {
          const size_t c = G.c(gi, d.j_begin + j, k);
          s.rho_d(i, j, k) = b.rho_d[c];
          s.rho_theta_m(i, j, k) = b.rho_theta_m[c];
          s.mom_u.storage()(xi, j, k) = f.mom_u[G.xf(d.i_begin + xi, d.j_begin + j, k)];
        }
      for (int fj = 0; fj <= d.ny_local(); ++fj)
        for (int k = 0; k < d.nz; ++k)
          s.mom_v.storage()(i, fj, k) = f.mom_v[G.yf(gi, d.j_begin + fj, k)];
      for (int j = 0; j < d.ny_local(); ++j)
        for (int fk = 0; fk <= d.nz; ++fk)
          s.mom_w.storage()(i, j, fk) = f.mom_w[G.zf(gi, d.j_begin + j, fk)];
    };
    auto row_strip = [&](int j, int yj) {
      const int gj = d.j_begin + j;
      for (int i = 0; i < d.nx_local(); ++i)
        for (int k = 0; k < d.nz; ++k) {
          const size_t c = G.c(d.i_begin + i, gj, k);
          s.rho_d(i, j, k) = b.rho_d[c];
          s.rho_theta_m(i, j, k) = b.rho_theta_m[c];
          s.mom_v.storage()(i, yj, k) = f.mom_v[G.yf(d.i_begin + i, d.j_begin + yj, k)];
        }
      for (int fi = 0; fi <= d.nx_local(); ++fi)
        for (int k = 0; k < d.nz; ++k)
          s.mom_u.storage()(fi, j, k) = f.mom_u[G.xf(d.i_begin + fi, gj, k)];
      for (int i = 0; i < d.nx_local(); ++i)
        for (int fk = 0; fk <= d.nz; ++fk)
          s.mom_w.storage()(i, j, fk) = f.mom_w[G.zf(d.i_begin + i, gj, fk)];
    };
    if (dycore::touches_west_boundary(d)) column_strip(0, 0);
    if (dycore::touches_east_boundary(d)) column_strip(d.nx_local() - 1, d.nx_local());
    if (dycore::touches_south_boundary(d)) row_strip(0, 0);
    if (dycore::touches_north_boundary(d)) row_strip(d.ny_local() - 1, d.ny_local());
  }
}

// Tracer strips: rho_d (balanced boundary) times the snapshot's mixing ratio,
// for the tracers the snapshot carries (prepared_case_init.cpp:567-638).
inline void apply_prepared_tracer_strips(std::vector<state::TracerState>& states,
                                         const std::vector<domain::SubdomainDescriptor>& layout,
                                         const BoundarySnapshotIR& snap, int gnx, int gny,
                                         const std::vector<real>& rho_d) {
  require(states.size() == layout.size(),
          "Tracer/layout size mismatch in apply_cached_tracer_boundary_strips");
  const detail::GlobalIndex G{gnx, gny};
  for (size_t r = 0; r < layout.size(); ++r) {
    auto& s = states[r];
    const auto& d = layout[r];
    for (size_t q = 0; q < s.size(); ++q) {
      const auto it = snap.atmosphere.values.find(s.registry().at(static_cast<int>(q)).name);
      if (it == snap.atmosphere.values.end()) continue;
      const auto& ratio = it->second;
      auto& rq = s.mass(static_cast<int>(q));
      auto cell = [&](int i, int j) {
        for (int k = 0; k < d.nz; ++k) {
          const size_t c = G.c(d.i_begin + i, d.j_begin + j, k);
          rq(i, j, k) = rho_d[c] * ratio[c];
        }
      };
      if (dycore::touches_west_boundary(d))
        for (int j = 0; j < d.ny_local(); ++j) cell(0, j);
      if (dycore::touches_east_boundary(d))
        for (int j = 0; j < d.ny_local(); ++j) cell(d.nx_local() - 1, j);
      if (dycore::touches_south_boundary(d))
        for (int i = 0; i < d.nx_local(); ++i) cell(i, 0);
      if (dycore::touches_north_boundary(d))
        for (int i = 0; i < d.nx_local(); ++i) cell(i, d.ny_local() - 1);
    }
  }
}

class PreparedCaseBoundaryUpdater final : public dycore::BoundaryUpdater {
 public:
  PreparedCaseBoundaryUpdater(AnalysisStateIR analysis, BoundaryCacheIR cache,
                              PreparedCaseInitConfig config)
      : analysis_(std::move(analysis)), cache_(std::move(cache)), config_(config),
        metrics_(make_prepared_case_metrics(analysis_, config_)) {}
  void set_step_start_time(real t) { step_start_ = std::max(t, 0.0f); }
  void apply(std::vector<dycore::DryState>& states,
             const std::vector<domain::SubdomainDescriptor>& layout, real sim_time) override {
    if (cache_.empty()) return;
    proj_.refresh(analysis_, cache_, metrics_, step_start_, sim_time);
    apply_prepared_dry_strips(states, layout, analysis_.grid.nx, analysis_.grid.ny, proj_.columns,
                              proj_.faces);
  }

 private:
  AnalysisStateIR analysis_;
  BoundaryCacheIR cache_;
  PreparedCaseInitConfig config_{};
  domain::GridMetrics metrics_{};
  real step_start_ = 0.0f;
  detail::BoundaryProjectionCache proj_;
};

class PreparedCaseTracerBoundaryUpdater final : public dycore::TracerBoundaryUpdater {
 public:
  PreparedCaseTracerBoundaryUpdater(AnalysisStateIR analysis, BoundaryCacheIR cache,
                                    PreparedCaseInitConfig config)
      : analysis_(std::move(analysis)), cache_(std::move(cache)), config_(config),
        metrics_(make_prepared_case_metrics(analysis_, config_)) {}
  void set_step_start_time(real t) { step_start_ = std::max(t, 0.0f); }
  void apply(std::vector<state::TracerState>& states,
             const std::vector<domain::SubdomainDescriptor>& layout, real sim_time) override {
    if (cache_.empty() || states.empty()) return;
    proj_.refresh(analysis_, cache_, metrics_, step_start_, sim_time);
    apply_prepared_tracer_strips(states, layout, proj_.snapshot, analysis_.grid.nx,
                                 analysis_.grid.ny, proj_.columns.rho_d);
  }

 private:
  AnalysisStateIR analysis_;
  BoundaryCacheIR cache_;
  PreparedCaseInitConfig config_{};
  domain::GridMetrics metrics_{};
  real step_start_ = 0.0f;
  detail::BoundaryProjectionCache proj_;
};

}  // namespace gwm::ingest

// ---------------------------------------------------------------------------
// Runtime summary (include/gwm/core/runtime_summary.hpp:12-55,
// src/core/runtime_summary.cpp)
// ---------------------------------------------------------------------------
namespace gwm::core {

struct TracerSummary {
  std::string name, units;
  bool positive = true;
  double column_integrated_sum_kg_m2 = 0.0;
  dycore::ScalarRange mixing_ratio;
};

struct MoistureBudgetSummary {
  double vapor_water_path_sum_kg_m2 = 0.0;
  double cloud_water_path_sum_kg_m2 = 0.0;
  double rain_water_path_sum_kg_m2 = 0.0;
  double condensed_water_path_sum_kg_m2 = 0.0;
  double total_water_path_sum_kg_m2 = 0.0;
  double accumulated_surface_precipitation_sum_mm = 0.0;
  double mean_surface_precipitation_mm = 0.0;
  double max_surface_precipitation_mm = 0.0;
  int total_surface_cell_count = 0;
  int precipitating_surface_cell_count = 0;
  double precipitating_surface_fraction = 0.0;
  double mean_precipitating_surface_precipitation_mm = 0.0;
};

struct RuntimeStateSummary {
  dycore::DryStateSummary dry;
  std::vector<TracerSummary> tracers;
  MoistureBudgetSummary moisture;
};

inline RuntimeStateSummary summarize_runtime_state(
    const std::vector<dycore::DryState>& states, const std::vector<state::TracerState>& tracers,
    const std::vector<domain::SubdomainDescriptor>& layout, const domain::GridMetrics& metrics,
    const std::vector<physics::WarmRainSurfaceAccumulation>* accumulations = nullptr) {
  require(states.size() == tracers.size(), "State/tracer size mismatch in summarize_runtime_state");
  require(states.size() == layout.size(), "State/layout size mismatch in summarize_runtime_state");
  RuntimeStateSummary out;
  out.dry = dycore::summarize_dry_states(states);
  if (states.empty() || tracers.empty()) return out;
  const auto& reg = tracers.front().registry();
  for (size_t q = 0; q < reg.size(); ++q) {
    const auto& spec = reg.at(static_cast<int>(q));
    TracerSummary t;
    t.name = spec.name;
    t.units = spec.units;
    t.positive = spec.positive;
    t.mixing_ratio = {std::numeric_limits<double>::max(), -std::numeric_limits<double>::max()};
    out.tracers.push_back(std::move(t));
  }
  bool any = false;
  for (size_t r = 0; r < states.size(); ++r) {
    require(tracers[r].size() == out.tracers.size(),
            "Tracer registry mismatch in summarize_runtime_state");
    const auto& d = layout[r];
    for (size_t q = 0; q < out.tracers.size(); ++q) {
      require(reg.at(static_cast<int>(q)).name == tracers[r].registry().at(static_cast<int>(q)).name,
              "Tracer ordering mismatch in summarize_runtime_state");
      const auto& rq = tracers[r].mass(static_cast<int>(q));
      for (int k = 0; k < d.nz; ++k)
        for (int j = 0; j < d.ny_local(); ++j)
          for (int i = 0; i < d.nx_local(); ++i)
            out.tracers[q].column_integrated_sum_kg_m2 +=
                static_cast<double>(rq(i, j, k)) *
                (1.0 / static_cast<double>(metrics.inv_dz_cell(d.i_begin + i, d.j_begin + j, k)));
    }
    const auto& rho = states[r].rho_d;
    for (int k = 0; k < rho.nz(); ++k)
      for (int j = 0; j < rho.ny(); ++j)
        for (int i = 0; i < rho.nx(); ++i) {
          any = true;
          const double rr = std::max(static_cast<double>(rho(i, j, k)), 1.0e-12);
          for (size_t q = 0; q < out.tracers.size(); ++q) {
            const double v =
                static_cast<double>(tracers[r].mass(static_cast<int>(q))(i, j, k)) / rr;
            auto& mr = out.tracers[q].mixing_ratio;
            mr.min = std::min(mr.min, v);
            mr.max = std::max(mr.max, v);
          }
        }
  }
  if (!any)
    for (auto& t : out.tracers) t.mixing_ratio = {0.0, 0.0};
  auto path_of = [&out](const char* name) {
    for (const auto& t : out.tracers)
      if (t.name == name) return t.column_integrated_sum_kg_m2;
    return 0.0;
  };
  auto& mo = out.moisture;
  mo.vapor_water_path_sum_kg_m2 = path_of(state::kSpecificHumidityTracerName);
  mo.cloud_water_path_sum_kg_m2 = path_of(state::kCloudWaterTracerName);
  mo.rain_water_path_sum_kg_m2 = path_of(state::kRainWaterTracerName);
  mo.condensed_water_path_sum_kg_m2 = mo.cloud_water_path_sum_kg_m2 + mo.rain_water_path_sum_kg_m2;
  mo.total_water_path_sum_kg_m2 = mo.vapor_water_path_sum_kg_m2 + mo.condensed_water_path_sum_kg_m2;
  if (accumulations) {
    const auto w = physics::summarize_warm_rain(tracers, layout, metrics, accumulations);
    mo.accumulated_surface_precipitation_sum_mm = w.accumulated_surface_precipitation_sum_mm;
    mo.mean_surface_precipitation_mm = w.mean_surface_precipitation_mm;
    mo.max_surface_precipitation_mm = w.max_surface_precipitation_mm;
    double wet = 0.0;
    for (const auto& acc : *accumulations)
      for (int j = 0; j < acc.ny; ++j)
        for (int i = 0; i < acc.nx; ++i) {
          ++mo.total_surface_cell_count;
          const double v = acc.liquid_precipitation_kg_m2(i, j, 0);
          if (v > 0.0) {
            ++mo.precipitating_surface_cell_count;
            wet += v;
          }
        }
    if (mo.total_surface_cell_count > 0)
      mo.precipitating_surface_fraction = static_cast<double>(mo.precipitating_surface_cell_count) /
                                          static_cast<double>(mo.total_surface_cell_count);
    if (mo.precipitating_surface_cell_count > 0)
      mo.mean_precipitating_surface_precipitation_mm =
          wet / static_cast<double>(mo.precipitating_surface_cell_count);
  }
  return out;
}

inline std::string runtime_state_summary_to_json(const RuntimeStateSummary& s,
                                                 const std::string& in = "  ") {
  std::ostringstream o;
  o << std::setprecision(17);
  const auto& d = s.dry;
  auto range = [&](const char* key, const dycore::ScalarRange& r) {
    o << in << "\"" << key << "\": {\"min\": " << r.min << ", \"max\": " << r.max << "},\n";
  };
  o << "{\n";
  o << in << "\"total_dry_mass\": " << d.total_dry_mass << ",\n";
  o << in << "\"total_rho_theta_m\": " << d.total_rho_theta_m << ",\n";
  o << in << "\"total_horizontal_momentum_x\": " << d.total_horizontal_momentum_x << ",\n";
  o << in << "\"total_horizontal_momentum_y\": " << d.total_horizontal_momentum_y << ",\n";
  o << in << "\"total_vertical_momentum\": " << d.total_vertical_momentum << ",\n";
  range("rho_d", d.rho_d);
  range("theta_m", d.theta_m);
  range("u_face", d.u_face);
  range("v_face", d.v_face);
  range("w_face", d.w_face);
  const auto& m = s.moisture;
  o << in << "\"moisture\": {\n";
  o << in << "  \"vapor_water_path_sum_kg_m2\": " << m.vapor_water_path_sum_kg_m2 << ",\n";
  o << in << "  \"cloud_water_path_sum_kg_m2\": " << m.cloud_water_path_sum_kg_m2 << ",\n";
  o << in << "  \"rain_water_path_sum_kg_m2\": " << m.rain_water_path_sum_kg_m2 << ",\n";
  o << in << "  \"condensed_water_path_sum_kg_m2\": " << m.condensed_water_path_sum_kg_m2 << ",\n";
  o << in << "  \"total_water_path_sum_kg_m2\": " << m.total_water_path_sum_kg_m2 << ",\n";
  o << in << "  \"accumulated_surface_precipitation_sum_mm\": "
    << m.accumulated_surface_precipitation_sum_mm << ",\n";
  o << in << "  \"mean_surface_precipitation_mm\": " << m.mean_surface_precipitation_mm << ",\n";
  o << in << "  \"max_surface_precipitation_mm\": " << m.max_surface_precipitation_mm << ",\n";
  o << in << "  \"total_surface_cell_count\": " << m.total_surface_cell_count << ",\n";
  o << in << "  \"precipitating_surface_cell_count\": " << m.precipitating_surface_cell_count
    << ",\n";
  o << in << "  \"precipitating_surface_fraction\": " << m.precipitating_surface_fraction << ",\n";
  o << in << "  \"mean_precipitating_surface_precipitation_mm\": "
    << m.mean_precipitating_surface_precipitation_mm << "\n";
  o << in << "},\n";
  o << in << "\"tracers\": {\n";
  for (size_t q = 0; q < s.tracers.size(); ++q) {
    const auto& t = s.tracers[q];
    o << in << "  \"" << t.name << "\": {\n";
    o << in << "    \"units\": \"" << t.units << "\",\n";
    o << in << "    \"positive\": " << (t.positive ? "true" : "false") << ",\n";
    o << in << "    \"column_integrated_sum_kg_m2\": " << t.column_integrated_sum_kg_m2 << ",\n";
    o << in << "    \"mixing_ratio\": {\"min\": " << t.mixing_ratio.min
      << ", \"max\": " << t.mixing_ratio.max << "}\n";
    o << in << "  }" << (q + 1 != s.tracers.size() ? "," : "") << "\n";
  }
  o << in << "}\n";
  o << "}";
  return o.str();
}

}  // namespace gwm::core
