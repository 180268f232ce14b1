// gwm_ingest.hpp -- prepared-case input path of the C++ host mirror.
//
// gwm_prepared_case_driver starts from two JSON documents, an analysis state
// ("gwm-next-analysis-state/v1") and a lateral-boundary cache
// ("gwm-next-boundary-cache/v1"), as the reference's driver does
// (src/core/gwm_prepared_case_driver.cpp:354-509).  This header is organised
// as a pipeline of its own:
//
//   1. schema    the canonical field table of both documents (canonical
//                names and the HRRR / RRFS source names that feed them);
//   2. reading   a reader that walks a document with its key path and stops
//                at the first problem, then one consistency check per IR;
//   3. time      the snapshots around a forecast second and their blend;
//   4. balance   a per-column hydrostatic balancer, face momenta from the
//                balanced density, the barotropic projection of fully open
//                domains (lexicographic SOR on a zero-padded potential) and
//                w from column continuity;
//   5. states    dry states and tracers of a layout, start-up diagnostics;
//   6. strips    strip masks of the open sides and the boundary updaters,
//                which also pack the strips for the device stepper;
//   7. summary   runtime summary and its JSON.
//
// The reference's API names (types, free functions, updater classes) are
// kept so callers compile unchanged.  Numerics follow the reference where
// the driver's results depend on them: the EOS, the hydrostatic trapezoid
// recurrence with four fixed-point passes (prepared_case_init.cpp:320-384),
// the rho*theta inversion with two corrections (38-50), the SOR constants
// (omega 1.85, at most 2000 sweeps, stop at 1e-10; 126-170), up to three
// projection passes stopping at 1e-8, and the linear snapshot blend
// (runtime_case.cpp:205-228).  Arithmetic is FP32 (`real`) as there.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <iomanip>
#include <limits>
#include <optional>
#include <sstream>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "gwm/gwm_dycore.hpp"
#include "gwm/gwm_json.hpp"

// ---------------------------------------------------------------------------
// Host thermodynamics in FP32 (include/gwm/core/dry_thermo.hpp:15-44).
// ---------------------------------------------------------------------------
namespace gwm::core {

constexpr real kDryGasConstant = 287.05f;
constexpr real kReferencePressure = 100000.0f;
constexpr real kKappa = 0.2854f;

inline real dry_cp() { return kDryGasConstant / kKappa; }
inline real dry_theta_from_conserved(real rho_d, real rho_theta_m) {
  return rho_theta_m / std::max(rho_d, 1.0e-6f);
}
// p = p0 (R rho theta / p0)^(1 / (1 - kappa)), the argument floored at 1e-6.
inline real dry_pressure_from_rho_theta_m(real rho_d, real rho_theta_m) {
  const real arg = rho_d * kDryGasConstant * dry_theta_from_conserved(rho_d, rho_theta_m) /
                   kReferencePressure;
  return kReferencePressure * std::pow(std::max(arg, 1.0e-6f), 1.0f / (1.0f - kKappa));
}
// rho = p / (R theta Pi), Pi = (p / p0)^kappa, both floored.
inline real dry_rho_from_pressure_theta(real pressure, real theta) {
  const real exner = std::pow(std::max(pressure / kReferencePressure, 1.0e-6f), kKappa);
  return pressure / std::max(kDryGasConstant * theta * exner, 1.0e-6f);
}

}  // namespace gwm::core

namespace gwm::ingest {

// ===========================================================================
// 1. Schema
// ===========================================================================
enum class SourceKind { HRRR, RRFS, RAP, GFS, ERA5, ECMWF };

struct HorizontalGridIR {
  int nx = 0, ny = 0, nz = 0;
  double dx = 0.0, dy = 0.0, z_top = 0.0;
  double ref_lat = 0.0, ref_lon = 0.0, truelat1 = 0.0, truelat2 = 0.0, stand_lon = 0.0;
  std::size_t cell_count_2d() const { return static_cast<std::size_t>(nx) * ny; }
  std::size_t cell_count_3d() const { return cell_count_2d() * static_cast<std::size_t>(nz); }
};

// Named FP32 fields, x fastest, (nz, ny, nx) or (ny, nx).
struct FieldBundle3D {
  std::unordered_map<std::string, std::vector<real>> values;
  bool has(const std::string& n) const { return values.find(n) != values.end(); }
  std::size_t field_count() const { return values.size(); }
};
struct FieldBundle2D {
  std::unordered_map<std::string, std::vector<real>> values;
  bool has(const std::string& n) const { return values.find(n) != values.end(); }
  std::size_t field_count() const { return values.size(); }
};

struct SourceFieldBinding {
  std::string source_name;
  std::string canonical_name;
};

struct AnalysisStateIR {
  SourceKind source = SourceKind::HRRR;
  HorizontalGridIR grid;
  std::string cycle_time_utc, valid_time_utc;
  int forecast_offset_seconds = 0;
  std::unordered_map<std::string, std::string> metadata;
  FieldBundle3D atmosphere;
  FieldBundle2D surface, static_surface;
};

struct BoundarySnapshotIR {
  int forecast_offset_seconds = 0;
  std::string valid_time_utc;
  FieldBundle3D atmosphere;
  FieldBundle2D surface;
};

struct BoundaryCacheIR {
  SourceKind source = SourceKind::HRRR;
  HorizontalGridIR grid;
  std::string cycle_time_utc;
  int boundary_interval_seconds = 3600;
  std::unordered_map<std::string, std::string> metadata;
  std::vector<BoundarySnapshotIR> snapshots;
  bool empty() const { return snapshots.empty(); }
};

struct PreparedRuntimeCase {
  AnalysisStateIR analysis;
  BoundaryCacheIR boundary_cache;
};

namespace schema {

enum class Group { Atmosphere, Surface, Static };

// Canonical fields read by the runtime, with the source field each of the
// two supported sources maps onto it (canonical_ir.hpp's HRRR and RRFS
// adapters).  Static fields are part of the analysis state only.
struct Row {
  Group group;
  const char* canonical;
  const char* hrrr;
  const char* rrfs;
};
inline const std::vector<Row>& table() {
  static const std::vector<Row> rows = {
      {Group::Atmosphere, "u_wind", "UGRD", "UGRD"},
      {Group::Atmosphere, "v_wind", "VGRD", "VGRD"},
      {Group::Atmosphere, "w_wind", "VVEL", "DZDT"},
      {Group::Atmosphere, "air_temperature", "T", "TMP"},
      {Group::Atmosphere, "specific_humidity", "SPFH", "SPFH"},
      {Group::Atmosphere, "air_pressure", "PRES", "PRES"},
      {Group::Atmosphere, "geopotential_height", "HGT", "HGT"},
      {Group::Surface, "surface_pressure", "PSFC", "PRES"},
      {Group::Surface, "air_temperature_2m", "T2", "TMP_2M"},
      {Group::Surface, "specific_humidity_2m", "Q2", "SPFH_2M"},
      {Group::Surface, "u_wind_10m", "U10", "UGRD_10M"},
      {Group::Surface, "v_wind_10m", "V10", "VGRD_10M"},
      {Group::Surface, "skin_temperature", "TSK", "TMP_SFC"},
      {Group::Static, "terrain_height", "HGT", "HGT_SFC"},
      {Group::Static, "land_mask", "LANDMASK", "LANDMASK"},
      {Group::Static, "land_use_index", "LU_INDEX", "LANDUSE"},
  };
  return rows;
}

constexpr std::array<SourceKind, 6> kAllSources = {SourceKind::HRRR, SourceKind::RRFS,
                                                   SourceKind::RAP,  SourceKind::GFS,
                                                   SourceKind::ERA5, SourceKind::ECMWF};
constexpr std::array<const char*, 6> kSourceNames = {"HRRR", "RRFS", "RAP", "GFS", "ERA5", "ECMWF"};

// HRRR and RRFS carry runtime field bindings; the others are catalogued only.
inline bool has_bindings(SourceKind k) { return k == SourceKind::HRRR || k == SourceKind::RRFS; }

inline const char* group_label(Group g) {
  return g == Group::Atmosphere ? "atmosphere" : g == Group::Surface ? "surface" : "static_surface";
}

}  // namespace schema

inline std::string to_string(SourceKind k) {
  return schema::kSourceNames[static_cast<std::size_t>(k)];
}

// Source-field bindings of one group for a source with runtime bindings
// (empty for the catalogued-only sources).
inline std::vector<SourceFieldBinding> source_field_bindings(SourceKind k, schema::Group g) {
  std::vector<SourceFieldBinding> out;
  if (!schema::has_bindings(k)) return out;
  for (const auto& r : schema::table())
    if (r.group == g) out.push_back({k == SourceKind::HRRR ? r.hrrr : r.rrfs, r.canonical});
  return out;
}

// ===========================================================================
// 2. Reading and checking
// ===========================================================================
namespace read {

// A JSON node with the key path that led to it, so every problem names the
// document and the key.
class Node {
 public:
  Node(const json::Value& v, std::string where) : v_(&v), where_(std::move(where)) {}
  Node operator[](const std::string& key) const {
    if (!v_->is_object()) fail("is not an object");
    const auto& o = v_->as_object();
    const auto it = o.find(key);
    if (it == o.end()) throw std::runtime_error(where_ + ": required key '" + key + "' is absent");
    return Node(it->second, where_ + "." + key);
  }
  bool has(const std::string& key) const { return v_->has(key); }
  int integer() const {
    const double d = number();
    if (d != std::floor(d) || std::fabs(d) > 2.0e9) fail("is not an integer");
    return static_cast<int>(d);
  }
  double number() const {
    if (!v_->is_number()) fail("is not a number");
    return v_->as_number();
  }
  std::string text() const {
    if (!v_->is_string()) fail("is not a string");
    return v_->as_string();
  }
  std::unordered_map<std::string, std::string> labels() const {
    if (!v_->is_object()) fail("is not an object");
    return json::to_string_map(*v_);
  }
  template <class Bundle>
  Bundle fields() const {
    if (!v_->is_object()) fail("is not an object of named fields");
    Bundle b;
    for (const auto& kv : v_->as_object()) {
      if (!(kv.second.is_numbers() || kv.second.is_array()))
        throw std::runtime_error(where_ + "." + kv.first + ": a field must be a list of numbers");
      b.values.emplace(kv.first, json::to_real_vector(kv.second));
    }
    return b;
  }
  std::vector<Node> items() const {
    if (!(v_->is_array() || v_->is_numbers())) fail("is not a list");
    std::vector<Node> out;
    const auto& a = v_->as_array();
    for (std::size_t n = 0; n < a.size(); ++n)
      out.emplace_back(a[n], where_ + "[" + std::to_string(n) + "]");
    return out;
  }
  [[noreturn]] void fail(const std::string& what) const {
    throw std::runtime_error(where_ + " " + what);
  }

 private:
  const json::Value* v_;
  std::string where_;
};

inline SourceKind source_of(const Node& n) {
  const std::string s = n.text();
  for (std::size_t q = 0; q < schema::kAllSources.size(); ++q)
    if (s == schema::kSourceNames[q]) return schema::kAllSources[q];
  n.fail("names unknown source '" + s + "' (known: HRRR, RRFS, RAP, GFS, ERA5, ECMWF)");
}

inline HorizontalGridIR grid_of(const Node& n) {
  HorizontalGridIR g;
  g.nx = n["nx"].integer();
  g.ny = n["ny"].integer();
  g.nz = n["nz"].integer();
  g.dx = n["dx"].number();
  g.dy = n["dy"].number();
  double* optional[] = {&g.z_top, &g.ref_lat, &g.ref_lon, &g.truelat1, &g.truelat2, &g.stand_lon};
  const char* keys[] = {"z_top", "ref_lat", "ref_lon", "truelat1", "truelat2", "stand_lon"};
  for (int q = 0; q < 6; ++q)
    if (n.has(keys[q])) *optional[q] = n[keys[q]].number();
  return g;
}

inline void expect_schema(const Node& root, const char* schema) {
  const std::string got = root["schema_version"].text();
  if (got != schema)
    throw std::runtime_error(std::string("document declares schema '") + got + "', expected '" +
                             schema + "'");
}

// The first problem of a field bundle against the schema and the grid, or
// an empty string.
template <class Bundle>
std::string bundle_problem(const Bundle& b, SourceKind src, schema::Group g, std::size_t n,
                           const std::string& owner) {
  for (const auto& r : schema::table())
    if (r.group == g && !b.has(r.canonical))
      return owner + " " + schema::group_label(g) + " lacks the canonical field '" +
             r.canonical + "' (" + to_string(src) + " " +
             (src == SourceKind::HRRR ? r.hrrr : r.rrfs) + ")";
  for (const auto& kv : b.values)
    if (kv.second.size() != n)
      return owner + " " + schema::group_label(g) + " field '" + kv.first + "' holds " +
             std::to_string(kv.second.size()) + " values; the grid needs " + std::to_string(n);
  return {};
}

inline std::string grid_problem(const HorizontalGridIR& g, const std::string& owner,
                                bool need_top) {
  if (g.nx <= 0 || g.ny <= 0 || g.nz <= 0) return owner + " grid needs positive nx, ny and nz";
  if (!(g.dx > 0.0) || !(g.dy > 0.0)) return owner + " grid needs positive dx and dy";
  if (need_top && !(g.z_top > 0.0)) return owner + " grid needs a positive z_top";
  return {};
}

inline void raise_if(const std::string& problem) {
  if (!problem.empty()) throw std::runtime_error(problem);
}

}  // namespace read

// Consistency of an analysis state (runtime_case.cpp:230-253 checks the same
// properties): grid, source with runtime bindings, every canonical field of
// the three groups, field sizes.
inline void validate_runtime_analysis(const AnalysisStateIR& a) {
  const std::string who = "analysis";
  read::raise_if(read::grid_problem(a.grid, who, true));
  if (!schema::has_bindings(a.source))
    throw std::runtime_error("analysis source " + to_string(a.source) +
                             " has no runtime field bindings (HRRR and RRFS do)");
  read::raise_if(read::bundle_problem(a.atmosphere, a.source, schema::Group::Atmosphere,
                                      a.grid.cell_count_3d(), who));
  read::raise_if(read::bundle_problem(a.surface, a.source, schema::Group::Surface,
                                      a.grid.cell_count_2d(), who));
  read::raise_if(read::bundle_problem(a.static_surface, a.source, schema::Group::Static,
                                      a.grid.cell_count_2d(), who));
}

// Consistency of a boundary cache (runtime_case.cpp:255-274 checks the same
// properties): grid, interval, source with runtime bindings, at least one
// snapshot, strictly increasing offsets, every snapshot's fields.
inline void validate_runtime_boundary_cache(const BoundaryCacheIR& c) {
  const std::string who = "boundary cache";
  read::raise_if(read::grid_problem(c.grid, who, false));
  if (c.boundary_interval_seconds <= 0)
    throw std::runtime_error("boundary cache needs a positive boundary_interval_seconds");
  if (!schema::has_bindings(c.source))
    throw std::runtime_error("boundary-cache source " + to_string(c.source) +
                             " has no runtime field bindings (HRRR and RRFS do)");
  if (c.snapshots.empty()) throw std::runtime_error("boundary cache holds no snapshots");
  for (std::size_t n = 0; n < c.snapshots.size(); ++n) {
    const auto& s = c.snapshots[n];
    if (n > 0 && s.forecast_offset_seconds <= c.snapshots[n - 1].forecast_offset_seconds)
      throw std::runtime_error("boundary snapshots must have strictly increasing forecast "
                               "offsets (snapshot " + std::to_string(n) + " at " +
                               std::to_string(s.forecast_offset_seconds) + " s follows " +
                               std::to_string(c.snapshots[n - 1].forecast_offset_seconds) + " s)");
    const std::string tag = "boundary snapshot " + std::to_string(n);
    read::raise_if(read::bundle_problem(s.atmosphere, c.source, schema::Group::Atmosphere,
                                        c.grid.cell_count_3d(), tag));
    read::raise_if(read::bundle_problem(s.surface, c.source, schema::Group::Surface,
                                        c.grid.cell_count_2d(), tag));
  }
}

inline AnalysisStateIR load_analysis_state_json(const std::string& path) {
  const json::Value doc = json::parse_file(path);
  const read::Node root(doc, path);
  read::expect_schema(root, "gwm-next-analysis-state/v1");
  AnalysisStateIR a;
  a.source = read::source_of(root["source"]);
  a.grid = read::grid_of(root["grid"]);
  a.cycle_time_utc = root["cycle_time_utc"].text();
  a.valid_time_utc = root["valid_time_utc"].text();
  a.forecast_offset_seconds = root["forecast_offset_seconds"].integer();
  a.metadata = root["metadata"].labels();
  a.atmosphere = root["atmosphere"].fields<FieldBundle3D>();
  a.surface = root["surface"].fields<FieldBundle2D>();
  a.static_surface = root["static_surface"].fields<FieldBundle2D>();
  validate_runtime_analysis(a);
  return a;
}

inline BoundaryCacheIR load_boundary_cache_json(const std::string& path) {
  const json::Value doc = json::parse_file(path);
  const read::Node root(doc, path);
  read::expect_schema(root, "gwm-next-boundary-cache/v1");
  BoundaryCacheIR c;
  c.source = read::source_of(root["source"]);
  c.grid = read::grid_of(root["grid"]);
  c.cycle_time_utc = root["cycle_time_utc"].text();
  c.boundary_interval_seconds = root["boundary_interval_seconds"].integer();
  c.metadata = root["metadata"].labels();
  for (const auto& s : root["snapshots"].items())
    c.snapshots.push_back(BoundarySnapshotIR{s["forecast_offset_seconds"].integer(),
                                             s["valid_time_utc"].text(),
                                             s["atmosphere"].fields<FieldBundle3D>(),
                                             s["surface"].fields<FieldBundle2D>()});
  validate_runtime_boundary_cache(c);
  return c;
}

inline PreparedRuntimeCase load_prepared_runtime_case(const std::string& analysis_path,
                                                      const std::string& boundary_path) {
  PreparedRuntimeCase rc{load_analysis_state_json(analysis_path),
                         load_boundary_cache_json(boundary_path)};
  const auto& ga = rc.analysis.grid;
  const auto& gb = rc.boundary_cache.grid;
  if (rc.analysis.source != rc.boundary_cache.source)
    throw std::runtime_error("analysis source " + to_string(rc.analysis.source) +
                             " and boundary-cache source " + to_string(rc.boundary_cache.source) +
                             " differ");
  if (ga.nx != gb.nx || ga.ny != gb.ny || ga.nz != gb.nz)
    throw std::runtime_error("analysis grid " + std::to_string(ga.nx) + "x" +
                             std::to_string(ga.ny) + "x" + std::to_string(ga.nz) +
                             " and boundary-cache grid " + std::to_string(gb.nx) + "x" +
                             std::to_string(gb.ny) + "x" + std::to_string(gb.nz) + " differ");
  return rc;
}

// ===========================================================================
// 3. Time: the snapshot at a forecast second
// ===========================================================================
namespace time_blend {

// (1 - alpha) a + alpha b for every field of a; b must hold the same fields.
template <class Bundle>
Bundle mix(const Bundle& a, const Bundle& b, real alpha) {
  if (a.values.size() != b.values.size())
    throw std::runtime_error("boundary snapshots to blend hold different field sets");
  Bundle out;
  for (const auto& kv : a.values) {
    const auto it = b.values.find(kv.first);
    if (it == b.values.end() || it->second.size() != kv.second.size())
      throw std::runtime_error("boundary snapshots to blend differ in field '" + kv.first + "'");
    std::vector<real> v(kv.second.size());
    const real keep = 1.0f - alpha;
    for (std::size_t n = 0; n < v.size(); ++n) v[n] = keep * kv.second[n] + alpha * it->second[n];
    out.values.emplace(kv.first, std::move(v));
  }
  return out;
}

}  // namespace time_blend

// The boundary snapshot at forecast second t: the latest snapshot at or
// before t and the earliest at or after it, blended linearly; before the
// first or after the last snapshot, that snapshot (runtime_case.cpp:205-228).
inline BoundarySnapshotIR interpolate_boundary_snapshot(const BoundaryCacheIR& c, int t) {
  validate_runtime_boundary_cache(c);
  const auto& s = c.snapshots;
  // first snapshot with offset >= t (offsets increase strictly)
  const auto hi = std::lower_bound(s.begin(), s.end(), t, [](const BoundarySnapshotIR& x, int v) {
    return x.forecast_offset_seconds < v;
  });
  if (hi == s.end()) return s.back();
  if (hi->forecast_offset_seconds == t || hi == s.begin()) return *hi;
  const auto lo = hi - 1;
  const real alpha = static_cast<real>(t - lo->forecast_offset_seconds) /
                     static_cast<real>(hi->forecast_offset_seconds - lo->forecast_offset_seconds);
  BoundarySnapshotIR out;
  out.forecast_offset_seconds = t;
  out.valid_time_utc = lo->valid_time_utc + "->" + hi->valid_time_utc;
  out.atmosphere = time_blend::mix(lo->atmosphere, hi->atmosphere, alpha);
  out.surface = time_blend::mix(lo->surface, hi->surface, alpha);
  return out;
}

// ===========================================================================
// 4. Balance
// ===========================================================================
struct PreparedCaseInitConfig {
  int halo = 1;
  int ranks_x = 1;
  int ranks_y = 1;
  double terrain_taper_eta = 0.25;
  bool periodic_x = false;
  bool periodic_y = false;
};

namespace balance {

inline constexpr real kGravity = 9.81f;

// Global (haloless) fields of a grid: cells (nz, ny, nx), x-faces
// (nz, ny, nx+1), y-faces (nz, ny+1, nx), z-faces (nz+1, ny, nx).
struct Dims {
  int nx, ny, nz;
  std::size_t cell(int i, int j, int k) const {
    return (static_cast<std::size_t>(k) * ny + j) * nx + i;
  }
  std::size_t xface(int fi, int j, int k) const {
    return (static_cast<std::size_t>(k) * ny + j) * (nx + 1) + fi;
  }
  std::size_t yface(int i, int fj, int k) const {
    return (static_cast<std::size_t>(k) * (ny + 1) + fj) * nx + i;
  }
  std::size_t zface(int i, int j, int fk) const { return cell(i, j, fk); }
  std::size_t col(int i, int j) const { return static_cast<std::size_t>(j) * nx + i; }
  std::size_t cells() const { return static_cast<std::size_t>(nx) * ny * nz; }
};

// The balanced dry fields of one analysis (or boundary snapshot).
struct Fields {
  std::vector<real> pressure, theta, rho, rho_theta;  // cells
  std::vector<real> mom_u, mom_v, mom_w;              // faces
};

// rho*theta whose EOS pressure at rho = 1 is p: the closed-form inverse of
// the EOS without its floors, then two multiplicative corrections through
// the floored EOS (prepared_case_init.cpp:38-50).
inline real rho_theta_for_pressure(real p) {
  using namespace gwm::core;
  const real expo = 1.0f - kKappa;
  real guess = std::pow(std::max(p / kReferencePressure, 1.0e-6f), expo) *
               (kReferencePressure / kDryGasConstant);
  for (int pass = 0; pass < 2; ++pass) {
    const real implied = std::max(dry_pressure_from_rho_theta_m(1.0f, guess), 1.0f);
    guess *= std::pow(std::max(p / implied, 1.0e-6f), expo);
  }
  return guess;
}

// One column: potential temperature from the analysis (T, p), the analysis
// pressure at the lowest level, and above it the trapezoid hydrostatic
// recurrence p_k = p_{k-1} - g (rho_{k-1} + rho_k) dz / 2 solved by four
// fixed-point passes from an explicit first guess, all floored at 1 Pa
// (prepared_case_init.cpp:320-384).  In: p_src, t_src, z (nz each).
struct ColumnBalance {
  std::vector<real> p, theta, rho, rho_theta;
  void solve(const real* p_src, const real* t_src, const real* z, int nz) {
    using namespace gwm::core;
    p.resize(nz);
    theta.resize(nz);
    rho.resize(nz);
    rho_theta.resize(nz);
    for (int k = 0; k < nz; ++k) {
      const real pk = std::max(p_src[k], 1.0f);
      theta[k] = t_src[k] * std::pow(kReferencePressure / pk, kKappa);
    }
    p[0] = std::max(p_src[0], 1.0f);
    rho[0] = dry_rho_from_pressure_theta(p[0], theta[0]);
    for (int k = 1; k < nz; ++k) {
      const real dz = z[k] - z[k - 1];
      const real below = p[k - 1], rho_below = rho[k - 1];
      real pk = std::max(1.0f, below - kGravity * rho_below * dz);
      for (int pass = 0; pass < 4; ++pass) {
        const real rho_here = dry_rho_from_pressure_theta(pk, theta[k]);
        pk = std::max(below - kGravity * 0.5f * (rho_below + rho_here) * dz, 1.0f);
      }
      p[k] = pk;
      rho[k] = dry_rho_from_pressure_theta(pk, theta[k]);
    }
    for (int k = 0; k < nz; ++k) rho_theta[k] = rho_theta_for_pressure(p[k]);
  }
};

// Successive over-relaxation for lap(phi) = rhs on the cell centres of an
// nx x ny grid with phi = 0 outside (five-point Laplacian), lexicographic
// order, omega 1.85, at most 2000 sweeps, stopping once no update moves phi
// by more than 1e-10 (prepared_case_init.cpp:126-170).  phi is held with a
// ring of zeros so every point reads four neighbours.
inline std::vector<real> sor_potential(const std::vector<real>& rhs, int nx, int ny, real dx,
                                       real dy) {
  const int px = nx + 2;
  std::vector<real> pad(static_cast<std::size_t>(px) * (ny + 2), 0.0f);
  std::vector<real> phi(static_cast<std::size_t>(nx) * ny, 0.0f);
  if (nx <= 0 || ny <= 0) return phi;
  const real cx = 1.0f / std::max(dx * dx, 1.0e-6f);
  const real cy = 1.0f / std::max(dy * dy, 1.0e-6f);
  const real diag = 2.0f * (cx + cy);
  if (!(diag > 0.0f)) return phi;
  constexpr real omega = 1.85f;
  for (int sweep = 0; sweep < 2000; ++sweep) {
    real largest = 0.0f;
    for (int j = 0; j < ny; ++j) {
      real* row = pad.data() + static_cast<std::size_t>(j + 1) * px + 1;
      const real* rs = rhs.data() + static_cast<std::size_t>(j) * nx;
      for (int i = 0; i < nx; ++i) {
        const real target =
            ((row[i - 1] + row[i + 1]) * cx + (row[i - px] + row[i + px]) * cy - rs[i]) / diag;
        const real next = row[i] + omega * (target - row[i]);
        largest = std::max(largest, std::fabs(next - row[i]));
        row[i] = next;
      }
    }
    if (largest <= 1.0e-10f) break;
  }
  for (int j = 0; j < ny; ++j)
    std::copy_n(pad.data() + static_cast<std::size_t>(j + 1) * px + 1, nx,
                phi.data() + static_cast<std::size_t>(j) * nx);
  return phi;
}

// Layer thicknesses dz = 1 / max(inv_dz_cell, 1e-6) of every cell.
inline std::vector<real> thicknesses(const domain::GridMetrics& m, const Dims& d) {
  std::vector<real> dz(d.cells());
  for (int k = 0; k < d.nz; ++k)
    for (int j = 0; j < d.ny; ++j)
      for (int i = 0; i < d.nx; ++i)
        dz[d.cell(i, j, k)] = 1.0f / std::max(m.inv_dz_cell(i, j, k), 1.0e-6f);
  return dz;
}

// Horizontal mass-flux divergence of every cell.
inline std::vector<real> horizontal_divergence(const std::vector<real>& mu,
                                               const std::vector<real>& mv, const Dims& d,
                                               real dx, real dy) {
  std::vector<real> out(d.cells());
  for (int k = 0; k < d.nz; ++k)
    for (int j = 0; j < d.ny; ++j)
      for (int i = 0; i < d.nx; ++i)
        out[d.cell(i, j, k)] = (mu[d.xface(i + 1, j, k)] - mu[d.xface(i, j, k)]) / dx +
                               (mv[d.yface(i, j + 1, k)] - mv[d.yface(i, j, k)]) / dy;
  return out;
}

// Column means of a cell field weighted by the layer thickness.
inline std::vector<real> column_means(const std::vector<real>& f, const std::vector<real>& dz,
                                      const Dims& d) {
  std::vector<real> acc(static_cast<std::size_t>(d.nx) * d.ny, 0.0f), depth(acc.size(), 0.0f);
  for (int k = 0; k < d.nz; ++k)
    for (std::size_t c = 0; c < acc.size(); ++c) {
      const std::size_t e = static_cast<std::size_t>(k) * acc.size() + c;
      acc[c] += f[e] * dz[e];
      depth[c] += dz[e];
    }
  for (std::size_t c = 0; c < acc.size(); ++c) acc[c] /= std::max(depth[c], 1.0e-6f);
  return acc;
}

// The balanced dry state of an analysis: hydrostatic columns, face momenta
// rho_f * u_f with the density and wind averaged to each face (wrapped on
// periodic axes, clamped on open ones), on fully open domains up to three
// barotropic projections that remove the column-mean horizontal divergence
// (stopping once it is at most 1e-8), and w from column continuity with the
// column mean removed, zero at the ground and the lid
// (prepared_case_init.cpp:172-318, 320-384).
inline Fields build(const AnalysisStateIR& a, const domain::GridMetrics& m) {
  const Dims d{a.grid.nx, a.grid.ny, a.grid.nz};
  const real dx = static_cast<real>(a.grid.dx), dy = static_cast<real>(a.grid.dy);
  const auto& p_src = a.atmosphere.values.at("air_pressure");
  const auto& t_src = a.atmosphere.values.at("air_temperature");
  const auto& u_src = a.atmosphere.values.at("u_wind");
  const auto& v_src = a.atmosphere.values.at("v_wind");
  Fields f;
  f.pressure.resize(d.cells());
  f.theta.resize(d.cells());
  f.rho.resize(d.cells());
  f.rho_theta.resize(d.cells());
  {
    std::vector<real> cp(d.nz), ct(d.nz), cz(d.nz);
    ColumnBalance col;
    for (int j = 0; j < d.ny; ++j)
      for (int i = 0; i < d.nx; ++i) {
        for (int k = 0; k < d.nz; ++k) {
          cp[k] = p_src[d.cell(i, j, k)];
          ct[k] = t_src[d.cell(i, j, k)];
          cz[k] = m.z_center(i, j, k);
        }
        col.solve(cp.data(), ct.data(), cz.data(), d.nz);
        for (int k = 0; k < d.nz; ++k) {
          const std::size_t c = d.cell(i, j, k);
          f.pressure[c] = col.p[k];
          f.theta[c] = col.theta[k];
          f.rho[c] = col.rho[k];
          f.rho_theta[c] = col.rho_theta[k];
        }
      }
  }
  auto side = [](int q, int n, bool periodic) {
    if (periodic) return ((q % n) + n) % n;
    return std::min(std::max(q, 0), n - 1);
  };
  f.mom_u.assign(static_cast<std::size_t>(d.nx + 1) * d.ny * d.nz, 0.0f);
  f.mom_v.assign(static_cast<std::size_t>(d.nx) * (d.ny + 1) * d.nz, 0.0f);
  f.mom_w.assign(static_cast<std::size_t>(d.nx) * d.ny * (d.nz + 1), 0.0f);
  for (int k = 0; k < d.nz; ++k)
    for (int j = 0; j < d.ny; ++j) {
      for (int fi = 0; fi <= d.nx; ++fi) {
        const std::size_t l = d.cell(side(fi - 1, d.nx, m.periodic_x), j, k);
        const std::size_t r = d.cell(side(fi, d.nx, m.periodic_x), j, k);
        f.mom_u[d.xface(fi, j, k)] = 0.5f * (f.rho[l] + f.rho[r]) * (0.5f * (u_src[l] + u_src[r]));
      }
    }
  for (int k = 0; k < d.nz; ++k)
    for (int fj = 0; fj <= d.ny; ++fj)
      for (int i = 0; i < d.nx; ++i) {
        const std::size_t s = d.cell(i, side(fj - 1, d.ny, m.periodic_y), k);
        const std::size_t n = d.cell(i, side(fj, d.ny, m.periodic_y), k);
        f.mom_v[d.yface(i, fj, k)] = 0.5f * (f.rho[s] + f.rho[n]) * (0.5f * (v_src[s] + v_src[n]));
      }
  const std::vector<real> dz = thicknesses(m, d);
  if (!m.periodic_x && !m.periodic_y) {
    for (int pass = 0; pass < 3; ++pass) {
      const auto mean =
          column_means(horizontal_divergence(f.mom_u, f.mom_v, d, dx, dy), dz, d);
      real worst = 0.0f;
      for (real v : mean) worst = std::max(worst, std::fabs(v));
      if (worst <= 1.0e-8f) break;
      const auto phi = sor_potential(mean, d.nx, d.ny, dx, dy);
      // phi is 0 outside the domain, so the outer faces see one side only
      auto at = [&](int i, int j) { return (i < 0 || i >= d.nx || j < 0 || j >= d.ny) ? 0.0f : phi[d.col(i, j)]; };
      for (int j = 0; j < d.ny; ++j)
        for (int fi = 0; fi <= d.nx; ++fi) {
          const real grad = (at(fi, j) - at(fi - 1, j)) / dx;
          for (int k = 0; k < d.nz; ++k) f.mom_u[d.xface(fi, j, k)] -= grad;
        }
      for (int fj = 0; fj <= d.ny; ++fj)
        for (int i = 0; i < d.nx; ++i) {
          const real grad = (at(i, fj) - at(i, fj - 1)) / dy;
          for (int k = 0; k < d.nz; ++k) f.mom_v[d.yface(i, fj, k)] -= grad;
        }
    }
  }
  const auto hdiv = horizontal_divergence(f.mom_u, f.mom_v, d, dx, dy);
  const auto mean = column_means(hdiv, dz, d);
  for (int j = 0; j < d.ny; ++j)
    for (int i = 0; i < d.nx; ++i) {
      real w = 0.0f;
      f.mom_w[d.zface(i, j, 0)] = 0.0f;
      for (int k = 0; k < d.nz; ++k) {
        const std::size_t c = d.cell(i, j, k);
        w = w - (hdiv[c] - mean[d.col(i, j)]) * dz[c];
        f.mom_w[d.zface(i, j, k + 1)] = w;
      }
      f.mom_w[d.zface(i, j, d.nz)] = 0.0f;
    }
  return f;
}

// An analysis whose atmosphere and surface are a boundary snapshot's.
inline AnalysisStateIR with_snapshot(const AnalysisStateIR& a, const BoundarySnapshotIR& s) {
  AnalysisStateIR out = a;
  out.valid_time_utc = s.valid_time_utc;
  out.forecast_offset_seconds = s.forecast_offset_seconds;
  out.atmosphere = s.atmosphere;
  out.surface = s.surface;
  return out;
}

}  // namespace balance

// ===========================================================================
// 5. States of a layout and the start-up diagnostics
// ===========================================================================
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
inline void same_shape(const AnalysisStateIR& a, const domain::GridMetrics& m) {
  if (m.nx != a.grid.nx || m.ny != a.grid.ny || m.nz != a.grid.nz)
    throw std::runtime_error("grid metrics and analysis differ in shape");
}
// Running maximum.
struct Peak {
  real v = 0.0f;
  void see(real x) { v = std::max(v, x); }
};
}  // namespace detail

inline std::vector<domain::SubdomainDescriptor> make_prepared_case_layout(
    const AnalysisStateIR& a, const PreparedCaseInitConfig& c) {
  validate_runtime_analysis(a);
  if (c.halo <= 0) throw std::runtime_error("prepared-case halo must be at least 1");
  if (c.ranks_x <= 0 || c.ranks_y <= 0)
    throw std::runtime_error("prepared-case rank counts must be at least 1");
  return comm::VirtualRankLayout::build(a.grid.nx, a.grid.ny, a.grid.nz, c.halo, c.ranks_x,
                                        c.ranks_y, c.periodic_x, c.periodic_y);
}

inline domain::GridMetrics make_prepared_case_metrics(const AnalysisStateIR& a,
                                                      const PreparedCaseInitConfig& c) {
  validate_runtime_analysis(a);
  return domain::GridMetrics::make_hybrid_height(
      a.grid.nx, a.grid.ny, a.grid.nz, a.grid.dx, a.grid.dy, a.grid.z_top,
      a.static_surface.values.at("terrain_height"), c.terrain_taper_eta, c.periodic_x,
      c.periodic_y, {}, c.halo);
}

namespace detail {
// Copies balanced global fields into the ranks of a layout.
inline void fill_dry(std::vector<dycore::DryState>& states,
                     const std::vector<domain::SubdomainDescriptor>& layout,
                     const balance::Fields& f, const balance::Dims& d) {
  for (std::size_t r = 0; r < layout.size(); ++r) {
    const auto& L = layout[r];
    auto& s = states[r];
    const int nxl = L.nx_local(), nyl = L.ny_local();
    for (int k = 0; k < L.nz; ++k)
      for (int j = 0; j < nyl; ++j) {
        for (int i = 0; i < nxl; ++i) {
          const std::size_t c = d.cell(L.i_begin + i, L.j_begin + j, k);
          s.rho_d(i, j, k) = f.rho[c];
          s.rho_theta_m(i, j, k) = f.rho_theta[c];
        }
        for (int fi = 0; fi <= nxl; ++fi)
          s.mom_u.storage()(fi, j, k) = f.mom_u[d.xface(L.i_begin + fi, L.j_begin + j, k)];
      }
    for (int k = 0; k < L.nz; ++k)
      for (int fj = 0; fj <= nyl; ++fj)
        for (int i = 0; i < nxl; ++i)
          s.mom_v.storage()(i, fj, k) = f.mom_v[d.yface(L.i_begin + i, L.j_begin + fj, k)];
    for (int fk = 0; fk <= L.nz; ++fk)
      for (int j = 0; j < nyl; ++j)
        for (int i = 0; i < nxl; ++i)
          s.mom_w.storage()(i, j, fk) = f.mom_w[d.zface(L.i_begin + i, L.j_begin + j, fk)];
  }
}
}  // namespace detail

// The balanced dry state of the analysis on each rank of the layout
// (make_dry_states_from_analysis, prepared_case_init.cpp:644-744).
inline std::vector<dycore::DryState> make_dry_states_from_analysis(
    const AnalysisStateIR& a, const domain::GridMetrics& m,
    const std::vector<domain::SubdomainDescriptor>& layout,
    const std::string& label_prefix = "prepared_case") {
  validate_runtime_analysis(a);
  if (layout.empty()) throw std::runtime_error("prepared-case layout has no ranks");
  detail::same_shape(a, m);
  const balance::Dims d{a.grid.nx, a.grid.ny, a.grid.nz};
  const balance::Fields f = balance::build(a, m);
  std::vector<dycore::DryState> states;
  states.reserve(layout.size());
  for (const auto& L : layout) {
    states.emplace_back(L.nx_local(), L.ny_local(), L.nz, L.halo,
                        label_prefix + "_rank_" + std::to_string(L.rank));
    states.back().fill_constant(0.0f, 300.0f, 0.0f, 0.0f, 0.0f);
  }
  detail::fill_dry(states, layout, f, d);
  return states;
}

// Tracer masses rho_d * q of every registry tracer the analysis carries
// (specific humidity is required), zero (or the baseline's) otherwise
// (make_tracers_from_analysis, prepared_case_init.cpp:923-1007).
inline std::vector<state::TracerState> make_tracers_from_analysis(
    const AnalysisStateIR& a, const std::vector<dycore::DryState>& dry,
    const std::vector<domain::SubdomainDescriptor>& layout, const state::TracerRegistry& registry,
    const std::string& label_prefix = "prepared_case_tracer",
    const std::vector<state::TracerState>* baseline = nullptr) {
  validate_runtime_analysis(a);
  if (dry.size() != layout.size() || (baseline && baseline->size() != layout.size()))
    throw std::runtime_error("tracer start-up: dry states, baseline and layout disagree in rank "
                             "count");
  const balance::Dims d{a.grid.nx, a.grid.ny, a.grid.nz};
  std::vector<state::TracerState> out;
  out.reserve(layout.size());
  for (std::size_t r = 0; r < layout.size(); ++r) {
    const auto& L = layout[r];
    out.emplace_back(registry, L.nx_local(), L.ny_local(), L.nz, L.halo,
                     label_prefix + "_rank_" + std::to_string(L.rank));
    out.back().fill_zero();
    if (baseline)
      for (std::size_t q = 0; q < registry.size(); ++q)
        out.back().mass(static_cast<int>(q)).copy_all_from((*baseline)[r].mass(static_cast<int>(q)));
  }
  if (!a.atmosphere.has(state::kSpecificHumidityTracerName))
    throw std::runtime_error("tracer start-up needs the analysis specific_humidity");
  for (std::size_t q = 0; q < registry.size(); ++q) {
    const std::string& name = registry.at(static_cast<int>(q)).name;
    const auto it = a.atmosphere.values.find(name);
    if (it == a.atmosphere.values.end()) continue;
    const std::vector<real>& ratio = it->second;
    if (ratio.size() != d.cells())
      throw std::runtime_error("analysis tracer field '" + name + "' does not match the grid");
    for (std::size_t r = 0; r < layout.size(); ++r) {
      const auto& L = layout[r];
      auto& mass = out[r].mass(static_cast<int>(q));
      for (int k = 0; k < L.nz; ++k)
        for (int j = 0; j < L.ny_local(); ++j)
          for (int i = 0; i < L.nx_local(); ++i)
            mass(i, j, k) = dry[r].rho_d(i, j, k) * ratio[d.cell(L.i_begin + i, L.j_begin + j, k)];
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

// How well a state holds the balance of its analysis
// (diagnose_prepared_case_balance, prepared_case_init.cpp:746-908 defines
// the quantities):
//  - max_rel_eos: |rho - rho(p_bal, theta)| / max(|rho(p_bal, theta)|, 1);
//  - max_rel_hydrostatic: |p - p_bal| / max(|p_bal|, 1);
//  - max_rel_fast_vertical: per interior w face, the largest of the two
//    cells' |p - p_bal| / max(|p_bal|, 1) and the face's |rho' averaged| /
//    max(|rho_bal averaged|, 1);
//  - mass divergence: |div m| and |div m| / max(|d/dx|, |d/dy|, |d/dz|, 1e-12);
//  - |m_w| at the ground and the lid;
//  - tracer closure |rho q - rho_d q_analysis| for the tracers the analysis has;
//  - source heights against the metric level heights, when the analysis
//    carries geopotential_height.
inline PreparedCaseBalanceDiagnostics diagnose_prepared_case_balance(
    const AnalysisStateIR& a, const domain::GridMetrics& m,
    const std::vector<dycore::DryState>& dry,
    const std::vector<domain::SubdomainDescriptor>& layout,
    const std::vector<state::TracerState>* tracers = nullptr) {
  using namespace gwm::core;
  validate_runtime_analysis(a);
  detail::same_shape(a, m);
  if (dry.size() != layout.size() || (tracers && tracers->size() != layout.size()))
    throw std::runtime_error("balance diagnostics: states, tracers and layout disagree in rank "
                             "count");
  const balance::Dims d{a.grid.nx, a.grid.ny, a.grid.nz};
  const balance::Fields bal = balance::build(a, m);
  // the state's rho, theta and p on the global grid
  std::vector<real> rho(d.cells()), theta(d.cells()), p(d.cells());
  detail::Peak eos, hydro, fastv, mdiv_abs, mdiv_rel, w_bot, w_top, closure;
  const real dx = static_cast<real>(a.grid.dx), dy = static_cast<real>(a.grid.dy);
  for (std::size_t r = 0; r < layout.size(); ++r) {
    const auto& L = layout[r];
    const auto& s = dry[r];
    for (int k = 0; k < L.nz; ++k)
      for (int j = 0; j < L.ny_local(); ++j)
        for (int i = 0; i < L.nx_local(); ++i) {
          const int gi = L.i_begin + i, gj = L.j_begin + j;
          const std::size_t c = d.cell(gi, gj, k);
          const real rr = s.rho_d(i, j, k), rt = s.rho_theta_m(i, j, k);
          rho[c] = rr;
          theta[c] = dry_theta_from_conserved(rr, rt);
          p[c] = dry_pressure_from_rho_theta_m(rr, rt);
          const real ddx = (s.mom_u.storage()(i + 1, j, k) - s.mom_u.storage()(i, j, k)) / dx;
          const real ddy = (s.mom_v.storage()(i, j + 1, k) - s.mom_v.storage()(i, j, k)) / dy;
          const real ddz = (s.mom_w.storage()(i, j, k + 1) - s.mom_w.storage()(i, j, k)) *
                           m.inv_dz_cell(gi, gj, k);
          const real div = std::fabs(ddx + ddy + ddz);
          mdiv_abs.see(div);
          mdiv_rel.see(div / std::max({std::fabs(ddx), std::fabs(ddy), std::fabs(ddz), 1.0e-12f}));
        }
    for (int j = 0; j < L.ny_local(); ++j)
      for (int i = 0; i < L.nx_local(); ++i) {
        w_bot.see(std::fabs(s.mom_w.storage()(i, j, 0)));
        w_top.see(std::fabs(s.mom_w.storage()(i, j, L.nz)));
      }
  }
  auto rel = [](real x, real ref) { return std::fabs(x - ref) / std::max(std::fabs(ref), 1.0f); };
  for (std::size_t c = 0; c < d.cells(); ++c) {
    eos.see(rel(rho[c], dry_rho_from_pressure_theta(bal.pressure[c], theta[c])));
    hydro.see(rel(p[c], bal.pressure[c]));
  }
  for (int k = 1; k < d.nz; ++k)
    for (int j = 0; j < d.ny; ++j)
      for (int i = 0; i < d.nx; ++i) {
        const std::size_t lo = d.cell(i, j, k - 1), hi = d.cell(i, j, k);
        const real rho_dep = 0.5f * ((rho[lo] - bal.rho[lo]) + (rho[hi] - bal.rho[hi]));
        const real rho_ref = 0.5f * (bal.rho[lo] + bal.rho[hi]);
        fastv.see(std::max({rel(p[lo], bal.pressure[lo]), rel(p[hi], bal.pressure[hi]),
                            std::fabs(rho_dep) / std::max(std::fabs(rho_ref), 1.0f), 0.0f}));
      }
  PreparedCaseBalanceDiagnostics out;
  const auto zs = a.atmosphere.values.find("geopotential_height");
  if (zs != a.atmosphere.values.end()) {
    detail::Peak miss;
    for (int k = 0; k < d.nz; ++k)
      for (int j = 0; j < d.ny; ++j)
        for (int i = 0; i < d.nx; ++i)
          miss.see(std::fabs(zs->second[d.cell(i, j, k)] - m.z_center(i, j, k)));
    out.max_abs_z_src_minus_metric = miss.v;
  }
  if (tracers)
    for (std::size_t r = 0; r < layout.size(); ++r) {
      const auto& L = layout[r];
      const auto& t = (*tracers)[r];
      for (std::size_t q = 0; q < t.size(); ++q) {
        const auto it = a.atmosphere.values.find(t.registry().at(static_cast<int>(q)).name);
        if (it == a.atmosphere.values.end()) continue;
        const auto& mass = t.mass(static_cast<int>(q));
        for (int k = 0; k < L.nz; ++k)
          for (int j = 0; j < L.ny_local(); ++j)
            for (int i = 0; i < L.nx_local(); ++i)
              closure.see(std::fabs(mass(i, j, k) -
                                    dry[r].rho_d(i, j, k) *
                                        it->second[d.cell(L.i_begin + i, L.j_begin + j, k)]));
      }
    }
  out.max_rel_eos = eos.v;
  out.max_rel_hydrostatic = hydro.v;
  out.max_rel_fast_vertical = fastv.v;
  out.max_abs_mass_divergence = mdiv_abs.v;
  out.max_rel_mass_divergence = mdiv_rel.v;
  out.max_abs_mom_w_bottom = w_bot.v;
  out.max_abs_mom_w_top = w_top.v;
  out.max_abs_tracer_closure = closure.v;
  return out;
}

// ===========================================================================
// 6. Lateral strips
// ===========================================================================
namespace strips {

// Which open sides a rank touches, and membership of its cells and faces in
// the strips that the prepared boundary overwrites (the reference's column
// and row strips, prepared_case_init.cpp:424-565, described as masks):
//  - a cell (and each of its z-faces) is in a strip when it is an outermost
//    owned cell on a touched side;
//  - an x-face is overwritten when it is the outer face of a touched west /
//    east side, or lies on a row of a touched south / north strip;
//  - a y-face likewise with the axes exchanged.
struct Sides {
  bool w, e, s, n;
  int nx, ny;  // local owned cells
  explicit Sides(const domain::SubdomainDescriptor& L)
      : w(dycore::touches_west_boundary(L)), e(dycore::touches_east_boundary(L)),
        s(dycore::touches_south_boundary(L)), n(dycore::touches_north_boundary(L)),
        nx(L.nx_local()), ny(L.ny_local()) {}
  bool any() const { return w || e || s || n; }
  bool cell(int i, int j) const {
    return (w && i == 0) || (e && i == nx - 1) || (s && j == 0) || (n && j == ny - 1);
  }
  bool xface(int fi, int j) const {
    return (w && fi == 0) || (e && fi == nx) || (s && j == 0) || (n && j == ny - 1);
  }
  bool yface(int i, int fj) const {
    return (s && fj == 0) || (n && fj == ny) || (w && i == 0) || (e && i == nx - 1);
  }
};

// Overwrites the strips of every rank with balanced global fields.
inline void dry(std::vector<dycore::DryState>& states,
                const std::vector<domain::SubdomainDescriptor>& layout, const balance::Dims& d,
                const balance::Fields& f) {
  if (states.size() != layout.size())
    throw std::runtime_error("prepared boundary: states and layout disagree in rank count");
  for (std::size_t r = 0; r < layout.size(); ++r) {
    const auto& L = layout[r];
    const Sides S(L);
    if (!S.any()) continue;
    auto& st = states[r];
    for (int k = 0; k < L.nz; ++k)
      for (int j = 0; j < S.ny; ++j) {
        const int gj = L.j_begin + j;
        for (int i = 0; i < S.nx; ++i)
          if (S.cell(i, j)) {
            const std::size_t c = d.cell(L.i_begin + i, gj, k);
            st.rho_d(i, j, k) = f.rho[c];
            st.rho_theta_m(i, j, k) = f.rho_theta[c];
          }
        for (int fi = 0; fi <= S.nx; ++fi)
          if (S.xface(fi, j)) st.mom_u.storage()(fi, j, k) = f.mom_u[d.xface(L.i_begin + fi, gj, k)];
      }
    for (int k = 0; k < L.nz; ++k)
      for (int fj = 0; fj <= S.ny; ++fj)
        for (int i = 0; i < S.nx; ++i)
          if (S.yface(i, fj))
            st.mom_v.storage()(i, fj, k) = f.mom_v[d.yface(L.i_begin + i, L.j_begin + fj, k)];
    for (int j = 0; j < S.ny; ++j)
      for (int i = 0; i < S.nx; ++i)
        if (S.cell(i, j))
          for (int fk = 0; fk <= L.nz; ++fk)
            st.mom_w.storage()(i, j, fk) = f.mom_w[d.zface(L.i_begin + i, L.j_begin + j, fk)];
  }
}

// Tracer strips: balanced rho_d times the snapshot's mixing ratio, for the
// tracers the snapshot carries (prepared_case_init.cpp:567-638).
inline void tracers(std::vector<state::TracerState>& states,
                    const std::vector<domain::SubdomainDescriptor>& layout, const balance::Dims& d,
                    const BoundarySnapshotIR& snap, const std::vector<real>& rho) {
  if (states.size() != layout.size())
    throw std::runtime_error("prepared tracer boundary: tracers and layout disagree in rank count");
  for (std::size_t r = 0; r < layout.size(); ++r) {
    const auto& L = layout[r];
    const Sides S(L);
    if (!S.any()) continue;
    for (std::size_t q = 0; q < states[r].size(); ++q) {
      const auto it = snap.atmosphere.values.find(states[r].registry().at(static_cast<int>(q)).name);
      if (it == snap.atmosphere.values.end()) continue;
      auto& mass = states[r].mass(static_cast<int>(q));
      for (int k = 0; k < L.nz; ++k)
        for (int j = 0; j < S.ny; ++j)
          for (int i = 0; i < S.nx; ++i)
            if (S.cell(i, j)) {
              const std::size_t c = d.cell(L.i_begin + i, L.j_begin + j, k);
              mass(i, j, k) = rho[c] * it->second[c];
            }
    }
  }
}

// Balanced boundary fields for a forecast second, rebuilt only when the
// second changes (the reference re-projects at the same points,
// prepared_case_init.cpp:396-422).
class Source {
 public:
  Source(AnalysisStateIR analysis, BoundaryCacheIR cache, const PreparedCaseInitConfig& cfg)
      : analysis_(std::move(analysis)), cache_(std::move(cache)),
        metrics_(make_prepared_case_metrics(analysis_, cfg)),
        dims_{analysis_.grid.nx, analysis_.grid.ny, analysis_.grid.nz} {}
  bool empty() const { return cache_.empty(); }
  // Brings the balanced fields to forecast second round(step_start + t).
  // The fields are a function of the second alone, so the last few seconds
  // are kept: the RK stages visit t0, t0 + dt, t0 + dt/2 and the next step
  // starts at t0 + dt again, which the reference rebalances each time.
  void at(real step_start, real t) {
    const long second = std::lround(step_start + t);
    if (have_ && second == second_) return;
    for (auto it = recent_.begin(); it != recent_.end(); ++it)
      if (it->second == second) {
        std::swap(fields_, it->fields);
        std::swap(snapshot_, it->snapshot);
        it->second = second_;  // the fields swapped out are the previous second's
        if (!have_) recent_.erase(it);
        second_ = second;
        have_ = true;
        return;
      }
    if (have_) {
      recent_.push_back(Entry{second_, std::move(fields_), std::move(snapshot_)});
      if (recent_.size() > kRecent) recent_.erase(recent_.begin());
    }
    snapshot_ = interpolate_boundary_snapshot(cache_, static_cast<int>(second));
    fields_ = balance::build(balance::with_snapshot(analysis_, snapshot_), metrics_);
    second_ = second;
    have_ = true;
  }
  const balance::Fields& fields() const { return fields_; }
  const BoundarySnapshotIR& snapshot() const { return snapshot_; }
  const balance::Dims& dims() const { return dims_; }
  long second() const { return second_; }

 private:
  struct Entry {
    long second;
    balance::Fields fields;
    BoundarySnapshotIR snapshot;
  };
  static constexpr std::size_t kRecent = 3;
  AnalysisStateIR analysis_;
  BoundaryCacheIR cache_;
  domain::GridMetrics metrics_;
  balance::Dims dims_;
  balance::Fields fields_;
  BoundarySnapshotIR snapshot_;
  long second_ = 0;
  bool have_ = false;
  std::vector<Entry> recent_;
};

}  // namespace strips

// BoundaryUpdater of the prepared driver: overwrites the lateral strips with
// the balanced state of the boundary snapshot at forecast second
// round(step start + stage time) (PreparedCaseBoundaryUpdater,
// prepared_case_init.cpp:1019-1047).
class PreparedCaseBoundaryUpdater final : public dycore::BoundaryUpdater {
 public:
  PreparedCaseBoundaryUpdater(AnalysisStateIR analysis, BoundaryCacheIR cache,
                              PreparedCaseInitConfig config)
      : src_(std::move(analysis), std::move(cache), config) {}
  void set_step_start_time(real t) { step_start_ = std::max(t, 0.0f); }
  real step_start_time() const { return step_start_; }
  void apply(std::vector<dycore::DryState>& states,
             const std::vector<domain::SubdomainDescriptor>& layout, real sim_time) override {
    if (src_.empty()) return;
    src_.at(step_start_, sim_time);
    strips::dry(states, layout, src_.dims(), src_.fields());
  }
  // The balanced fields at stage time t of the current step (for callers
  // that apply the strips themselves, such as the device run below), and
  // the forecast second they belong to.
  const balance::Fields* fields_at(real sim_time, long* second = nullptr) {
    if (src_.empty()) return nullptr;
    src_.at(step_start_, sim_time);
    if (second) *second = src_.second();
    return &src_.fields();
  }
  const balance::Dims& dims() const { return src_.dims(); }

 private:
  strips::Source src_;
  real step_start_ = 0.0f;
};

// TracerBoundaryUpdater of the prepared driver: vapour (and any tracer the
// snapshot carries) on the strips, from the balanced density of the same
// snapshot (PreparedCaseTracerBoundaryUpdater, prepared_case_init.cpp:1049-1083).
class PreparedCaseTracerBoundaryUpdater final : public dycore::TracerBoundaryUpdater {
 public:
  PreparedCaseTracerBoundaryUpdater(AnalysisStateIR analysis, BoundaryCacheIR cache,
                                    PreparedCaseInitConfig config)
      : src_(std::move(analysis), std::move(cache), config) {}
  void set_step_start_time(real t) { step_start_ = std::max(t, 0.0f); }
  void apply(std::vector<state::TracerState>& states,
             const std::vector<domain::SubdomainDescriptor>& layout, real sim_time) override {
    if (src_.empty() || states.empty()) return;
    src_.at(step_start_, sim_time);
    strips::tracers(states, layout, src_.dims(), src_.snapshot(), src_.fields().rho);
  }
  // The snapshot and balanced density at stage time t (device run below).
  bool at(real sim_time, long* second) {
    if (src_.empty()) return false;
    src_.at(step_start_, sim_time);
    *second = src_.second();
    return true;
  }
  const BoundarySnapshotIR& snapshot() const { return src_.snapshot(); }
  const std::vector<real>& balanced_rho() const { return src_.fields().rho; }
  const balance::Dims& dims() const { return src_.dims(); }

 private:
  strips::Source src_;
  real step_start_ = 0.0f;
};

// ---------------------------------------------------------------------------
// The prepared run on the device.  States, tracers and the surface bucket
// are uploaded once into one device domain and stay there between steps;
// each RK stage's lateral strips are applied on the device (STRIPS against
// the boundary reference) after a host callback that refreshes the
// reference only when the stage's forecast second changes: the balanced
// fields of that second are host work, as in the reference, but the stage
// state never leaves the device (the host plug-in path downloads and
// re-uploads it at every stage).  Results equal the host updaters', which
// overwrite the same strips with the same values.
// ---------------------------------------------------------------------------
class PreparedCaseDeviceRun {
 public:
  PreparedCaseDeviceRun(const domain::GridMetrics& metrics,
                        const std::vector<domain::SubdomainDescriptor>& layout,
                        const std::vector<dycore::DryState>& states,
                        const std::vector<state::TracerState>& tracers,
                        const std::vector<physics::WarmRainSurfaceAccumulation>& acc,
                        PreparedCaseBoundaryUpdater* dry_bc, PreparedCaseTracerBoundaryUpdater* tr_bc)
      : layout_(layout), G_(gwm::detail::global_of(layout, metrics)),
        ntr_(tracers.empty() ? 0 : static_cast<int>(tracers.front().size())),
        lease_(metrics.device.get(), ntr_), dry_bc_(dry_bc), tr_bc_(tr_bc) {
    gwm_state* s = lease_;
    dycore::detail::upload_dry(s, states, layout, G_);
    if (ntr_ > 0) dycore::detail::upload_tracers(s, tracers, layout, G_);
    std::vector<real> bucket(static_cast<std::size_t>(G_.nx) * G_.ny, 0.0f);
    for (std::size_t n = 0; n < layout.size() && n < acc.size(); ++n)
      for (int j = 0; j < layout[n].ny_local(); ++j)
        for (int i = 0; i < layout[n].nx_local(); ++i)
          bucket[static_cast<std::size_t>(layout[n].j_begin + j) * G_.nx + layout[n].i_begin + i] =
              acc[n].liquid_precipitation_kg_m2(i, j, 0);
    gwm::detail::check(gwm_state_upload(s, GWM_PRECIP, bucket.data()));
    if (dry_bc_ || tr_bc_) gwm::detail::check(gwm_state_capture_boundary(s));
    if (tr_bc_ && !tracers.empty()) names_ = tracers.front().registry();
  }
  PreparedCaseDeviceRun(const PreparedCaseDeviceRun&) = delete;
  PreparedCaseDeviceRun& operator=(const PreparedCaseDeviceRun&) = delete;

  // advance_dry_state_ssprk3 with the prepared boundary; fast == false runs
  // the step without fast modes (a null FastModeIntegrator).
  void dry_step(const dycore::DryStepperConfig& config, bool fast) {
    const auto c = dycore::detail::cfg_of(config);
    gwm_boundary_config b{dry_bc_ ? GWM_BOUNDARY_STRIPS : GWM_BOUNDARY_NONE, 0.0,
                          dry_bc_ ? &PreparedCaseDeviceRun::dry_refresh : nullptr, this};
    if (fast)
      gwm::detail::check(gwm_ssprk3_step(lease_, &c, &b));
    else
      gwm::detail::check(gwm_ssprk3_step_custom_fast(lease_, &c, &b, &PreparedCaseDeviceRun::no_fast,
                                                     nullptr));
    rethrow();
  }
  // advance_passive_tracers_ssprk3 with the prepared tracer boundary.
  void tracer_step(const dycore::DryStepperConfig& config) {
    if (ntr_ == 0) return;
    const auto c = dycore::detail::cfg_of(config);
    if (tr_bc_)
      gwm::detail::check(gwm_tracers_ssprk3_callback(lease_, &c, GWM_BOUNDARY_STRIPS,
                                                     &PreparedCaseDeviceRun::tracer_refresh, this));
    else
      gwm::detail::check(gwm_tracers_ssprk3(lease_, &c, GWM_BOUNDARY_NONE));
    rethrow();
  }
  // apply_warm_rain_microphysics (layout/metric overload).
  void warm_rain(const physics::WarmRainConfig& config) {
    const auto c = physics::detail::cfg_of(config);
    gwm::detail::check(gwm_warm_rain(lease_, &c, 1));
  }
  // The device state back into the host states, tracers and buckets.
  void download(std::vector<dycore::DryState>& states, std::vector<state::TracerState>& tracers,
                std::vector<physics::WarmRainSurfaceAccumulation>& acc) {
    gwm_state* s = lease_;
    dycore::detail::download_dry(s, states, layout_, G_);
    if (ntr_ > 0) dycore::detail::download_tracers(s, tracers, layout_, G_);
    std::vector<real> bucket(static_cast<std::size_t>(G_.nx) * G_.ny);
    gwm::detail::check(gwm_state_download(s, GWM_PRECIP, bucket.data()));
    for (std::size_t n = 0; n < layout_.size() && n < acc.size(); ++n)
      for (int j = 0; j < layout_[n].ny_local(); ++j)
        for (int i = 0; i < layout_[n].nx_local(); ++i)
          acc[n].liquid_precipitation_kg_m2(i, j, 0) =
              bucket[static_cast<std::size_t>(layout_[n].j_begin + j) * G_.nx + layout_[n].i_begin + i];
  }
  // Boundary-reference refreshes so far (host balance work happened at
  // most this often; the stage strips themselves run on the device).
  int dry_refreshes() const { return dry_uploads_; }
  int tracer_refreshes() const { return tracer_uploads_; }

 private:
  // Global haloless field -> the device storage layout of field f.
  void to_storage(const std::vector<real>& src, int ox, int oy, int oz, std::vector<real>& out) const {
    out.assign(G_.size(ox, oy, oz), 0.0f);
    const int fx = G_.nx + ox, fy = G_.ny + oy, fz = G_.nz + oz;
    for (int k = 0; k < fz; ++k)
      for (int j = 0; j < fy; ++j)
        for (int i = 0; i < fx; ++i)
          out[G_.index(ox, oy, i, j, k)] = src[(static_cast<std::size_t>(k) * fy + j) * fx + i];
  }
  static int no_fast(gwm_state*, double, void*) { return GWM_OK; }
  static int dry_refresh(gwm_state* stage, double t, void* user) {
    auto* self = static_cast<PreparedCaseDeviceRun*>(user);
    try {
      long second = 0;
      const balance::Fields* f = self->dry_bc_->fields_at(static_cast<real>(t), &second);
      if (f == nullptr || (self->have_dry_ && second == self->dry_second_)) return GWM_OK;
      const std::vector<real>* src[5] = {&f->rho, &f->rho_theta, &f->mom_u, &f->mom_v, &f->mom_w};
      std::vector<real> buf;
      for (int q = 0; q < 5; ++q) {
        const auto* o = gwm::detail::kOff[q];
        self->to_storage(*src[q], o[0], o[1], o[2], buf);
        gwm::detail::check(gwm_state_set_boundary(stage, q, buf.data()));
      }
      self->dry_second_ = second;
      self->have_dry_ = true;
      ++self->dry_uploads_;
      return GWM_OK;
    } catch (...) {
      self->error_ = std::current_exception();
      return GWM_E_INVALID;
    }
  }
  static int tracer_refresh(gwm_state* stage, double t, void* user) {
    auto* self = static_cast<PreparedCaseDeviceRun*>(user);
    try {
      long second = 0;
      unsigned mask = 0;
      if (!self->tr_bc_->at(static_cast<real>(t), &second)) {
        gwm::detail::check(gwm_state_set_tracer_strip_mask(stage, 0u));
        return GWM_OK;
      }
      if (self->have_tr_ && second == self->tr_second_) return GWM_OK;
      const auto& snap = self->tr_bc_->snapshot();
      const auto& rho = self->tr_bc_->balanced_rho();
      std::vector<real> mass(rho.size()), buf;
      for (int q = 0; q < self->ntr_; ++q) {
        const auto it = snap.atmosphere.values.find(self->names_.at(q).name);
        if (it == snap.atmosphere.values.end()) continue;
        for (std::size_t c = 0; c < rho.size(); ++c) mass[c] = rho[c] * it->second[c];
        self->to_storage(mass, 0, 0, 0, buf);
        gwm::detail::check(gwm_state_set_boundary(stage, GWM_TRACER0 + q, buf.data()));
        if (q < 32) mask |= 1u << q;
      }
      gwm::detail::check(gwm_state_set_tracer_strip_mask(stage, mask));
      self->tr_second_ = second;
      self->have_tr_ = true;
      ++self->tracer_uploads_;
      return GWM_OK;
    } catch (...) {
      self->error_ = std::current_exception();
      return GWM_E_INVALID;
    }
  }
  void rethrow() {
    if (error_) {
      auto e = error_;
      error_ = nullptr;
      std::rethrow_exception(e);
    }
  }

  std::vector<domain::SubdomainDescriptor> layout_;
  gwm::detail::Global G_;
  int ntr_;
  gwm::detail::StateLease lease_;
  PreparedCaseBoundaryUpdater* dry_bc_;
  PreparedCaseTracerBoundaryUpdater* tr_bc_;
  state::TracerRegistry names_;
  long dry_second_ = 0, tr_second_ = 0;
  bool have_dry_ = false, have_tr_ = false;
  int dry_uploads_ = 0, tracer_uploads_ = 0;
  std::exception_ptr error_;
};

}  // namespace gwm::ingest

// ===========================================================================
// 7. Runtime summary (include/gwm/core/runtime_summary.hpp:12-55; the
// quantities of src/core/runtime_summary.cpp:27-178, the JSON keys of
// runtime_summary.cpp:180-251)
// ===========================================================================
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

// Dry summary, per tracer the FP64 column path sum of rho q dz and the
// range of q = rho q / max(rho, 1e-12), the water paths, and (with
// accumulations) the surface precipitation statistics.
inline RuntimeStateSummary summarize_runtime_state(
    const std::vector<dycore::DryState>& states, const std::vector<state::TracerState>& tracers,
    const std::vector<domain::SubdomainDescriptor>& layout, const domain::GridMetrics& metrics,
    const std::vector<physics::WarmRainSurfaceAccumulation>* accumulations = nullptr) {
  if (states.size() != tracers.size() || states.size() != layout.size())
    throw std::runtime_error("runtime summary: states, tracers and layout disagree in rank count");
  RuntimeStateSummary out;
  out.dry = dycore::summarize_dry_states(states);
  if (states.empty() || tracers.empty()) return out;
  const auto& reg = tracers.front().registry();
  const std::size_t nq = reg.size();
  std::vector<double> path(nq, 0.0), qmin(nq, std::numeric_limits<double>::max()),
      qmax(nq, -std::numeric_limits<double>::max());
  bool cells = false;
  for (std::size_t r = 0; r < states.size(); ++r) {
    const auto& L = layout[r];
    const auto& tr = tracers[r];
    if (tr.size() != nq) throw std::runtime_error("runtime summary: tracer sets differ by rank");
    for (std::size_t q = 0; q < nq; ++q)
      if (tr.registry().at(static_cast<int>(q)).name != reg.at(static_cast<int>(q)).name)
        throw std::runtime_error("runtime summary: tracer order differs by rank");
    const auto& rho = states[r].rho_d;
    for (int k = 0; k < L.nz; ++k)
      for (int j = 0; j < L.ny_local(); ++j)
        for (int i = 0; i < L.nx_local(); ++i) {
          cells = true;
          const double dz =
              1.0 / static_cast<double>(metrics.inv_dz_cell(L.i_begin + i, L.j_begin + j, k));
          const double rr = std::max(static_cast<double>(rho(i, j, k)), 1.0e-12);
          for (std::size_t q = 0; q < nq; ++q) {
            const double m = static_cast<double>(tr.mass(static_cast<int>(q))(i, j, k));
            path[q] += m * dz;
            qmin[q] = std::min(qmin[q], m / rr);
            qmax[q] = std::max(qmax[q], m / rr);
          }
        }
  }
  for (std::size_t q = 0; q < nq; ++q) {
    const auto& spec = reg.at(static_cast<int>(q));
    TracerSummary t;
    t.name = spec.name;
    t.units = spec.units;
    t.positive = spec.positive;
    t.column_integrated_sum_kg_m2 = path[q];
    t.mixing_ratio = cells ? dycore::ScalarRange{qmin[q], qmax[q]} : dycore::ScalarRange{0.0, 0.0};
    out.tracers.push_back(std::move(t));
  }
  auto& mo = out.moisture;
  for (const auto& t : out.tracers) {
    if (t.name == state::kSpecificHumidityTracerName) mo.vapor_water_path_sum_kg_m2 = t.column_integrated_sum_kg_m2;
    if (t.name == state::kCloudWaterTracerName) mo.cloud_water_path_sum_kg_m2 = t.column_integrated_sum_kg_m2;
    if (t.name == state::kRainWaterTracerName) mo.rain_water_path_sum_kg_m2 = t.column_integrated_sum_kg_m2;
  }
  mo.condensed_water_path_sum_kg_m2 = mo.cloud_water_path_sum_kg_m2 + mo.rain_water_path_sum_kg_m2;
  mo.total_water_path_sum_kg_m2 = mo.vapor_water_path_sum_kg_m2 + mo.condensed_water_path_sum_kg_m2;
  if (accumulations) {
    const auto w = physics::summarize_warm_rain(tracers, layout, metrics, accumulations);
    mo.accumulated_surface_precipitation_sum_mm = w.accumulated_surface_precipitation_sum_mm;
    mo.mean_surface_precipitation_mm = w.mean_surface_precipitation_mm;
    mo.max_surface_precipitation_mm = w.max_surface_precipitation_mm;
    double wet_sum = 0.0;
    for (const auto& acc : *accumulations) {
      mo.total_surface_cell_count += acc.nx * acc.ny;
      for (int j = 0; j < acc.ny; ++j)
        for (int i = 0; i < acc.nx; ++i) {
          const double v = acc.liquid_precipitation_kg_m2(i, j, 0);
          if (v > 0.0) {
            mo.precipitating_surface_cell_count += 1;
            wet_sum += v;
          }
        }
    }
    if (mo.total_surface_cell_count > 0)
      mo.precipitating_surface_fraction = static_cast<double>(mo.precipitating_surface_cell_count) /
                                          mo.total_surface_cell_count;
    if (mo.precipitating_surface_cell_count > 0)
      mo.mean_precipitating_surface_precipitation_mm = wet_sum / mo.precipitating_surface_cell_count;
  }
  return out;
}

namespace detail {
// A small JSON object writer with a fixed indent.
class JsonObject {
 public:
  JsonObject(std::ostringstream& o, std::string indent) : o_(o), in_(std::move(indent)) {
    o_ << "{\n";
  }
  template <class V>
  JsonObject& num(const char* key, V v) {
    sep();
    o_ << in_ << "\"" << key << "\": " << v;
    return *this;
  }
  JsonObject& flag(const char* key, bool v) {
    sep();
    o_ << in_ << "\"" << key << "\": " << (v ? "true" : "false");
    return *this;
  }
  JsonObject& text(const char* key, const std::string& v) {
    sep();
    o_ << in_ << "\"" << key << "\": \"" << v << "\"";
    return *this;
  }
  JsonObject& range(const char* key, const dycore::ScalarRange& r) {
    sep();
    o_ << in_ << "\"" << key << "\": {\"min\": " << r.min << ", \"max\": " << r.max << "}";
    return *this;
  }
  // Opens a nested object under `key`; the caller closes it with close().
  JsonObject open(const char* key) {
    sep();
    o_ << in_ << "\"" << key << "\": ";
    return JsonObject(o_, in_ + "  ");
  }
  void close(const std::string& outer_indent) {
    o_ << "\n" << outer_indent << "}";
  }

 private:
  void sep() {
    if (!first_) o_ << ",\n";
    first_ = false;
  }
  std::ostringstream& o_;
  std::string in_;
  bool first_ = true;
};
}  // namespace detail

inline std::string runtime_state_summary_to_json(const RuntimeStateSummary& s,
                                                 const std::string& in = "  ") {
  std::ostringstream o;
  o << std::setprecision(17);
  const std::string outer = in.size() >= 2 ? in.substr(0, in.size() - 2) : std::string();
  detail::JsonObject top(o, in);
  const auto& d = s.dry;
  top.num("total_dry_mass", d.total_dry_mass)
      .num("total_rho_theta_m", d.total_rho_theta_m)
      .num("total_horizontal_momentum_x", d.total_horizontal_momentum_x)
      .num("total_horizontal_momentum_y", d.total_horizontal_momentum_y)
      .num("total_vertical_momentum", d.total_vertical_momentum)
      .range("rho_d", d.rho_d)
      .range("theta_m", d.theta_m)
      .range("u_face", d.u_face)
      .range("v_face", d.v_face)
      .range("w_face", d.w_face);
  const auto& m = s.moisture;
  auto mo = top.open("moisture");
  mo.num("vapor_water_path_sum_kg_m2", m.vapor_water_path_sum_kg_m2)
      .num("cloud_water_path_sum_kg_m2", m.cloud_water_path_sum_kg_m2)
      .num("rain_water_path_sum_kg_m2", m.rain_water_path_sum_kg_m2)
      .num("condensed_water_path_sum_kg_m2", m.condensed_water_path_sum_kg_m2)
      .num("total_water_path_sum_kg_m2", m.total_water_path_sum_kg_m2)
      .num("accumulated_surface_precipitation_sum_mm", m.accumulated_surface_precipitation_sum_mm)
      .num("mean_surface_precipitation_mm", m.mean_surface_precipitation_mm)
      .num("max_surface_precipitation_mm", m.max_surface_precipitation_mm)
      .num("total_surface_cell_count", m.total_surface_cell_count)
      .num("precipitating_surface_cell_count", m.precipitating_surface_cell_count)
      .num("precipitating_surface_fraction", m.precipitating_surface_fraction)
      .num("mean_precipitating_surface_precipitation_mm",
           m.mean_precipitating_surface_precipitation_mm);
  mo.close(in);
  auto tr = top.open("tracers");
  for (const auto& t : s.tracers) {
    auto one = tr.open(t.name.c_str());
    one.text("units", t.units)
        .flag("positive", t.positive)
        .num("column_integrated_sum_kg_m2", t.column_integrated_sum_kg_m2)
        .range("mixing_ratio", t.mixing_ratio);
    one.close(in + "  ");
  }
  tr.close(in);
  top.close(outer);
  return o.str();
}

}  // namespace gwm::core
