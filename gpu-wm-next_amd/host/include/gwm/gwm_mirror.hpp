// gwm_mirror.hpp -- C++ host mirror of the reference's hot-path API on top of
// the C-ABI in include/gwm_hip.h.
//
// Restores the reference's names, argument meanings and exception behaviour
// (FahrenheitResearch/gpu-wm-next include/gwm/**), so code written against
// the reference compiles against this header set:
//   state::Field3D / FaceField (include/gwm/state/field3d.hpp, face_field.hpp)
//   domain::GridMetrics, SubdomainDescriptor, build_rectilinear_domain
//   comm::VirtualRankLayout, HaloExchange, find_cartesian_neighbors
//   dycore::DryState, DryStepperConfig, BoundaryUpdater, FastModeIntegrator,
//     compute_slow_tendencies, advance_dry_state_ssprk3,
//     apply_local_split_explicit_fast_modes, idealized cases, diagnostics,
//     advance_passive_tracers_ssprk3
//   physics::WarmRainConfig, apply_warm_rain_microphysics
// Fields live in host memory (the reference's managed memory was host-
// indexable); each operator gathers the virtual-rank layout into one device
// domain, runs the HIP path and scatters back, so the state is host-readable
// after every call.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <functional>
#include <memory>
#include <optional>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "gwm_hip.h"

namespace gwm {

using real = float;
using index_t = int;

struct Int2 {
  int x = 0;
  int y = 0;
};
struct Int3 {
  int x = 0;
  int y = 0;
  int z = 0;
};

inline void require(bool condition, const std::string& message) {
  if (!condition) throw std::runtime_error(message);
}

namespace detail {
inline void check(int rc) {
  if (rc != GWM_OK) throw std::runtime_error(std::string("gwm_hip: ") + gwm_last_error());
}
// One device context per process (FP32, the reference's `real`).
struct Device {
  gwm_ctx* ctx = nullptr;
  Device() {
    const char* d = std::getenv("GWM_DEVICE");
    check(gwm_ctx_create(d ? std::atoi(d) : 0, GWM_F32, &ctx));
  }
  ~Device() {
    if (ctx) gwm_ctx_destroy(ctx);
  }
  static Device& get() {
    static Device dev;
    return dev;
  }
};
// Cached device states per (grid, tracer count); destroyed before the grid.
// A state is leased for the duration of one operator call, so an operator
// called from inside another (a host plug-in) gets a state of its own.
struct StateCache {
  struct Entry {
    const gwm_grid* grid;
    int ntr;
    gwm_state* state;
    bool in_use;
  };
  std::vector<Entry> entries;
  ~StateCache() { clear(); }
  void clear() {
    for (auto& e : entries) gwm_state_destroy(e.state);
    entries.clear();
  }
  gwm_state* acquire(const gwm_grid* g, int ntr) {
    for (auto& e : entries)
      if (e.grid == g && e.ntr == ntr && !e.in_use) {
        e.in_use = true;
        return e.state;
      }
    gwm_state* s = nullptr;
    check(gwm_state_create(Device::get().ctx, g, ntr, &s));
    entries.push_back(Entry{g, ntr, s, true});
    return s;
  }
  void release(gwm_state* s) {
    for (auto& e : entries)
      if (e.state == s) e.in_use = false;
  }
  void drop(const gwm_grid* g) {
    for (auto it = entries.begin(); it != entries.end();) {
      if (it->grid == g) {
        gwm_state_destroy(it->state);
        it = entries.erase(it);
      } else {
        ++it;
      }
    }
  }
  static StateCache& get() {
    Device::get();  // the context outlives the cache
    static StateCache cache;
    return cache;
  }
};

class StateLease {
 public:
  StateLease(const gwm_grid* g, int ntr) : s_(StateCache::get().acquire(g, ntr)) {}
  ~StateLease() { StateCache::get().release(s_); }
  StateLease(const StateLease&) = delete;
  StateLease& operator=(const StateLease&) = delete;
  operator gwm_state*() const { return s_; }

 private:
  gwm_state* s_;
};

}  // namespace detail

// --------------------------------------------------------------------------
// state
// --------------------------------------------------------------------------
namespace state {

template <class T>
class Field3D {
 public:
  Field3D() = default;
  Field3D(int nx, int ny, int nz, int halo, std::string label = {})
      : nx_(nx), ny_(ny), nz_(nz), halo_(halo), label_(std::move(label)),
        data_(static_cast<size_t>(nx + 2 * halo) * (ny + 2 * halo) * nz, T{}) {}
  int nx() const { return nx_; }
  int ny() const { return ny_; }
  int nz() const { return nz_; }
  int halo() const { return halo_; }
  int pitch_x() const { return nx_ + 2 * halo_; }
  int pitch_y() const { return ny_ + 2 * halo_; }
  size_t total_size() const { return data_.size(); }
  T* data() { return data_.data(); }
  const T* data() const { return data_.data(); }
  size_t storage_index(int i, int j, int k) const {
    return (static_cast<size_t>(k) * pitch_y() + static_cast<size_t>(j + halo_)) *
               static_cast<size_t>(pitch_x()) +
           static_cast<size_t>(i + halo_);
  }
  T& operator()(int i, int j, int k) { return data_[storage_index(i, j, k)]; }
  const T& operator()(int i, int j, int k) const { return data_[storage_index(i, j, k)]; }
  void fill(T v) { std::fill(data_.begin(), data_.end(), v); }
  void copy_all_from(const Field3D& o) {
    require(nx_ == o.nx_ && ny_ == o.ny_ && nz_ == o.nz_ && halo_ == o.halo_,
            "Field3D shape mismatch in copy_all_from");
    data_ = o.data_;
  }
  Field3D clone_empty_like(const std::string& suffix = "_tmp") const {
    return Field3D(nx_, ny_, nz_, halo_, label_ + suffix);
  }
  double owned_sum() const {
    double s = 0.0;
    for (int k = 0; k < nz_; ++k)
      for (int j = 0; j < ny_; ++j)
        for (int i = 0; i < nx_; ++i) s += static_cast<double>((*this)(i, j, k));
    return s;
  }
  void sync() const {}

 private:
  int nx_ = 0, ny_ = 0, nz_ = 0, halo_ = 0;
  std::string label_;
  std::vector<T> data_;
};

enum class FaceOrientation { X, Y, Z };

template <class T>
class FaceField {
 public:
  FaceField() = default;
  FaceField(int nx, int ny, int nz, int halo, FaceOrientation o, std::string label = {})
      : orientation_(o),
        storage_(nx + (o == FaceOrientation::X), ny + (o == FaceOrientation::Y),
                 nz + (o == FaceOrientation::Z), halo, std::move(label)) {}
  FaceOrientation orientation() const { return orientation_; }
  Field3D<T>& storage() { return storage_; }
  const Field3D<T>& storage() const { return storage_; }

 private:
  FaceOrientation orientation_ = FaceOrientation::X;
  Field3D<T> storage_;
};

struct TracerSpec {
  std::string name;
  std::string units;
  bool positive = true;
};

inline constexpr const char* kSpecificHumidityTracerName = "specific_humidity";
inline constexpr const char* kCloudWaterTracerName = "cloud_water_mixing_ratio";
inline constexpr const char* kRainWaterTracerName = "rain_water_mixing_ratio";

class TracerRegistry {
 public:
  int add(TracerSpec spec) {
    require(!find(spec.name).has_value(), "Tracer already exists in registry");
    specs_.push_back(std::move(spec));
    return static_cast<int>(specs_.size()) - 1;
  }
  std::optional<int> find(const std::string& name) const {
    for (size_t i = 0; i < specs_.size(); ++i)
      if (specs_[i].name == name) return static_cast<int>(i);
    return std::nullopt;
  }
  const TracerSpec& at(int i) const { return specs_.at(static_cast<size_t>(i)); }
  size_t size() const { return specs_.size(); }

 private:
  std::vector<TracerSpec> specs_;
};

inline TracerRegistry make_specific_humidity_registry() {
  TracerRegistry r;
  r.add({kSpecificHumidityTracerName, "kg/kg", true});
  return r;
}
inline TracerRegistry make_warm_rain_registry() {
  TracerRegistry r;
  r.add({kSpecificHumidityTracerName, "kg/kg", true});
  r.add({kCloudWaterTracerName, "kg/kg", true});
  r.add({kRainWaterTracerName, "kg/kg", true});
  return r;
}

class TracerState {
 public:
  TracerState() = default;
  TracerState(TracerRegistry registry, int nx, int ny, int nz, int halo, std::string label = {})
      : registry_(std::move(registry)) {
    for (size_t t = 0; t < registry_.size(); ++t) masses_.emplace_back(nx, ny, nz, halo, label);
  }
  int nx() const { return masses_.empty() ? 0 : masses_.front().nx(); }
  int ny() const { return masses_.empty() ? 0 : masses_.front().ny(); }
  int nz() const { return masses_.empty() ? 0 : masses_.front().nz(); }
  int halo() const { return masses_.empty() ? 0 : masses_.front().halo(); }
  const TracerRegistry& registry() const { return registry_; }
  size_t size() const { return masses_.size(); }
  std::optional<int> find(const std::string& n) const { return registry_.find(n); }
  Field3D<real>& mass(int i) { return masses_.at(static_cast<size_t>(i)); }
  const Field3D<real>& mass(int i) const { return masses_.at(static_cast<size_t>(i)); }
  Field3D<real>& mass(const std::string& n) {
    const auto i = registry_.find(n);
    require(i.has_value(), "Tracer not found in TracerState: " + n);
    return mass(*i);
  }
  const Field3D<real>& mass(const std::string& n) const {
    const auto i = registry_.find(n);
    require(i.has_value(), "Tracer not found in TracerState: " + n);
    return mass(*i);
  }
  void fill_zero() {
    for (auto& f : masses_) f.fill(0.0f);
  }
  double total_mass(int i) const { return mass(i).owned_sum(); }

 private:
  TracerRegistry registry_;
  std::vector<Field3D<real>> masses_;
};

}  // namespace state

// --------------------------------------------------------------------------
// domain
// --------------------------------------------------------------------------
namespace domain {

struct SubdomainDescriptor {
  int rank = 0, coord_x = 0, coord_y = 0, ranks_x = 1, ranks_y = 1;
  int global_nx = 0, global_ny = 0, nz = 0, halo = 0;
  int i_begin = 0, i_end = 0, j_begin = 0, j_end = 0;
  bool periodic_x = true, periodic_y = true;
  int nx_local() const { return i_end - i_begin; }
  int ny_local() const { return j_end - j_begin; }
};

// GridMetrics (include/gwm/domain/grid_metrics.hpp:12-240): the metric fields
// are built on the device by gwm_grid_create and mirrored on the host.
struct GridMetrics {
  int nx = 0, ny = 0, nz = 0, halo = 1;
  double dx = 0.0, dy = 0.0, z_top = 0.0, dz_nominal = 0.0, terrain_taper_eta = 0.25;
  bool periodic_x = true, periodic_y = true;
  std::vector<double> z_interfaces, z_centers, terrain_weights_interfaces,
      terrain_weights_centers;
  state::Field3D<real> terrain, terrain_slope_x, terrain_slope_y, z_interfaces_metric,
      z_centers_metric, inv_dz_cell_metric, inv_dz_face_metric;
  std::shared_ptr<gwm_grid> device;

  real terrain_height(int i, int j) const { return terrain(i, j, 0); }
  real z_interface(int i, int j, int k) const { return z_interfaces_metric(i, j, k); }
  real z_center(int i, int j, int k) const { return z_centers_metric(i, j, k); }
  real inv_dz_cell(int i, int j, int k) const { return inv_dz_cell_metric(i, j, k); }
  real inv_dz_face(int i, int j, int k) const { return inv_dz_face_metric(i, j, k); }
  double flat_z_interface(int k) const { return z_interfaces[static_cast<size_t>(k)]; }
  double flat_z_center(int k) const { return z_centers[static_cast<size_t>(k)]; }
  double terrain_weight_center(int k) const { return terrain_weights_centers[static_cast<size_t>(k)]; }

  // Extension: explicit (stretched) flat interfaces; pass an empty vector
  // for the reference's uniform eta levels.
  static GridMetrics make_hybrid_height(int nx, int ny, int nz, double dx, double dy, double z_top,
                                        const std::vector<real>& terrain_dyn,
                                        double terrain_taper_eta = 0.25, bool periodic_x = true,
                                        bool periodic_y = true,
                                        const std::vector<double>& z_levels = {}, int halo = 1) {
    require(static_cast<int>(terrain_dyn.size()) == nx * ny,
            "terrain_dyn shape mismatch in GridMetrics::make_hybrid_height");
    GridMetrics m;
    m.nx = nx;
    m.ny = ny;
    m.nz = nz;
    m.halo = halo;
    m.dx = dx;
    m.dy = dy;
    m.z_top = z_levels.empty() ? z_top : z_levels.back();
    m.dz_nominal = m.z_top / static_cast<double>(nz);
    m.terrain_taper_eta = terrain_taper_eta;
    m.periodic_x = periodic_x;
    m.periodic_y = periodic_y;
    gwm_grid_desc d{nx, ny, nz, halo, periodic_x ? 1 : 0, periodic_y ? 1 : 0, dx, dy, z_top,
                    terrain_taper_eta};
    gwm_grid* g = nullptr;
    detail::check(gwm_grid_create(detail::Device::get().ctx, &d,
                                  z_levels.empty() ? nullptr : z_levels.data(),
                                  terrain_dyn.data(), &g));
    m.device = std::shared_ptr<gwm_grid>(g, [](gwm_grid* p) {
      detail::StateCache::get().drop(p);
      gwm_grid_destroy(p);
    });
    m.z_interfaces.resize(nz + 1);
    m.z_centers.resize(nz);
    m.terrain_weights_interfaces.resize(nz + 1);
    m.terrain_weights_centers.resize(nz);
    detail::check(gwm_grid_download_metric(g, 7, m.z_interfaces.data()));
    detail::check(gwm_grid_download_metric(g, 8, m.z_centers.data()));
    detail::check(gwm_grid_download_metric(g, 9, m.terrain_weights_interfaces.data()));
    detail::check(gwm_grid_download_metric(g, 10, m.terrain_weights_centers.data()));
    auto field = [&](int which, int fz) {
      state::Field3D<real> f(nx, ny, fz, 0);
      detail::check(gwm_grid_download_metric(g, which, f.data()));
      return f;
    };
    m.z_interfaces_metric = field(0, nz + 1);
    m.z_centers_metric = field(1, nz);
    m.inv_dz_cell_metric = field(2, nz);
    m.inv_dz_face_metric = field(3, nz + 1);
    m.terrain = field(4, 1);
    m.terrain_slope_x = field(5, 1);
    m.terrain_slope_y = field(6, 1);
    return m;
  }
};

enum class TerrainProfileKind { Flat, CosineMountain };

struct RectilinearDomainConfig {
  int nx = 0, ny = 0, nz = 0, halo = 1, ranks_x = 1, ranks_y = 1;
  double dx = 1000.0, dy = 1000.0, z_top = 16000.0, terrain_taper_eta = 0.25;
  TerrainProfileKind terrain_kind = TerrainProfileKind::Flat;
  double mountain_center_x = 0.5, mountain_center_y = 0.5;
  double mountain_half_width_x = 0.15, mountain_half_width_y = 0.15;
  double mountain_height = 0.0;
  bool periodic_x = true, periodic_y = true;
};

struct IdealizedDomain {
  GridMetrics metrics;
  std::vector<SubdomainDescriptor> layout;
  std::vector<real> terrain_dyn;
  real terrain(int i, int j) const {
    return terrain_dyn[static_cast<size_t>(j) * metrics.nx + static_cast<size_t>(i)];
  }
};

}  // namespace domain

// --------------------------------------------------------------------------
// comm
// --------------------------------------------------------------------------
namespace comm {

struct CartesianNeighborRanks {
  int west = -1, east = -1, south = -1, north = -1;
};

inline int rank_at(const std::vector<domain::SubdomainDescriptor>& L, int cx, int cy, bool px,
                   bool py) {
  if (L.empty()) return -1;
  const int rx = L.front().ranks_x, ry = L.front().ranks_y;
  if (px) cx = ((cx % rx) + rx) % rx;
  else if (cx < 0 || cx >= rx) return -1;
  if (py) cy = ((cy % ry) + ry) % ry;
  else if (cy < 0 || cy >= ry) return -1;
  for (const auto& c : L)
    if (c.coord_x == cx && c.coord_y == cy) return c.rank;
  return -1;
}

// find_cartesian_neighbors (src/comm/cartesian_topology.cpp:57-74)
inline CartesianNeighborRanks find_cartesian_neighbors(
    const std::vector<domain::SubdomainDescriptor>& L, const domain::SubdomainDescriptor& d) {
  CartesianNeighborRanks n;
  n.west = rank_at(L, d.coord_x - 1, d.coord_y, d.periodic_x, d.periodic_y);
  n.east = rank_at(L, d.coord_x + 1, d.coord_y, d.periodic_x, d.periodic_y);
  n.south = rank_at(L, d.coord_x, d.coord_y - 1, d.periodic_x, d.periodic_y);
  n.north = rank_at(L, d.coord_x, d.coord_y + 1, d.periodic_x, d.periodic_y);
  return n;
}

class VirtualRankLayout {
 public:
  // VirtualRankLayout::build (src/comm/virtual_rank_layout.cpp:7-47)
  static std::vector<domain::SubdomainDescriptor> build(int gnx, int gny, int nz, int halo,
                                                        int rx, int ry, bool px = true,
                                                        bool py = true) {
    std::vector<domain::SubdomainDescriptor> L;
    const int bx = gnx / rx, remx = gnx % rx, by = gny / ry, remy = gny % ry;
    int jb = 0;
    for (int y = 0; y < ry; ++y) {
      const int nyl = by + (y < remy ? 1 : 0);
      int ib = 0;
      for (int x = 0; x < rx; ++x) {
        const int nxl = bx + (x < remx ? 1 : 0);
        domain::SubdomainDescriptor d;
        d.rank = static_cast<int>(L.size());
        d.coord_x = x;
        d.coord_y = y;
        d.ranks_x = rx;
        d.ranks_y = ry;
        d.global_nx = gnx;
        d.global_ny = gny;
        d.nz = nz;
        d.halo = halo;
        d.i_begin = ib;
        d.i_end = ib + nxl;
        d.j_begin = jb;
        d.j_end = jb + nyl;
        d.periodic_x = px;
        d.periodic_y = py;
        L.push_back(d);
        ib += nxl;
      }
      jb += nyl;
    }
    return L;
  }
  static std::vector<state::Field3D<real>> scatter_scalar(
      const std::vector<domain::SubdomainDescriptor>& L,
      const std::function<real(int, int, int)>& init, const std::string& label) {
    std::vector<state::Field3D<real>> out;
    for (const auto& d : L) {
      state::Field3D<real> f(d.nx_local(), d.ny_local(), d.nz, d.halo, label);
      for (int k = 0; k < d.nz; ++k)
        for (int j = 0; j < d.ny_local(); ++j)
          for (int i = 0; i < d.nx_local(); ++i) f(i, j, k) = init(d.i_begin + i, d.j_begin + j, k);
      out.push_back(std::move(f));
    }
    return out;
  }
  static state::Field3D<real> gather_scalar(const std::vector<state::Field3D<real>>& fields,
                                            const std::vector<domain::SubdomainDescriptor>& L) {
    require(fields.size() == L.size() && !L.empty(), "Field/layout size mismatch in gather");
    const auto& d0 = L.front();
    state::Field3D<real> g(d0.global_nx, d0.global_ny, d0.nz, d0.halo, "gathered");
    for (size_t n = 0; n < L.size(); ++n)
      for (int k = 0; k < L[n].nz; ++k)
        for (int j = 0; j < L[n].ny_local(); ++j)
          for (int i = 0; i < L[n].nx_local(); ++i)
            g(L[n].i_begin + i, L[n].j_begin + j, k) = fields[n](i, j, k);
    return g;
  }
};

namespace detail {
inline std::vector<int> desc13(const std::vector<domain::SubdomainDescriptor>& L) {
  std::vector<int> v;
  for (const auto& d : L) {
    const int a[13] = {d.rank, d.coord_x, d.coord_y, d.ranks_x, d.ranks_y, d.global_nx,
                       d.global_ny, d.nz, d.halo, d.i_begin, d.i_end, d.j_begin, d.j_end};
    v.insert(v.end(), a, a + 13);
  }
  return v;
}
// Runs the device multi-rank exchange on host fields.
inline void exchange(std::vector<state::Field3D<real>*>& F,
                     const std::vector<domain::SubdomainDescriptor>& L, int orient) {
  require(F.size() == L.size(), "Field/layout size mismatch in exchange");
  if (F.empty()) return;
  gwm_ctx* ctx = gwm::detail::Device::get().ctx;
  std::vector<void*> dev(F.size());
  for (size_t n = 0; n < F.size(); ++n) {
    const size_t bytes = F[n]->total_size() * sizeof(real);
    gwm::detail::check(gwm_device_alloc(ctx, bytes, &dev[n]));
    gwm::detail::check(gwm_memcpy_h2d(ctx, dev[n], F[n]->data(), bytes));
  }
  const auto d = desc13(L);
  const int rc = gwm_halo_exchange_layout(ctx, static_cast<int>(L.size()), d.data(),
                                          L.front().periodic_x, L.front().periodic_y, orient,
                                          dev.data());
  for (size_t n = 0; n < F.size(); ++n) {
    if (rc == GWM_OK)
      gwm::detail::check(gwm_memcpy_d2h(ctx, F[n]->data(), dev[n], F[n]->total_size() * sizeof(real)));
    gwm_device_free(ctx, dev[n]);
  }
  gwm::detail::check(rc);
}
inline int orient_of(state::FaceOrientation o) {
  return o == state::FaceOrientation::X ? 1 : o == state::FaceOrientation::Y ? 2 : 3;
}
}  // namespace detail

// HaloExchange (include/gwm/comm/halo_exchange.hpp:24-92), virtual-rank path.
class HaloExchange {
 public:
  static void exchange_scalar(std::vector<state::Field3D<real>>& fields,
                              const std::vector<domain::SubdomainDescriptor>& layout) {
    std::vector<state::Field3D<real>*> p;
    for (auto& f : fields) p.push_back(&f);
    detail::exchange(p, layout, 0);
  }
  static void exchange_face(std::vector<state::FaceField<real>>& fields,
                            const std::vector<domain::SubdomainDescriptor>& layout) {
    std::vector<state::FaceField<real>*> p;
    for (auto& f : fields) p.push_back(&f);
    exchange_face(p, layout);
  }
  static void exchange_face(const std::vector<state::FaceField<real>*>& fields,
                            const std::vector<domain::SubdomainDescriptor>& layout) {
    if (fields.empty()) return;
    const auto o = fields.front()->orientation();
    std::vector<state::Field3D<real>*> p;
    for (auto* f : fields) {
      require(f != nullptr, "Null field pointer in exchange_face");
      require(f->orientation() == o, "Face orientation mismatch in exchange_face");
      p.push_back(&f->storage());
    }
    detail::exchange(p, layout, detail::orient_of(o));
  }
  // synchronize_face_owned_interfaces_impl (halo_exchange.cpp:217-285)
  static void synchronize_owned_face_interfaces(std::vector<state::FaceField<real>>& fields,
                                                const std::vector<domain::SubdomainDescriptor>& L) {
    require(fields.size() == L.size(), "Field/layout size mismatch");
    if (fields.empty() || fields.front().orientation() == state::FaceOrientation::Z) return;
    const auto o = fields.front().orientation();
    for (size_t n = 0; n < L.size(); ++n) {
      const auto nb = find_cartesian_neighbors(L, L[n]);
      auto& s = fields[n].storage();
      if (o == state::FaceOrientation::X && nb.east >= 0 &&
          (static_cast<int>(n) < nb.east || nb.east == static_cast<int>(n))) {
        auto& p = fields[static_cast<size_t>(nb.east)].storage();
        for (int k = 0; k < s.nz(); ++k)
          for (int j = 0; j < L[n].ny_local(); ++j) {
            const real a = 0.5f * (s(L[n].nx_local(), j, k) + p(0, j, k));
            s(L[n].nx_local(), j, k) = a;
            p(0, j, k) = a;
          }
      }
      if (o == state::FaceOrientation::Y && nb.north >= 0 &&
          (static_cast<int>(n) < nb.north || nb.north == static_cast<int>(n))) {
        auto& p = fields[static_cast<size_t>(nb.north)].storage();
        for (int k = 0; k < s.nz(); ++k)
          for (int i = 0; i < L[n].nx_local(); ++i) {
            const real a = 0.5f * (s(i, L[n].ny_local(), k) + p(i, 0, k));
            s(i, L[n].ny_local(), k) = a;
            p(i, 0, k) = a;
          }
      }
    }
  }
};

}  // namespace comm

namespace domain {
// build_rectilinear_domain (src/domain/idealized_domain_builder.cpp:39-64)
inline IdealizedDomain build_rectilinear_domain(const RectilinearDomainConfig& c) {
  require(c.nx > 0 && c.ny > 0 && c.nz > 0, "RectilinearDomainConfig dimensions must be positive");
  require(c.dx > 0.0 && c.dy > 0.0 && c.z_top > 0.0,
          "RectilinearDomainConfig spacing/top must be positive");
  require(c.ranks_x > 0 && c.ranks_y > 0, "RectilinearDomainConfig rank counts must be positive");
  IdealizedDomain d;
  d.terrain_dyn.assign(static_cast<size_t>(c.nx) * c.ny, 0.0f);
  if (c.terrain_kind == TerrainProfileKind::CosineMountain)
    gwm::detail::check(gwm_terrain_cosine_mountain(c.nx, c.ny, c.mountain_center_x,
                                                   c.mountain_center_y, c.mountain_half_width_x,
                                                   c.mountain_half_width_y, c.mountain_height,
                                                   d.terrain_dyn.data()));
  d.metrics = GridMetrics::make_hybrid_height(c.nx, c.ny, c.nz, c.dx, c.dy, c.z_top, d.terrain_dyn,
                                              c.terrain_taper_eta, c.periodic_x, c.periodic_y, {},
                                              c.halo);
  d.layout = comm::VirtualRankLayout::build(c.nx, c.ny, c.nz, c.halo, c.ranks_x, c.ranks_y,
                                            c.periodic_x, c.periodic_y);
  return d;
}
}  // namespace domain

}  // namespace gwm
