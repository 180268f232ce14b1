// gwm_dycore.hpp -- dycore and warm-rain half of the C++ host mirror.
//
// Reference declarations mirrored here (FahrenheitResearch/gpu-wm-next):
//   include/gwm/dycore/dry_core.hpp:14-112, dry_fast_modes.hpp:9-13,
//   dry_diagnostics.hpp:9-34, idealized_cases.hpp:12-62,
//   passive_tracer.hpp:12-52, boundary_conditions.hpp:9-26,
//   include/gwm/physics/warm_rain.hpp:14-82.
//
// Every operator gathers the per-rank host fields into one device domain,
// runs the HIP path of libgwm_hip.so and scatters the result back, halos
// filled with the virtual-rank exchange semantics, so host code can index
// the state right after the call.  Device states are cached per grid.
#pragma once

#include <exception>
#include <iomanip>
#include <limits>
#include <tuple>

#include "gwm/gwm_mirror.hpp"

namespace gwm {
namespace detail {

// Shape of the device domain a layout tiles.
struct Global {
  int nx = 0, ny = 0, nz = 0, h = 0;
  size_t size(int ox, int oy, int oz) const {
    return static_cast<size_t>(nx + ox + 2 * h) * (ny + oy + 2 * h) * (nz + oz);
  }
  size_t index(int ox, int oy, int i, int j, int k) const {
    return (static_cast<size_t>(k) * (ny + oy + 2 * h) + static_cast<size_t>(j + h)) *
               static_cast<size_t>(nx + ox + 2 * h) +
           static_cast<size_t>(i + h);
  }
};

inline Global global_of(const std::vector<domain::SubdomainDescriptor>& L,
                        const domain::GridMetrics& m) {
  require(!L.empty(), "Layout must not be empty");
  const auto& d = L.front();
  require(d.global_nx == m.nx && d.global_ny == m.ny && d.nz == m.nz,
          "GridMetrics/layout mismatch");
  require(d.halo == m.halo, "Layout halo must match the GridMetrics halo");
  require(m.device != nullptr, "GridMetrics has no device grid (use make_hybrid_height)");
  return Global{m.nx, m.ny, m.nz, m.halo};
}

// Owned elements (faces at both ends included) of every rank -> global.
inline void gather(const std::vector<const state::Field3D<real>*>& loc,
                   const std::vector<domain::SubdomainDescriptor>& L, const Global& G, int ox,
                   int oy, int oz, std::vector<real>& out) {
  out.assign(G.size(ox, oy, oz), 0.0f);
  for (size_t n = 0; n < L.size(); ++n) {
    const auto& d = L[n];
    const auto& f = *loc[n];
    require(f.nx() == d.nx_local() + ox && f.ny() == d.ny_local() + oy && f.nz() == d.nz + oz &&
                f.halo() == G.h,
            "Field shape does not match the layout");
    for (int k = 0; k < d.nz + oz; ++k)
      for (int j = 0; j < d.ny_local() + oy; ++j)
        for (int i = 0; i < d.nx_local() + ox; ++i)
          out[G.index(ox, oy, d.i_begin + i, d.j_begin + j, k)] = f(i, j, k);
  }
}

// Global (halo-filled) -> every rank, ghosts included; `owned_only` skips them.
inline void scatter(const std::vector<real>& in, const std::vector<domain::SubdomainDescriptor>& L,
                    const Global& G, int ox, int oy, int oz,
                    const std::vector<state::Field3D<real>*>& loc, bool owned_only = false) {
  for (size_t n = 0; n < L.size(); ++n) {
    const auto& d = L[n];
    auto& f = *loc[n];
    const int h = owned_only ? 0 : G.h;
    for (int k = 0; k < d.nz + oz; ++k)
      for (int j = -h; j < d.ny_local() + oy + h; ++j)
        for (int i = -h; i < d.nx_local() + ox + h; ++i)
          f(i, j, k) = in[G.index(ox, oy, d.i_begin + i, d.j_begin + j, k)];
  }
}

constexpr int kOff[5][3] = {{0, 0, 0}, {0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};

}  // namespace detail

// --------------------------------------------------------------------------
// dycore
// --------------------------------------------------------------------------
namespace dycore {

struct DryState {
  state::Field3D<real> rho_d;
  state::Field3D<real> rho_theta_m;
  state::FaceField<real> mom_u;
  state::FaceField<real> mom_v;
  state::FaceField<real> mom_w;

  DryState() = default;
  DryState(int nx, int ny, int nz, int halo, const std::string& p)
      : rho_d(nx, ny, nz, halo, p + "_rho_d"),
        rho_theta_m(nx, ny, nz, halo, p + "_rho_theta_m"),
        mom_u(nx, ny, nz, halo, state::FaceOrientation::X, p + "_mom_u"),
        mom_v(nx, ny, nz, halo, state::FaceOrientation::Y, p + "_mom_v"),
        mom_w(nx, ny, nz, halo, state::FaceOrientation::Z, p + "_mom_w") {}
  DryState clone_empty_like(const std::string& suffix = "_tmp") const {
    return DryState(rho_d.nx(), rho_d.ny(), rho_d.nz(), rho_d.halo(),
                    suffix.empty() ? "dry_state" : suffix);
  }
  void copy_all_from(const DryState& o) {
    rho_d.copy_all_from(o.rho_d);
    rho_theta_m.copy_all_from(o.rho_theta_m);
    mom_u.storage().copy_all_from(o.mom_u.storage());
    mom_v.storage().copy_all_from(o.mom_v.storage());
    mom_w.storage().copy_all_from(o.mom_w.storage());
  }
  void fill_constant(real rho, real theta, real u, real v, real w) {
    rho_d.fill(rho);
    rho_theta_m.fill(rho * theta);
    mom_u.storage().fill(rho * u);
    mom_v.storage().fill(rho * v);
    mom_w.storage().fill(rho * w);
  }
  double total_dry_mass() const { return rho_d.owned_sum(); }
  double total_rho_theta_m() const { return rho_theta_m.owned_sum(); }
  double total_vertical_momentum() const { return mom_w.storage().owned_sum(); }

  // Field n of the C-ABI order (GWM_RHO .. GWM_MOM_W).
  state::Field3D<real>& field(int n) {
    switch (n) {
      case GWM_RHO: return rho_d;
      case GWM_RHO_THETA: return rho_theta_m;
      case GWM_MOM_U: return mom_u.storage();
      case GWM_MOM_V: return mom_v.storage();
      default: return mom_w.storage();
    }
  }
  const state::Field3D<real>& field(int n) const { return const_cast<DryState*>(this)->field(n); }
};

struct DrySlowTendencies {
  state::Field3D<real> rho_d;
  state::Field3D<real> rho_theta_m;
  state::FaceField<real> mom_u;
  state::FaceField<real> mom_v;
  state::FaceField<real> mom_w;

  DrySlowTendencies() = default;
  DrySlowTendencies(int nx, int ny, int nz, int halo, const std::string& p)
      : rho_d(nx, ny, nz, halo, p + "_rho_d"),
        rho_theta_m(nx, ny, nz, halo, p + "_rho_theta_m"),
        mom_u(nx, ny, nz, halo, state::FaceOrientation::X, p + "_mom_u"),
        mom_v(nx, ny, nz, halo, state::FaceOrientation::Y, p + "_mom_v"),
        mom_w(nx, ny, nz, halo, state::FaceOrientation::Z, p + "_mom_w") {}
  void fill_zero() {
    rho_d.fill(0.0f);
    rho_theta_m.fill(0.0f);
    mom_u.storage().fill(0.0f);
    mom_v.storage().fill(0.0f);
    mom_w.storage().fill(0.0f);
  }
  state::Field3D<real>& field(int n) {
    switch (n) {
      case GWM_RHO: return rho_d;
      case GWM_RHO_THETA: return rho_theta_m;
      case GWM_MOM_U: return mom_u.storage();
      case GWM_MOM_V: return mom_v.storage();
      default: return mom_w.storage();
    }
  }
};

struct DryStepperConfig {
  real dt = 2.0f;
  int fast_substeps = 1;
  real gravity = 9.81f;
  bool fast_update_horizontal_momentum = true;
  bool fast_update_density = true;
};

class BoundaryUpdater {
 public:
  virtual ~BoundaryUpdater() = default;
  virtual void apply(std::vector<DryState>& states,
                     const std::vector<domain::SubdomainDescriptor>& layout, real sim_time) = 0;
};

class NullBoundaryUpdater final : public BoundaryUpdater {
 public:
  void apply(std::vector<DryState>&, const std::vector<domain::SubdomainDescriptor>&,
             real) override {}
};

class FastModeIntegrator {
 public:
  virtual ~FastModeIntegrator() = default;
  virtual void apply(std::vector<DryState>& states,
                     const std::vector<domain::SubdomainDescriptor>& layout,
                     const domain::GridMetrics& metrics, const DryStepperConfig& config) = 0;
};

void apply_local_split_explicit_fast_modes(std::vector<DryState>& states,
                                           const std::vector<domain::SubdomainDescriptor>& layout,
                                           const domain::GridMetrics& metrics,
                                           const DryStepperConfig& config);

class LocalSplitExplicitFastMode final : public FastModeIntegrator {
 public:
  void apply(std::vector<DryState>& states, const std::vector<domain::SubdomainDescriptor>& layout,
             const domain::GridMetrics& metrics, const DryStepperConfig& config) override {
    apply_local_split_explicit_fast_modes(states, layout, metrics, config);
  }
};

namespace detail {
using gwm::detail::check;
using gwm::detail::Global;

inline gwm_step_config cfg_of(const DryStepperConfig& c) {
  return gwm_step_config{static_cast<double>(c.dt), c.fast_substeps, static_cast<double>(c.gravity),
                         c.fast_update_horizontal_momentum ? 1 : 0, c.fast_update_density ? 1 : 0};
}

inline void upload_dry(gwm_state* s, const std::vector<DryState>& st,
                       const std::vector<domain::SubdomainDescriptor>& L, const Global& G) {
  require(st.size() == L.size(), "State/layout size mismatch");
  std::vector<real> buf;
  for (int f = 0; f < 5; ++f) {
    std::vector<const state::Field3D<real>*> loc;
    for (const auto& x : st) loc.push_back(&x.field(f));
    const auto* o = gwm::detail::kOff[f];
    gwm::detail::gather(loc, L, G, o[0], o[1], o[2], buf);
    check(gwm_state_upload(s, f, buf.data()));
  }
}

inline void download_dry(gwm_state* s, std::vector<DryState>& st,
                         const std::vector<domain::SubdomainDescriptor>& L, const Global& G) {
  std::vector<real> buf;
  for (int f = 0; f < 5; ++f) {
    check(gwm_halo_fill(s, f));
    buf.resize(G.size(gwm::detail::kOff[f][0], gwm::detail::kOff[f][1], gwm::detail::kOff[f][2]));
    check(gwm_state_download(s, f, buf.data()));
    std::vector<state::Field3D<real>*> loc;
    for (auto& x : st) loc.push_back(&x.field(f));
    const auto* o = gwm::detail::kOff[f];
    gwm::detail::scatter(buf, L, G, o[0], o[1], o[2], loc);
  }
}

// Host plug-ins called from inside the device stepper.
struct PluginCall {
  const std::vector<DryState>* shape = nullptr;
  const std::vector<domain::SubdomainDescriptor>* layout = nullptr;
  const domain::GridMetrics* metrics = nullptr;
  const DryStepperConfig* config = nullptr;
  Global G;
  BoundaryUpdater* boundary = nullptr;
  FastModeIntegrator* fast = nullptr;
  std::exception_ptr error;

  std::vector<DryState> host_copy() const {
    std::vector<DryState> v;
    for (const auto& s : *shape) v.push_back(s.clone_empty_like("plugin_stage"));
    return v;
  }
};

inline int boundary_trampoline(gwm_state* stage, double t, void* user) {
  auto* c = static_cast<PluginCall*>(user);
  try {
    auto v = c->host_copy();
    download_dry(stage, v, *c->layout, c->G);
    c->boundary->apply(v, *c->layout, static_cast<real>(t));
    upload_dry(stage, v, *c->layout, c->G);
    return GWM_OK;
  } catch (...) {
    c->error = std::current_exception();
    return GWM_E_INVALID;
  }
}

inline int fast_trampoline(gwm_state* stage, double, void* user) {
  auto* c = static_cast<PluginCall*>(user);
  try {
    auto v = c->host_copy();
    download_dry(stage, v, *c->layout, c->G);
    c->fast->apply(v, *c->layout, *c->metrics, *c->config);
    upload_dry(stage, v, *c->layout, c->G);
    return GWM_OK;
  } catch (...) {
    c->error = std::current_exception();
    return GWM_E_INVALID;
  }
}
}  // namespace detail

inline void apply_local_split_explicit_fast_modes(
    std::vector<DryState>& states, const std::vector<domain::SubdomainDescriptor>& layout,
    const domain::GridMetrics& metrics, const DryStepperConfig& config) {
  const auto G = gwm::detail::global_of(layout, metrics);
  gwm::detail::StateLease s(metrics.device.get(), 0);
  detail::upload_dry(s, states, layout, G);
  const auto c = detail::cfg_of(config);
  detail::check(gwm_fast_modes(s, &c));
  detail::download_dry(s, states, layout, G);
}

inline void compute_slow_tendencies(const std::vector<DryState>& states,
                                    const std::vector<domain::SubdomainDescriptor>& layout,
                                    const domain::GridMetrics& metrics,
                                    const DryStepperConfig& config,
                                    std::vector<DrySlowTendencies>& out) {
  const auto G = gwm::detail::global_of(layout, metrics);
  require(out.size() == states.size(), "Tendency/state size mismatch");
  gwm::detail::StateLease s(metrics.device.get(), 0);
  detail::upload_dry(s, states, layout, G);
  const auto c = detail::cfg_of(config);
  detail::check(gwm_slow_tendencies(s, &c));
  std::vector<real> buf;
  for (int f = 0; f < 5; ++f) {
    const auto* o = gwm::detail::kOff[f];
    buf.resize(G.size(o[0], o[1], o[2]));
    detail::check(gwm_state_download(s, GWM_TEND_RHO + f, buf.data()));
    std::vector<state::Field3D<real>*> loc;
    for (auto& x : out) loc.push_back(&x.field(f));
    gwm::detail::scatter(buf, layout, G, o[0], o[1], o[2], loc, true);
  }
}

// advance_dry_state_ssprk3 (src/dycore/dry_core.cu:541-651).  The reference's
// own plug-ins run on the device; other BoundaryUpdater / FastModeIntegrator
// subclasses are called on host copies of each stage state.
inline void advance_dry_state_ssprk3(std::vector<DryState>& states,
                                     const std::vector<domain::SubdomainDescriptor>& layout,
                                     const domain::GridMetrics& metrics,
                                     const DryStepperConfig& config,
                                     BoundaryUpdater& boundary_updater,
                                     FastModeIntegrator& fast_modes) {
  const auto G = gwm::detail::global_of(layout, metrics);
  gwm::detail::StateLease s(metrics.device.get(), 0);
  detail::upload_dry(s, states, layout, G);
  const auto c = detail::cfg_of(config);
  detail::PluginCall call;
  call.shape = &states;
  call.layout = &layout;
  call.metrics = &metrics;
  call.config = &config;
  call.G = G;
  call.boundary = &boundary_updater;
  call.fast = &fast_modes;
  gwm_boundary_config b{GWM_BOUNDARY_NONE, 0.0, nullptr, nullptr};
  if (dynamic_cast<NullBoundaryUpdater*>(&boundary_updater) == nullptr) {
    b.callback = detail::boundary_trampoline;
    b.user = &call;
  }
  int rc;
  if (dynamic_cast<LocalSplitExplicitFastMode*>(&fast_modes) != nullptr)
    rc = gwm_ssprk3_step(s, &c, &b);
  else
    rc = gwm_ssprk3_step_custom_fast(s, &c, &b, detail::fast_trampoline, &call);
  if (call.error) std::rethrow_exception(call.error);
  detail::check(rc);
  detail::download_dry(s, states, layout, G);
}

inline std::vector<DryState> make_constant_dry_state(
    const std::vector<domain::SubdomainDescriptor>& layout, real rho, real theta, real u, real v,
    real w, const std::string& label_prefix) {
  std::vector<DryState> out;
  for (const auto& d : layout) {
    DryState s(d.nx_local(), d.ny_local(), d.nz, d.halo,
               label_prefix + "_rank_" + std::to_string(d.rank));
    s.fill_constant(rho, theta, u, v, w);
    out.push_back(std::move(s));
  }
  return out;
}

namespace detail {
template <class Init>
std::vector<DryState> device_init(const std::vector<domain::SubdomainDescriptor>& layout,
                                  const domain::GridMetrics& metrics, const std::string& prefix,
                                  Init init) {
  const auto G = gwm::detail::global_of(layout, metrics);
  gwm::detail::StateLease s(metrics.device.get(), 0);
  check(init(s));
  std::vector<DryState> out;
  for (const auto& d : layout)
    out.emplace_back(d.nx_local(), d.ny_local(), d.nz, d.halo,
                     prefix + "_rank_" + std::to_string(d.rank));
  download_dry(s, out, layout, G);
  return out;
}
}  // namespace detail

// make_hydrostatic_rest_state (src/dycore/dry_core.cu:363-398); the density
// scale height is unused there too.
inline std::vector<DryState> make_hydrostatic_rest_state(
    const std::vector<domain::SubdomainDescriptor>& layout, const domain::GridMetrics& metrics,
    real rho_surface, real theta_ref, real /*density_scale_height*/,
    const std::string& label_prefix) {
  return detail::device_init(layout, metrics, label_prefix, [&](gwm_state* s) {
    return gwm_init_hydrostatic_rest(s, rho_surface, theta_ref);
  });
}

struct ThermoBubbleConfig {
  real rho_background = 1.0f;
  real theta_background = 300.0f;
  real u_background = 0.0f;
  real v_background = 0.0f;
  real w_background = 0.0f;
  real center_x_fraction = 0.5f;
  real center_y_fraction = 0.5f;
  real center_z_fraction = 0.5f;
  real radius_x_fraction = 0.1f;
  real radius_y_fraction = 0.1f;
  real radius_z_fraction = 0.1f;
  real theta_perturbation = 2.0f;
};

struct MountainWaveBackgroundConfig {
  real rho_surface = 1.18f;
  real theta_ref = 300.0f;
  real density_scale_height = 8200.0f;
  real u_background = 15.0f;
  real v_background = 0.0f;
  real w_background = 0.0f;
};

namespace detail {
inline std::vector<DryState> bubble(const std::vector<domain::SubdomainDescriptor>& layout,
                                    const domain::GridMetrics& metrics,
                                    const ThermoBubbleConfig& c, const std::string& prefix,
                                    int sign) {
  const double p[12] = {c.rho_background,    c.theta_background,  c.u_background,
                        c.v_background,      c.w_background,      c.center_x_fraction,
                        c.center_y_fraction, c.center_z_fraction, c.radius_x_fraction,
                        c.radius_y_fraction, c.radius_z_fraction, c.theta_perturbation};
  return device_init(layout, metrics, prefix,
                     [&](gwm_state* s) { return gwm_init_bubble(s, p, sign); });
}
}  // namespace detail

// make_warm_bubble_state / make_density_current_state (src/dycore/idealized_cases.cpp:105-180)
inline std::vector<DryState> make_warm_bubble_state(
    const std::vector<domain::SubdomainDescriptor>& layout, const domain::GridMetrics& metrics,
    const ThermoBubbleConfig& config, const std::string& label_prefix) {
  return detail::bubble(layout, metrics, config, label_prefix, +1);
}
inline std::vector<DryState> make_density_current_state(
    const std::vector<domain::SubdomainDescriptor>& layout, const domain::GridMetrics& metrics,
    const ThermoBubbleConfig& config, const std::string& label_prefix) {
  return detail::bubble(layout, metrics, config, label_prefix, -1);
}

// make_mountain_wave_background_state (src/dycore/idealized_cases.cpp:182-223);
// the terrain comes from the metrics' device grid.
inline std::vector<DryState> make_mountain_wave_background_state(
    const std::vector<domain::SubdomainDescriptor>& layout, const domain::GridMetrics& metrics,
    const domain::IdealizedDomain& /*domain_topography*/, const MountainWaveBackgroundConfig& c,
    const std::string& label_prefix) {
  return detail::device_init(layout, metrics, label_prefix, [&](gwm_state* s) {
    return gwm_init_mountain_wave(s, c.rho_surface, c.theta_ref, c.u_background, c.v_background,
                                  c.w_background);
  });
}

// ---- diagnostics (include/gwm/dycore/dry_diagnostics.hpp) ----------------
struct ScalarRange {
  double min = 0.0;
  double max = 0.0;
};

struct DryStateSummary {
  double total_dry_mass = 0.0;
  double total_rho_theta_m = 0.0;
  double total_horizontal_momentum_x = 0.0;
  double total_horizontal_momentum_y = 0.0;
  double total_vertical_momentum = 0.0;
  ScalarRange rho_d, theta_m, u_face, v_face, w_face;
};

inline DryStateSummary summarize_dry_states(const std::vector<DryState>& states) {
  DryStateSummary s{};
  const double big = std::numeric_limits<double>::max();
  ScalarRange* r[5] = {&s.rho_d, &s.theta_m, &s.u_face, &s.v_face, &s.w_face};
  for (auto* x : r) *x = {states.empty() ? 0.0 : big, states.empty() ? 0.0 : -big};
  auto widen = [](ScalarRange& x, double v) {
    x.min = std::min(x.min, v);
    x.max = std::max(x.max, v);
  };
  for (const auto& st : states) {
    s.total_dry_mass += st.total_dry_mass();
    s.total_rho_theta_m += st.total_rho_theta_m();
    s.total_horizontal_momentum_x += st.mom_u.storage().owned_sum();
    s.total_horizontal_momentum_y += st.mom_v.storage().owned_sum();
    s.total_vertical_momentum += st.total_vertical_momentum();
    const auto& rho = st.rho_d;
    for (int k = 0; k < rho.nz(); ++k)
      for (int j = 0; j < rho.ny(); ++j)
        for (int i = 0; i < rho.nx(); ++i) {
          const double d = rho(i, j, k);
          widen(s.rho_d, d);
          widen(s.theta_m, static_cast<double>(st.rho_theta_m(i, j, k)) / std::max(d, 1.0e-12));
        }
    for (int f = 2; f < 5; ++f) {
      const auto& m = st.field(f);
      for (int k = 0; k < m.nz(); ++k)
        for (int j = 0; j < m.ny(); ++j)
          for (int i = 0; i < m.nx(); ++i) widen(*r[f], m(i, j, k));
    }
  }
  return s;
}

inline std::string summary_to_json(const DryStateSummary& s, const std::string& indent = "  ") {
  std::ostringstream o;
  auto range = [&](const char* key, const ScalarRange& x, bool last) {
    o << indent << "\"" << key << "\": {\"min\": " << x.min << ", \"max\": " << x.max << "}"
      << (last ? "\n" : ",\n");
  };
  o << "{\n";
  o << indent << "\"total_dry_mass\": " << s.total_dry_mass << ",\n";
  o << indent << "\"total_rho_theta_m\": " << s.total_rho_theta_m << ",\n";
  o << indent << "\"total_horizontal_momentum_x\": " << s.total_horizontal_momentum_x << ",\n";
  o << indent << "\"total_horizontal_momentum_y\": " << s.total_horizontal_momentum_y << ",\n";
  o << indent << "\"total_vertical_momentum\": " << s.total_vertical_momentum << ",\n";
  range("rho_d", s.rho_d, false);
  range("theta_m", s.theta_m, false);
  range("u_face", s.u_face, false);
  range("v_face", s.v_face, false);
  range("w_face", s.w_face, true);
  o << "}";
  return o.str();
}

// ---- boundaries (include/gwm/dycore/boundary_conditions.hpp) ---------------
inline bool touches_west_boundary(const domain::SubdomainDescriptor& d) {
  return !d.periodic_x && d.i_begin == 0;
}
inline bool touches_east_boundary(const domain::SubdomainDescriptor& d) {
  return !d.periodic_x && d.i_end == d.global_nx;
}
inline bool touches_south_boundary(const domain::SubdomainDescriptor& d) {
  return !d.periodic_y && d.j_begin == 0;
}
inline bool touches_north_boundary(const domain::SubdomainDescriptor& d) {
  return !d.periodic_y && d.j_end == d.global_ny;
}

// ---- passive tracers (include/gwm/dycore/passive_tracer.hpp) -------------
class TracerBoundaryUpdater {
 public:
  virtual ~TracerBoundaryUpdater() = default;
  virtual void apply(std::vector<state::TracerState>& states,
                     const std::vector<domain::SubdomainDescriptor>& layout, real sim_time) = 0;
};

class NullTracerBoundaryUpdater final : public TracerBoundaryUpdater {
 public:
  void apply(std::vector<state::TracerState>&, const std::vector<domain::SubdomainDescriptor>&,
             real) override {}
};

namespace detail {
inline std::vector<state::TracerState> tracers_from_global(
    const std::vector<DryState>& dry, const std::vector<domain::SubdomainDescriptor>& layout,
    int gnx, int gny, state::TracerRegistry reg, const std::vector<const std::vector<real>*>& q,
    const std::string& prefix) {
  require(dry.size() == layout.size(), "Dry-state/layout size mismatch");
  require(!layout.empty(), "Layout must not be empty");
  const size_t n3 = static_cast<size_t>(gnx) * gny * layout.front().nz;
  for (const auto* x : q)
    if (x != nullptr) require(x->size() == n3, "Tracer field size mismatch");
  std::vector<state::TracerState> out;
  for (const auto& d : layout) {
    state::TracerState t(reg, d.nx_local(), d.ny_local(), d.nz, d.halo,
                         prefix + "_rank_" + std::to_string(d.rank));
    t.fill_zero();
    const auto& rho = dry[static_cast<size_t>(d.rank)].rho_d;
    for (size_t n = 0; n < q.size(); ++n) {
      if (q[n] == nullptr) continue;
      auto& m = t.mass(static_cast<int>(n));
      for (int k = 0; k < d.nz; ++k)
        for (int j = 0; j < d.ny_local(); ++j)
          for (int i = 0; i < d.nx_local(); ++i)
            m(i, j, k) = rho(i, j, k) *
                         (*q[n])[(static_cast<size_t>(k) * gny + d.j_begin + j) * gnx + d.i_begin + i];
    }
    out.push_back(std::move(t));
  }
  return out;
}

inline void upload_tracers(gwm_state* s, const std::vector<state::TracerState>& tr,
                           const std::vector<domain::SubdomainDescriptor>& L, const Global& G) {
  std::vector<real> buf;
  for (size_t t = 0; t < tr.front().size(); ++t) {
    std::vector<const state::Field3D<real>*> loc;
    for (const auto& x : tr) loc.push_back(&x.mass(static_cast<int>(t)));
    gwm::detail::gather(loc, L, G, 0, 0, 0, buf);
    check(gwm_state_upload(s, GWM_TRACER0 + static_cast<int>(t), buf.data()));
  }
}

inline void download_tracers(gwm_state* s, std::vector<state::TracerState>& tr,
                             const std::vector<domain::SubdomainDescriptor>& L, const Global& G) {
  std::vector<real> buf(G.size(0, 0, 0));
  for (size_t t = 0; t < tr.front().size(); ++t) {
    check(gwm_halo_fill(s, GWM_TRACER0 + static_cast<int>(t)));
    check(gwm_state_download(s, GWM_TRACER0 + static_cast<int>(t), buf.data()));
    std::vector<state::Field3D<real>*> loc;
    for (auto& x : tr) loc.push_back(&x.mass(static_cast<int>(t)));
    gwm::detail::scatter(buf, L, G, 0, 0, 0, loc);
  }
}

// A host TracerBoundaryUpdater called from inside the device tracer stepper.
struct TracerPluginCall {
  const std::vector<state::TracerState>* shape = nullptr;
  const std::vector<domain::SubdomainDescriptor>* layout = nullptr;
  Global G;
  TracerBoundaryUpdater* boundary = nullptr;
  std::exception_ptr error;
};

// apply_tracer_boundaries_to_fields (src/dycore/dry_transport.cu:254-276):
// each tracer of the stage goes to the updater on its own, as one-tracer
// states whose registry holds only that tracer's spec.
inline int tracer_boundary_trampoline(gwm_state* stage, double t, void* user) {
  auto* c = static_cast<TracerPluginCall*>(user);
  try {
    const auto& shape = *c->shape;
    std::vector<real> buf(c->G.size(0, 0, 0));
    for (size_t q = 0; q < shape.front().size(); ++q) {
      const int f = GWM_TRACER0 + static_cast<int>(q);
      check(gwm_halo_fill(stage, f));
      check(gwm_state_download(stage, f, buf.data()));
      std::vector<state::TracerState> staged;
      staged.reserve(shape.size());
      for (const auto& x : shape) {
        state::TracerRegistry reg;
        reg.add(x.registry().at(static_cast<int>(q)));
        staged.emplace_back(std::move(reg), x.nx(), x.ny(), x.nz(), x.halo(), "stage_tracer");
      }
      std::vector<state::Field3D<real>*> loc;
      for (auto& x : staged) loc.push_back(&x.mass(0));
      gwm::detail::scatter(buf, *c->layout, c->G, 0, 0, 0, loc);
      c->boundary->apply(staged, *c->layout, static_cast<real>(t));
      std::vector<const state::Field3D<real>*> cl;
      for (const auto& x : staged) cl.push_back(&x.mass(0));
      gwm::detail::gather(cl, *c->layout, c->G, 0, 0, 0, buf);
      check(gwm_state_upload(stage, f, buf.data()));
    }
    return GWM_OK;
  } catch (...) {
    c->error = std::current_exception();
    return GWM_E_INVALID;
  }
}
}  // namespace detail

inline std::vector<state::TracerState> make_specific_humidity_tracers_from_global_field(
    const std::vector<DryState>& dry_states, const std::vector<domain::SubdomainDescriptor>& layout,
    int global_nx, int global_ny, const std::vector<real>& specific_humidity,
    const std::string& label_prefix = "specific_humidity") {
  return detail::tracers_from_global(dry_states, layout, global_nx, global_ny,
                                     state::make_specific_humidity_registry(),
                                     {&specific_humidity}, label_prefix);
}

inline std::vector<state::TracerState> make_warm_rain_tracers_from_global_fields(
    const std::vector<DryState>& dry_states, const std::vector<domain::SubdomainDescriptor>& layout,
    int global_nx, int global_ny, const std::vector<real>& specific_humidity,
    const std::vector<real>* cloud_water = nullptr, const std::vector<real>* rain_water = nullptr,
    const std::string& label_prefix = "warm_rain") {
  return detail::tracers_from_global(dry_states, layout, global_nx, global_ny,
                                     state::make_warm_rain_registry(),
                                     {&specific_humidity, cloud_water, rain_water}, label_prefix);
}

// advance_passive_tracers_ssprk3 (src/dycore/dry_transport.cu:440-518) with
// the post-dry density and mass fluxes frozen over the step.  The null
// tracer boundary stays on the device; any other TracerBoundaryUpdater runs
// on host copies of each stage input (gwm_tracers_ssprk3_callback).
inline void advance_passive_tracers_ssprk3(std::vector<state::TracerState>& tracers,
                                           const std::vector<DryState>& dry_states,
                                           const std::vector<domain::SubdomainDescriptor>& layout,
                                           const domain::GridMetrics& metrics,
                                           const DryStepperConfig& config,
                                           TracerBoundaryUpdater& boundary_updater) {
  require(tracers.size() == layout.size(), "Tracer/layout size mismatch");
  if (tracers.empty() || tracers.front().size() == 0) return;
  const auto G = gwm::detail::global_of(layout, metrics);
  const int ntr = static_cast<int>(tracers.front().size());
  gwm::detail::StateLease s(metrics.device.get(), ntr);
  detail::upload_dry(s, dry_states, layout, G);
  detail::upload_tracers(s, tracers, layout, G);
  const auto c = detail::cfg_of(config);
  if (dynamic_cast<NullTracerBoundaryUpdater*>(&boundary_updater) != nullptr) {
    detail::check(gwm_tracers_ssprk3(s, &c, GWM_BOUNDARY_NONE));
  } else {
    detail::TracerPluginCall call;
    call.shape = &tracers;
    call.layout = &layout;
    call.G = G;
    call.boundary = &boundary_updater;
    const int rc = gwm_tracers_ssprk3_callback(s, &c, GWM_BOUNDARY_NONE,
                                               detail::tracer_boundary_trampoline, &call);
    if (call.error) std::rethrow_exception(call.error);
    detail::check(rc);
  }
  detail::download_tracers(s, tracers, layout, G);
}

}  // namespace dycore

// --------------------------------------------------------------------------
// physics
// --------------------------------------------------------------------------
namespace physics {

struct WarmRainConfig {
  real dt = 2.0f;
  real condensation_relaxation = 1.0f;
  real cloud_autoconversion_threshold = 1.0e-4f;
  real cloud_autoconversion_rate = 1.0e-3f;
  real rain_evaporation_rate = 5.0e-3f;
  real rain_terminal_velocity = 0.0f;
  real sedimentation_layer_depth_m = 0.0f;
  bool enable_latent_heating = true;
};

struct WarmRainSurfaceAccumulation {
  int nx = 0;
  int ny = 0;
  double cell_area_m2 = 1.0;
  state::Field3D<real> liquid_precipitation_kg_m2;
  void reset(int nx_in, int ny_in, double cell_area_in = 1.0) {
    nx = nx_in;
    ny = ny_in;
    cell_area_m2 = cell_area_in;
    liquid_precipitation_kg_m2 = state::Field3D<real>(nx, ny, 1, 0, "warm_rain_surface_precip");
    liquid_precipitation_kg_m2.fill(0.0f);
  }
  bool matches(int nx_in, int ny_in) const {
    return nx == nx_in && ny == ny_in && liquid_precipitation_kg_m2.nx() == nx &&
           liquid_precipitation_kg_m2.ny() == ny && liquid_precipitation_kg_m2.nz() == 1 &&
           liquid_precipitation_kg_m2.halo() == 0;
  }
};

struct WarmRainSummary {
  double vapor_water_path_sum_kg_m2 = 0.0;
  double cloud_water_path_sum_kg_m2 = 0.0;
  double rain_water_path_sum_kg_m2 = 0.0;
  double accumulated_surface_precipitation_sum_mm = 0.0;
  double mean_surface_precipitation_mm = 0.0;
  double max_surface_precipitation_mm = 0.0;
};

namespace detail {
using gwm::detail::check;
inline gwm_warm_rain_config cfg_of(const WarmRainConfig& c) {
  return gwm_warm_rain_config{c.dt,
                              c.condensation_relaxation,
                              c.cloud_autoconversion_threshold,
                              c.cloud_autoconversion_rate,
                              c.rain_evaporation_rate,
                              c.rain_terminal_velocity,
                              c.sedimentation_layer_depth_m,
                              c.enable_latent_heating ? 1 : 0,
                              0.0};
}

inline void check_warm_rain_registry(const state::TracerState& t) {
  require(t.size() >= 3 && t.registry().at(GWM_QV).name == state::kSpecificHumidityTracerName &&
              t.registry().at(GWM_QC).name == state::kCloudWaterTracerName &&
              t.registry().at(GWM_QR).name == state::kRainWaterTracerName,
          "Warm-rain microphysics requires the warm-rain tracer registry");
}

// One device step over a gathered domain; bucket in/out is global (nx, ny).
inline void run(gwm_state* s, std::vector<dycore::DryState>& states,
                std::vector<state::TracerState>& tracers,
                const std::vector<domain::SubdomainDescriptor>& L, const gwm::detail::Global& G,
                const WarmRainConfig& config, int metric,
                std::vector<WarmRainSurfaceAccumulation>* acc) {
  dycore::detail::upload_dry(s, states, L, G);
  dycore::detail::upload_tracers(s, tracers, L, G);
  std::vector<real> bucket(static_cast<size_t>(G.nx) * G.ny, 0.0f);
  if (acc != nullptr)
    for (size_t n = 0; n < L.size(); ++n)
      for (int j = 0; j < L[n].ny_local(); ++j)
        for (int i = 0; i < L[n].nx_local(); ++i)
          bucket[static_cast<size_t>(L[n].j_begin + j) * G.nx + L[n].i_begin + i] =
              (*acc)[n].liquid_precipitation_kg_m2(i, j, 0);
  check(gwm_state_upload(s, GWM_PRECIP, bucket.data()));
  const auto c = cfg_of(config);
  check(gwm_warm_rain(s, &c, metric));
  std::vector<real> buf(G.size(0, 0, 0));
  check(gwm_halo_fill(s, GWM_RHO_THETA));
  check(gwm_state_download(s, GWM_RHO_THETA, buf.data()));
  std::vector<state::Field3D<real>*> loc;
  for (auto& x : states) loc.push_back(&x.rho_theta_m);
  gwm::detail::scatter(buf, L, G, 0, 0, 0, loc);
  dycore::detail::download_tracers(s, tracers, L, G);
  if (acc != nullptr) {
    check(gwm_state_download(s, GWM_PRECIP, bucket.data()));
    for (size_t n = 0; n < L.size(); ++n)
      for (int j = 0; j < L[n].ny_local(); ++j)
        for (int i = 0; i < L[n].nx_local(); ++i)
          (*acc)[n].liquid_precipitation_kg_m2(i, j, 0) =
              bucket[static_cast<size_t>(L[n].j_begin + j) * G.nx + L[n].i_begin + i];
  }
}
}  // namespace detail

// Nominal-fallout overload (src/physics/warm_rain.cu:213-271): every rank
// runs on its own, so each is stepped as a one-rank domain.
inline void apply_warm_rain_microphysics(std::vector<dycore::DryState>& states,
                                         std::vector<state::TracerState>& tracers,
                                         const WarmRainConfig& config,
                                         std::vector<WarmRainSurfaceAccumulation>* accumulations =
                                             nullptr) {
  require(states.size() == tracers.size(), "Warm-rain state/tracer size mismatch");
  if (accumulations != nullptr) {
    require(accumulations->size() == states.size(), "Warm-rain accumulation/state size mismatch");
    for (size_t n = 0; n < states.size(); ++n)
      if (!(*accumulations)[n].matches(states[n].rho_d.nx(), states[n].rho_d.ny()))
        (*accumulations)[n].reset(states[n].rho_d.nx(), states[n].rho_d.ny());
  }
  for (size_t n = 0; n < states.size(); ++n) {
    detail::check_warm_rain_registry(tracers[n]);
    const auto& r = states[n].rho_d;
    gwm_grid_desc d{r.nx(), r.ny(), r.nz(), r.halo(), 1, 1, 1.0, 1.0,
                    static_cast<double>(r.nz()), 0.25};
    gwm_grid* g = nullptr;
    detail::check(gwm_grid_create(gwm::detail::Device::get().ctx, &d, nullptr, nullptr, &g));
    std::shared_ptr<gwm_grid> grid(g, [](gwm_grid* p) {
      gwm::detail::StateCache::get().drop(p);
      gwm_grid_destroy(p);
    });
    const auto L = comm::VirtualRankLayout::build(r.nx(), r.ny(), r.nz(), r.halo(), 1, 1);
    const gwm::detail::Global G{r.nx(), r.ny(), r.nz(), r.halo()};
    gwm::detail::StateLease s(g, static_cast<int>(tracers[n].size()));
    std::vector<dycore::DryState> one;
    one.push_back(std::move(states[n]));
    std::vector<state::TracerState> tone;
    tone.push_back(std::move(tracers[n]));
    std::vector<WarmRainSurfaceAccumulation> aone;
    if (accumulations != nullptr) aone.push_back((*accumulations)[n]);
    std::exception_ptr err;
    try {
      detail::run(s, one, tone, L, G, config, 0, accumulations ? &aone : nullptr);
    } catch (...) {
      err = std::current_exception();
    }
    states[n] = std::move(one.front());
    tracers[n] = std::move(tone.front());
    if (accumulations != nullptr) (*accumulations)[n] = std::move(aone.front());
    if (err) std::rethrow_exception(err);
  }
}

// Layout/metric overload (src/physics/warm_rain.cu:273-353): fallout uses the
// terrain-following cell depths.
inline void apply_warm_rain_microphysics(std::vector<dycore::DryState>& states,
                                         std::vector<state::TracerState>& tracers,
                                         const std::vector<domain::SubdomainDescriptor>& layout,
                                         const domain::GridMetrics& metrics,
                                         const WarmRainConfig& config,
                                         std::vector<WarmRainSurfaceAccumulation>* accumulations) {
  require(states.size() == tracers.size() && states.size() == layout.size(),
          "Warm-rain state/tracer/layout size mismatch");
  if (states.empty()) return;
  const auto G = gwm::detail::global_of(layout, metrics);
  if (accumulations != nullptr) {
    require(accumulations->size() == states.size(), "Warm-rain accumulation/state size mismatch");
    for (size_t n = 0; n < states.size(); ++n)
      if (!(*accumulations)[n].matches(states[n].rho_d.nx(), states[n].rho_d.ny()))
        (*accumulations)[n].reset(states[n].rho_d.nx(), states[n].rho_d.ny(),
                                  metrics.dx * metrics.dy);
  }
  for (const auto& t : tracers) detail::check_warm_rain_registry(t);
  gwm::detail::StateLease s(metrics.device.get(),
                                                    static_cast<int>(tracers.front().size()));
  detail::run(s, states, tracers, layout, G, config, 1, accumulations);
}

// summarize_warm_rain (src/physics/warm_rain.cu:355-403)
inline WarmRainSummary summarize_warm_rain(
    const std::vector<state::TracerState>& tracers,
    const std::vector<domain::SubdomainDescriptor>& layout, const domain::GridMetrics& metrics,
    const std::vector<WarmRainSurfaceAccumulation>* accumulations = nullptr) {
  require(tracers.size() == layout.size(), "Tracer/layout size mismatch in summarize_warm_rain");
  WarmRainSummary s{};
  auto path = [&](const state::TracerState& t, const domain::SubdomainDescriptor& d, int q) {
    double total = 0.0;
    const auto& m = t.mass(q);
    for (int k = 0; k < d.nz; ++k)
      for (int j = 0; j < d.ny_local(); ++j)
        for (int i = 0; i < d.nx_local(); ++i)
          total += static_cast<double>(m(i, j, k)) /
                   static_cast<double>(metrics.inv_dz_cell(d.i_begin + i, d.j_begin + j, k));
    return total;
  };
  for (size_t n = 0; n < tracers.size(); ++n) {
    if (auto q = tracers[n].find(state::kSpecificHumidityTracerName))
      s.vapor_water_path_sum_kg_m2 += path(tracers[n], layout[n], *q);
    if (auto q = tracers[n].find(state::kCloudWaterTracerName))
      s.cloud_water_path_sum_kg_m2 += path(tracers[n], layout[n], *q);
    if (auto q = tracers[n].find(state::kRainWaterTracerName))
      s.rain_water_path_sum_kg_m2 += path(tracers[n], layout[n], *q);
  }
  if (accumulations != nullptr) {
    size_t cells = 0;
    for (const auto& a : *accumulations)
      for (int j = 0; j < a.ny; ++j)
        for (int i = 0; i < a.nx; ++i) {
          const double v = a.liquid_precipitation_kg_m2(i, j, 0);
          s.accumulated_surface_precipitation_sum_mm += v;
          s.max_surface_precipitation_mm = std::max(s.max_surface_precipitation_mm, v);
          ++cells;
        }
    if (cells > 0)
      s.mean_surface_precipitation_mm =
          s.accumulated_surface_precipitation_sum_mm / static_cast<double>(cells);
  }
  return s;
}

inline std::string warm_rain_summary_to_json(const WarmRainSummary& s,
                                             const std::string& indent = "") {
  std::ostringstream o;
  o << indent << "{\n";
  o << indent << "  \"vapor_water_path_sum_kg_m2\": " << s.vapor_water_path_sum_kg_m2 << ",\n";
  o << indent << "  \"cloud_water_path_sum_kg_m2\": " << s.cloud_water_path_sum_kg_m2 << ",\n";
  o << indent << "  \"rain_water_path_sum_kg_m2\": " << s.rain_water_path_sum_kg_m2 << ",\n";
  o << indent << "  \"accumulated_surface_precipitation_sum_mm\": "
    << s.accumulated_surface_precipitation_sum_mm << ",\n";
  o << indent << "  \"mean_surface_precipitation_mm\": " << s.mean_surface_precipitation_mm
    << ",\n";
  o << indent << "  \"max_surface_precipitation_mm\": " << s.max_surface_precipitation_mm << "\n";
  o << indent << "}";
  return o.str();
}

}  // namespace physics
}  // namespace gwm
