// gwm_oracle.cpp -- CPU oracle for the dry + warm-rain time step.
//
// TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline).  Nothing in
// the product links or loads this file.  It is a restatement, in plain
// host C++ templated on the working precision R (float or double), of the
// reference's algorithm; every routine cites the reference file:line it
// follows.  Where the reference lacks a feature that the benchmark
// configurations need (stretched levels, relaxation boundaries, accretion)
// the routine says "extension" and the reference-parity behaviour is the
// default.
//
// Deliberate deviation: the reference's flat-limit level means in the fast
// modes use FP32 atomicAdd (src/dycore/dry_fast_modes.cu:42-57), whose
// summation order is nondeterministic.  The oracle sums in FP64 in a fixed
// order and rounds once, which is what the HIP path does as well.

#include "gwm_oracle.h"

#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc {

thread_local std::string g_err;

inline void require(bool c, const char* msg) {
  if (!c) throw std::runtime_error(msg);
}

// ---------------------------------------------------------------------------
// Constants and thermodynamics
// (include/gwm/core/dry_thermo.hpp:17-44, include/gwm/core/moist_thermo.hpp:18-74)
// ---------------------------------------------------------------------------
template <class R>
struct K;
template <>
struct K<float> {
  static constexpr float Rd = 287.05f, p0 = 100000.0f, kappa = 0.2854f;
  static constexpr float Rv = 461.5f, Lv = 2.5e6f;
  static constexpr float eps = Rd / Rv;
  static constexpr float g_init = 9.81f;  // dry_core.cu:20, idealized_cases.cpp:12
  static constexpr float pi = 3.14159265358979323846f;
  static constexpr float mom_floor = -1.0e30f;
};
template <>
struct K<double> {
  static constexpr double Rd = 287.05, p0 = 100000.0, kappa = 0.2854;
  static constexpr double Rv = 461.5, Lv = 2.5e6;
  static constexpr double eps = Rd / Rv;
  static constexpr double g_init = 9.81;
  static constexpr double pi = 3.14159265358979323846;
  static constexpr double mom_floor = -1.0e30;
};

template <class R>
inline R theta_of(R rho, R rt) {  // dry_thermo.hpp:21-24
  return rt / std::max(rho, R(1.0e-6));
}
template <class R>
inline R pressure_of(R rho, R rt) {  // dry_thermo.hpp:26-33
  const R th = theta_of(rho, rt);
  const R ratio = std::max(rho * K<R>::Rd * th / K<R>::p0, R(1.0e-6));
  return K<R>::p0 * std::pow(ratio, R(1) / (R(1) - K<R>::kappa));
}
template <class R>
inline R rho_from_p_theta(R p, R th) {  // dry_thermo.hpp:35-40
  const R ex = std::pow(std::max(p / K<R>::p0, R(1.0e-6)), K<R>::kappa);
  return p / std::max(K<R>::Rd * th * ex, R(1.0e-6));
}
template <class R>
inline R cp() {  // dry_thermo.hpp:42-44
  return K<R>::Rd / K<R>::kappa;
}
template <class R>
inline R esat_pa(R T) {  // moist_thermo.hpp:23-28
  const R tc = T - R(273.15);
  const R es = R(6.112) * std::exp((R(17.67) * tc) / std::max(tc + R(243.5), R(1.0e-6)));
  return es * R(100);
}
template <class R>
inline R qsat(R T, R p) {  // moist_thermo.hpp:30-36
  const R es = esat_pa(T);
  const R denom = std::max(p - (R(1) - K<R>::eps) * es, R(1000));
  return std::min(std::max(K<R>::eps * es / denom, R(0)), R(0.05));
}
template <class R>
inline R temperature_of(R rho, R rt) {  // moist_thermo.hpp:56-63
  const R th = theta_of(rho, rt);
  const R p = pressure_of(rho, rt);
  const R ex = std::pow(std::max(p / K<R>::p0, R(1.0e-6)), K<R>::kappa);
  return th * ex;
}

// ---------------------------------------------------------------------------
// Layout (include/gwm/domain/subdomain_descriptor.hpp:7-26,
// src/comm/virtual_rank_layout.cpp:7-47, src/comm/cartesian_topology.cpp:57-74)
// ---------------------------------------------------------------------------
struct Desc {
  int rank = 0, cx = 0, cy = 0, rx = 1, ry = 1, gnx = 0, gny = 0, nz = 0, h = 0;
  int i0 = 0, i1 = 0, j0 = 0, j1 = 0;
  bool px = true, py = true;
  int nxl() const { return i1 - i0; }
  int nyl() const { return j1 - j0; }
};

std::vector<Desc> build_layout(int gnx, int gny, int nz, int h, int rx, int ry,
                               bool px, bool py) {
  std::vector<Desc> out;
  const int bx = gnx / rx, remx = gnx % rx, by = gny / ry, remy = gny % ry;
  int jb = 0;
  for (int y = 0; y < ry; ++y) {
    const int nyl = by + (y < remy ? 1 : 0);
    int ib = 0;
    for (int x = 0; x < rx; ++x) {
      const int nxl = bx + (x < remx ? 1 : 0);
      Desc d;
      d.rank = static_cast<int>(out.size());
      d.cx = x;
      d.cy = y;
      d.rx = rx;
      d.ry = ry;
      d.gnx = gnx;
      d.gny = gny;
      d.nz = nz;
      d.h = h;
      d.i0 = ib;
      d.i1 = ib + nxl;
      d.j0 = jb;
      d.j1 = jb + nyl;
      d.px = px;
      d.py = py;
      out.push_back(d);
      ib += nxl;
    }
    jb += nyl;
  }
  return out;
}

inline int wrap_mod(int v, int n) {
  int m = v % n;
  return m < 0 ? m + n : m;
}

struct Nbr {
  int w = -1, e = -1, s = -1, n = -1;
};

int rank_at(const std::vector<Desc>& L, int cx, int cy, bool px, bool py) {
  if (L.empty()) return -1;
  const int rx = L.front().rx, ry = L.front().ry;
  if (px) cx = wrap_mod(cx, rx);
  else if (cx < 0 || cx >= rx) return -1;
  if (py) cy = wrap_mod(cy, ry);
  else if (cy < 0 || cy >= ry) return -1;
  for (const auto& c : L)
    if (c.cx == cx && c.cy == cy) return c.rank;
  return -1;
}

Nbr neighbors(const std::vector<Desc>& L, const Desc& d) {
  Nbr n;
  n.w = rank_at(L, d.cx - 1, d.cy, d.px, d.py);
  n.e = rank_at(L, d.cx + 1, d.cy, d.px, d.py);
  n.s = rank_at(L, d.cx, d.cy - 1, d.px, d.py);
  n.n = rank_at(L, d.cx, d.cy + 1, d.px, d.py);
  return n;
}

// ---------------------------------------------------------------------------
// Halo-padded field (include/gwm/state/field3d.hpp:11-132): x fastest,
// idx = (k*(ny+2h) + j+h)*(nx+2h) + i+h, no halo in k.
// ---------------------------------------------------------------------------
enum Orient { CELL = 0, XF = 1, YF = 2, ZF = 3 };

template <class R>
struct Field {
  int nx = 0, ny = 0, nz = 0, h = 0;
  std::vector<R> d;
  Field() = default;
  Field(int nx_, int ny_, int nz_, int h_)
      : nx(nx_), ny(ny_), nz(nz_), h(h_),
        d(static_cast<size_t>(nx_ + 2 * h_) * (ny_ + 2 * h_) * nz_, R(0)) {}
  int px() const { return nx + 2 * h; }
  int py() const { return ny + 2 * h; }
  size_t idx(int i, int j, int k) const {
    return (static_cast<size_t>(k) * py() + (j + h)) * px() + (i + h);
  }
  R& operator()(int i, int j, int k) { return d[idx(i, j, k)]; }
  const R& operator()(int i, int j, int k) const { return d[idx(i, j, k)]; }
  size_t total() const { return d.size(); }
  void fill(R v) { std::fill(d.begin(), d.end(), v); }
  double owned_sum() const {  // field3d.hpp:82-92
    double s = 0.0;
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) s += static_cast<double>((*this)(i, j, k));
    return s;
  }
};

template <class R>
Field<R> make_field(const Desc& d, Orient o) {  // face_field.hpp:16-23
  return Field<R>(d.nxl() + (o == XF ? 1 : 0), d.nyl() + (o == YF ? 1 : 0),
                  d.nz + (o == ZF ? 1 : 0), d.h);
}

// ---------------------------------------------------------------------------
// Halo exchange, virtual-rank path (src/comm/halo_exchange.cpp:32-116,
// 148-285, 313-373, 563-620).  Absent neighbours leave zero-filled receive
// buffers in the ghosts (halo_exchange.cpp:289-301).
// ---------------------------------------------------------------------------
template <class R>
void exchange(std::vector<Field<R>*>& F, const std::vector<Desc>& L, Orient o) {
  const size_t nr = L.size();
  require(F.size() == nr, "field/layout size mismatch in exchange");
  // x phase: pack all, then unpack all.
  std::vector<std::vector<R>> ws(nr), es(nr);
  for (size_t n = 0; n < nr; ++n) {
    const Desc& d = L[n];
    const Field<R>& f = *F[n];
    for (int k = 0; k < f.nz; ++k)
      for (int j = 0; j < f.ny; ++j)
        for (int hh = 0; hh < d.h; ++hh) {
          int wi = hh, ei = f.nx - d.h + hh;
          if (o == XF) {
            wi = 1 + hh;
            ei = d.nxl() - d.h + hh;
          }
          ws[n].push_back(f(wi, j, k));
          es[n].push_back(f(ei, j, k));
        }
  }
  for (size_t n = 0; n < nr; ++n) {
    const Desc& d = L[n];
    const Nbr nb = neighbors(L, d);
    Field<R>& f = *F[n];
    const std::vector<R> zero(ws[n].size(), R(0));
    const std::vector<R>& wr = nb.w >= 0 ? es[nb.w] : zero;
    const std::vector<R>& er = nb.e >= 0 ? ws[nb.e] : zero;
    size_t a = 0;
    for (int k = 0; k < f.nz; ++k)
      for (int j = 0; j < f.ny; ++j)
        for (int hh = 0; hh < d.h; ++hh, ++a) {
          f(-d.h + hh, j, k) = wr[a];
          f(f.nx + hh, j, k) = er[a];
        }
  }
  // y phase (rows include the x ghosts).
  std::vector<std::vector<R>> ss(nr), ns(nr);
  for (size_t n = 0; n < nr; ++n) {
    const Desc& d = L[n];
    const Field<R>& f = *F[n];
    for (int k = 0; k < f.nz; ++k)
      for (int hh = 0; hh < d.h; ++hh)
        for (int i = -d.h; i < f.nx + d.h; ++i) {
          int sj = hh, nj = f.ny - d.h + hh;
          if (o == YF) {
            sj = 1 + hh;
            nj = d.nyl() - d.h + hh;
          }
          ss[n].push_back(f(i, sj, k));
          ns[n].push_back(f(i, nj, k));
        }
  }
  for (size_t n = 0; n < nr; ++n) {
    const Desc& d = L[n];
    const Nbr nb = neighbors(L, d);
    Field<R>& f = *F[n];
    const std::vector<R> zero(ss[n].size(), R(0));
    const std::vector<R>& sr = nb.s >= 0 ? ns[nb.s] : zero;
    const std::vector<R>& nr_ = nb.n >= 0 ? ss[nb.n] : zero;
    size_t a = 0;
    for (int k = 0; k < f.nz; ++k)
      for (int hh = 0; hh < d.h; ++hh)
        for (int i = -d.h; i < f.nx + d.h; ++i, ++a) {
          f(i, -d.h + hh, k) = sr[a];
          f(i, f.ny + hh, k) = nr_[a];
        }
  }
}

// Owned-interface averaging (halo_exchange.cpp:217-285), including the
// reference's wrap-interface skip for ranks >= 2 (condition n < east).
template <class R>
void sync_interfaces(std::vector<Field<R>*>& F, const std::vector<Desc>& L, Orient o) {
  if (o != XF && o != YF) return;
  for (size_t n = 0; n < L.size(); ++n) {
    const Desc& d = L[n];
    const Nbr nb = neighbors(L, d);
    Field<R>& f = *F[n];
    if (o == XF && nb.e >= 0 && (static_cast<int>(n) < nb.e || nb.e == static_cast<int>(n))) {
      Field<R>& p = *F[nb.e];
      for (int k = 0; k < f.nz; ++k)
        for (int j = 0; j < d.nyl(); ++j) {
          const R avg = R(0.5) * (f(d.nxl(), j, k) + p(0, j, k));
          f(d.nxl(), j, k) = avg;
          p(0, j, k) = avg;
        }
    }
    if (o == YF && nb.n >= 0 && (static_cast<int>(n) < nb.n || nb.n == static_cast<int>(n))) {
      Field<R>& p = *F[nb.n];
      for (int k = 0; k < f.nz; ++k)
        for (int i = 0; i < d.nxl(); ++i) {
          const R avg = R(0.5) * (f(i, d.nyl(), k) + p(i, 0, k));
          f(i, d.nyl(), k) = avg;
          p(i, 0, k) = avg;
        }
    }
  }
}

// ---------------------------------------------------------------------------
// Metrics (include/gwm/domain/grid_metrics.hpp:87-239).  Extension: an
// explicit monotone set of flat interface heights (stretched levels).
// ---------------------------------------------------------------------------
template <class R>
struct Metrics {
  int nx = 0, ny = 0, nz = 0;
  double dx = 0, dy = 0, z_top = 0, taper = 0.25;
  bool px = true, py = true;
  std::vector<double> eta_if, eta_c, w_if, w_c, z_if, z_c;
  std::vector<R> terrain, slope_x, slope_y;  // nx*ny
  std::vector<R> zif_m, zc_m, idz_c, idz_f;  // (k*ny+j)*nx+i
  bool terrain_active = false;
  size_t mi(int i, int j, int k) const {
    return (static_cast<size_t>(k) * ny + j) * nx + i;
  }
};

// K11 column hydrostatic reference of one column (dry_fast_modes.cu:108-158):
// p_ref(0) = p_bottom, rho_ref from (p_ref, theta) at every level, and above
// the bottom the trapezoid recurrence p_k = p_{k-1} - g (rho_{k-1} + rho_k)
// dz / 2 solved by four fixed-point passes from the explicit guess, floored
// at 1 Pa.  th[k], zc[k]: theta and centre height of level k.
template <class R>
void column_reference(int nz, const R* th, const R* zc, R g, R p_bottom, R* pref, R* rref) {
  pref[0] = p_bottom;
  rref[0] = rho_from_p_theta(pref[0], th[0]);
  for (int k = 1; k < nz; ++k) {
    const R dz = zc[k] - zc[k - 1];
    R pn = std::max(R(1), pref[k - 1] - g * rref[k - 1] * dz);
    for (int it = 0; it < 4; ++it) {
      const R rn = rho_from_p_theta(pn, th[k]);
      pn = pref[k - 1] - g * R(0.5) * (rref[k - 1] + rn) * dz;
      pn = std::max(pn, R(1));
    }
    pref[k] = pn;
    rref[k] = rho_from_p_theta(pn, th[k]);
  }
}

double taper_weight(double eta, double taper) {  // grid_metrics.hpp:126-144
  if (eta <= 0.0) return 1.0;
  if (eta >= 1.0) return 0.0;
  if (taper <= 0.0) {
    const double s = std::clamp(eta, 0.0, 1.0);
    return 1.0 - s * s * (3.0 - 2.0 * s);
  }
  if (eta <= taper) return 1.0;
  const double s = std::clamp((eta - taper) / std::max(1.0e-9, 1.0 - taper), 0.0, 1.0);
  return 1.0 - s * s * (3.0 - 2.0 * s);
}

template <class R>
Metrics<R> build_metrics(int nx, int ny, int nz, double dx, double dy, double z_top,
                         const double* zin, const float* terr, double taper, bool px,
                         bool py) {
  Metrics<R> m;
  m.nx = nx;
  m.ny = ny;
  m.nz = nz;
  m.dx = dx;
  m.dy = dy;
  m.taper = taper;
  m.px = px;
  m.py = py;
  m.eta_if.resize(nz + 1);
  m.eta_c.resize(nz);
  m.w_if.resize(nz + 1);
  m.w_c.resize(nz);
  m.z_if.resize(nz + 1);
  m.z_c.resize(nz);
  if (zin == nullptr) {  // grid_metrics.hpp:146-161 (uniform eta)
    m.z_top = z_top;
    for (int k = 0; k <= nz; ++k) {
      m.eta_if[k] = static_cast<double>(k) / static_cast<double>(nz);
      m.w_if[k] = taper_weight(m.eta_if[k], taper);
      m.z_if[k] = m.eta_if[k] * z_top;
      if (k < nz) {
        m.eta_c[k] = (static_cast<double>(k) + 0.5) / static_cast<double>(nz);
        m.w_c[k] = taper_weight(m.eta_c[k], taper);
        m.z_c[k] = m.eta_c[k] * z_top;
      }
    }
  } else {  // extension: stretched flat interfaces, eta = z / z_top
    m.z_top = zin[nz];
    for (int k = 0; k <= nz; ++k) {
      m.z_if[k] = zin[k];
      m.eta_if[k] = zin[k] / m.z_top;
      m.w_if[k] = taper_weight(m.eta_if[k], taper);
    }
    for (int k = 0; k < nz; ++k) {
      m.z_c[k] = 0.5 * (zin[k] + zin[k + 1]);
      m.eta_c[k] = m.z_c[k] / m.z_top;
      m.w_c[k] = taper_weight(m.eta_c[k], taper);
    }
  }
  const size_t n2 = static_cast<size_t>(nx) * ny;
  m.terrain.assign(n2, R(0));
  m.slope_x.assign(n2, R(0));
  m.slope_y.assign(n2, R(0));
  m.zif_m.assign(n2 * (nz + 1), R(0));
  m.zc_m.assign(n2 * nz, R(0));
  m.idz_c.assign(n2 * nz, R(0));
  m.idz_f.assign(n2 * (nz + 1), R(0));
  // terrain_dyn is vector<real> (float) in the reference.
  auto tr = [&](int i, int j) -> float {
    return terr ? terr[static_cast<size_t>(j) * nx + i] : 0.0f;
  };
  auto wi = [&](int i) { return px ? wrap_mod(i, nx) : std::clamp(i, 0, nx - 1); };
  auto wj = [&](int j) { return py ? wrap_mod(j, ny) : std::clamp(j, 0, ny - 1); };
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const float th = tr(i, j);
      m.terrain[static_cast<size_t>(j) * nx + i] = static_cast<R>(th);
      if (std::fabs(static_cast<double>(th)) > 1.0e-6) m.terrain_active = true;
      const double ddx = (static_cast<double>(tr(wi(i + 1), j)) -
                          static_cast<double>(tr(wi(i - 1), j))) /
                         (px || (i > 0 && i + 1 < nx) ? 2.0 * dx : dx);
      const double ddy = (static_cast<double>(tr(i, wj(j + 1))) -
                          static_cast<double>(tr(i, wj(j - 1)))) /
                         (py || (j > 0 && j + 1 < ny) ? 2.0 * dy : dy);
      m.slope_x[static_cast<size_t>(j) * nx + i] = static_cast<R>(ddx);
      m.slope_y[static_cast<size_t>(j) * nx + i] = static_cast<R>(ddy);
      const double thd = static_cast<double>(th);
      for (int k = 0; k <= nz; ++k) {
        m.zif_m[m.mi(i, j, k)] = static_cast<R>(m.z_if[k] + m.w_if[k] * thd);
        if (k == 0 || k == nz) {
          m.idz_f[m.mi(i, j, k)] = R(0);
        } else {
          const double dzf = (m.z_c[k] + m.w_c[k] * thd) - (m.z_c[k - 1] + m.w_c[k - 1] * thd);
          m.idz_f[m.mi(i, j, k)] = static_cast<R>(1.0 / std::max(dzf, 1.0e-6));
        }
      }
      for (int k = 0; k < nz; ++k) {
        m.zc_m[m.mi(i, j, k)] = static_cast<R>(m.z_c[k] + m.w_c[k] * thd);
        const double dzc = static_cast<double>(m.zif_m[m.mi(i, j, k + 1)]) -
                           static_cast<double>(m.zif_m[m.mi(i, j, k)]);
        m.idz_c[m.mi(i, j, k)] = static_cast<R>(1.0 / std::max(dzc, 1.0e-6));
      }
    }
  return m;
}

// ---------------------------------------------------------------------------
// World: layout + per-rank fields (DryState, dry_core.hpp:14-34; tendencies
// dry_core.hpp:36-48; TracerState tracer_state.hpp:12-119; precipitation
// bucket warm_rain.hpp:25-46).
// ---------------------------------------------------------------------------
struct WorldBase {
  virtual ~WorldBase() = default;
  int prec = 4;
  int n_tracers = 0;
  unsigned tracer_strip_mask = ~0u;  // tracers the strip boundary overwrites
  std::vector<Desc> L;
  virtual long storage_size(int rank, int field, int* dims) = 0;
  virtual void get_rank(int rank, int field, void* out) = 0;
  virtual void set_rank(int rank, int field, const void* in) = 0;
  virtual void get_global(int field, void* out) = 0;
  virtual void set_global(int field, const void* in) = 0;
  virtual void fill(int field, double v) = 0;
  virtual void get_metric(int which, void* out) = 0;
  virtual void exch(int field) = 0;
  virtual void sync(int field) = 0;
  virtual void pressure() = 0;
  virtual void slow(const orc_step_cfg& c) = 0;
  virtual void mflux_only() = 0;
  virtual void pgf_only() = 0;
  virtual void fast(const orc_step_cfg& c) = 0;
  virtual void set_boundary() = 0;
  virtual void boundary_now(int mode, double w) = 0;
  virtual void ssprk3(const orc_step_cfg& c, int bmode, double w, double tau) = 0;
  virtual void tracers(const orc_step_cfg& c, int bmode) = 0;
  virtual void warm_rain(const orc_warm_rain_cfg& c, int metric) = 0;
  virtual void summary(double* out) = 0;
  virtual int plan_view(int slice_k, int steps, double dt, int with_precip, double* out) = 0;
  virtual void init_constant(double rho, double th, double u, double v, double w) = 0;
  virtual void init_hydro(double rho_s, double th) = 0;
  virtual void init_bubble(const double* p, int sign) = 0;
  virtual void init_mountain(double rho_s, double th, double u, double v, double w) = 0;
};

template <class R>
struct State {
  Field<R> rho, rt, u, v, w;
};

template <class R>
struct World : WorldBase {
  Metrics<R> M;
  std::vector<State<R>> S;
  std::vector<Field<R>> P;                 // pressure output (compute_dry_pressure_fields)
  std::vector<State<R>> T;                 // slow tendencies
  std::vector<std::vector<Field<R>>> Q;    // tracers (rho*q)
  std::vector<Field<R>> precip;            // (nx_local, ny_local, 1, h=0)
  std::vector<State<R>> B;                 // boundary reference state
  std::vector<std::vector<Field<R>>> QB;   // boundary tracer state
  bool have_boundary = false;

  State<R> make_state(const Desc& d) {
    State<R> s;
    s.rho = make_field<R>(d, CELL);
    s.rt = make_field<R>(d, CELL);
    s.u = make_field<R>(d, XF);
    s.v = make_field<R>(d, YF);
    s.w = make_field<R>(d, ZF);
    return s;
  }

  World(int prec_, int nx, int ny, int nz, int h, int rx, int ry, bool px, bool py,
        double dx, double dy, double z_top, const double* zin, const float* terr,
        double taper, int ntr) {
    prec = prec_;
    n_tracers = ntr;
    require(nx > 0 && ny > 0 && nz > 0 && h >= 1, "bad dimensions");
    require(rx > 0 && ry > 0, "bad rank counts");
    L = build_layout(nx, ny, nz, h, rx, ry, px, py);
    M = build_metrics<R>(nx, ny, nz, dx, dy, z_top, zin, terr, taper, px, py);
    for (const auto& d : L) {
      S.push_back(make_state(d));
      T.push_back(make_state(d));
      P.push_back(make_field<R>(d, CELL));
      Q.emplace_back();
      for (int t = 0; t < ntr; ++t) Q.back().push_back(make_field<R>(d, CELL));
      precip.emplace_back(d.nxl(), d.nyl(), 1, 0);
    }
  }

  Orient orient_of(int field) const {
    switch (field) {
      case ORC_MOM_U:
      case ORC_TEND_U:
        return XF;
      case ORC_MOM_V:
      case ORC_TEND_V:
        return YF;
      case ORC_MOM_W:
      case ORC_TEND_W:
        return ZF;
      default:
        return CELL;
    }
  }

  Field<R>& fld(int rank, int field) {
    require(rank >= 0 && rank < static_cast<int>(L.size()), "bad rank");
    State<R>& s = S[rank];
    State<R>& t = T[rank];
    switch (field) {
      case ORC_RHO: return s.rho;
      case ORC_RHO_THETA: return s.rt;
      case ORC_MOM_U: return s.u;
      case ORC_MOM_V: return s.v;
      case ORC_MOM_W: return s.w;
      case ORC_PRESSURE: return P[rank];
      case ORC_TEND_RHO: return t.rho;
      case ORC_TEND_RHO_THETA: return t.rt;
      case ORC_TEND_U: return t.u;
      case ORC_TEND_V: return t.v;
      case ORC_TEND_W: return t.w;
      case ORC_PRECIP: return precip[rank];
      default:
        break;
    }
    const int tr = field - ORC_TRACER0;
    require(tr >= 0 && tr < n_tracers, "bad field id");
    return Q[rank][tr];
  }

  long storage_size(int rank, int field, int* dims) override {
    Field<R>& f = fld(rank, field);
    if (dims) {
      dims[0] = f.nx;
      dims[1] = f.ny;
      dims[2] = f.nz;
      dims[3] = f.h;
    }
    return static_cast<long>(f.total());
  }
  void get_rank(int rank, int field, void* out) override {
    Field<R>& f = fld(rank, field);
    std::memcpy(out, f.d.data(), f.total() * sizeof(R));
  }
  void set_rank(int rank, int field, const void* in) override {
    Field<R>& f = fld(rank, field);
    std::memcpy(f.d.data(), in, f.total() * sizeof(R));
  }
  void global_dims(int field, int& gx, int& gy, int& gz) {
    const Orient o = orient_of(field);
    gx = M.nx + (o == XF);
    gy = M.ny + (o == YF);
    gz = M.nz + (o == ZF);
    if (field == ORC_PRECIP) gz = 1;
  }
  // VirtualRankLayout::gather_face / gather_scalar (virtual_rank_layout.cpp:73-135)
  void get_global(int field, void* out) override {
    int gx, gy, gz;
    global_dims(field, gx, gy, gz);
    R* o = static_cast<R*>(out);
    for (size_t n = 0; n < L.size(); ++n) {
      Field<R>& f = fld(static_cast<int>(n), field);
      for (int k = 0; k < f.nz; ++k)
        for (int j = 0; j < f.ny; ++j)
          for (int i = 0; i < f.nx; ++i)
            o[(static_cast<size_t>(k) * gy + L[n].j0 + j) * gx + L[n].i0 + i] = f(i, j, k);
    }
  }
  void set_global(int field, const void* in) override {
    int gx, gy, gz;
    global_dims(field, gx, gy, gz);
    const R* o = static_cast<const R*>(in);
    for (size_t n = 0; n < L.size(); ++n) {
      Field<R>& f = fld(static_cast<int>(n), field);
      for (int k = 0; k < f.nz; ++k)
        for (int j = 0; j < f.ny; ++j)
          for (int i = 0; i < f.nx; ++i)
            f(i, j, k) = o[(static_cast<size_t>(k) * gy + L[n].j0 + j) * gx + L[n].i0 + i];
    }
  }
  void fill(int field, double v) override {
    for (size_t n = 0; n < L.size(); ++n) fld(static_cast<int>(n), field).fill(static_cast<R>(v));
  }
  void get_metric(int which, void* out) override {
    R* o = static_cast<R*>(out);
    double* od = static_cast<double*>(out);
    switch (which) {
      case ORC_METRIC_Z_INTERFACES: std::copy(M.zif_m.begin(), M.zif_m.end(), o); break;
      case ORC_METRIC_Z_CENTERS: std::copy(M.zc_m.begin(), M.zc_m.end(), o); break;
      case ORC_METRIC_INV_DZ_CELL: std::copy(M.idz_c.begin(), M.idz_c.end(), o); break;
      case ORC_METRIC_INV_DZ_FACE: std::copy(M.idz_f.begin(), M.idz_f.end(), o); break;
      case ORC_METRIC_TERRAIN: std::copy(M.terrain.begin(), M.terrain.end(), o); break;
      case ORC_METRIC_SLOPE_X: std::copy(M.slope_x.begin(), M.slope_x.end(), o); break;
      case ORC_METRIC_SLOPE_Y: std::copy(M.slope_y.begin(), M.slope_y.end(), o); break;
      case ORC_METRIC_FLAT_Z_INTERFACES: std::copy(M.z_if.begin(), M.z_if.end(), od); break;
      case ORC_METRIC_FLAT_Z_CENTERS: std::copy(M.z_c.begin(), M.z_c.end(), od); break;
      case ORC_METRIC_WEIGHT_INTERFACES: std::copy(M.w_if.begin(), M.w_if.end(), od); break;
      case ORC_METRIC_WEIGHT_CENTERS: std::copy(M.w_c.begin(), M.w_c.end(), od); break;
      default: throw std::runtime_error("bad metric id");
    }
  }

  template <class Get>
  std::vector<Field<R>*> ptrs(Get get) {
    std::vector<Field<R>*> v;
    for (size_t n = 0; n < L.size(); ++n) v.push_back(&get(n));
    return v;
  }
  void exch(int field) override {
    auto v = ptrs([&](size_t n) -> Field<R>& { return fld(static_cast<int>(n), field); });
    exchange(v, L, orient_of(field));
  }
  void sync(int field) override {
    auto v = ptrs([&](size_t n) -> Field<R>& { return fld(static_cast<int>(n), field); });
    sync_interfaces(v, L, orient_of(field));
  }

  // compute_dry_pressure_fields (dry_pressure_gradient.cu:177-202), K20 over
  // the full storage including the halo.
  static void eos_total(const Field<R>& rho, const Field<R>& rt, Field<R>& p) {
    const size_t n = rho.total();
#pragma omp parallel for schedule(static)
    for (long a = 0; a < static_cast<long>(n); ++a) p.d[a] = pressure_of(rho.d[a], rt.d[a]);
  }
  void pressure() override {
    for (size_t n = 0; n < L.size(); ++n) eos_total(S[n].rho, S[n].rt, P[n]);
  }

  // Halo clones of a state with interface synchronisation
  // (dry_core.cu:417-457, dry_transport.cu:189-230).
  std::vector<State<R>> halo_clones(const std::vector<State<R>>& src) {
    std::vector<State<R>> c = src;
    auto fr = ptrs([&](size_t n) -> Field<R>& { return c[n].rho; });
    auto ft = ptrs([&](size_t n) -> Field<R>& { return c[n].rt; });
    auto fu = ptrs([&](size_t n) -> Field<R>& { return c[n].u; });
    auto fv = ptrs([&](size_t n) -> Field<R>& { return c[n].v; });
    auto fw = ptrs([&](size_t n) -> Field<R>& { return c[n].w; });
    exchange(fr, L, CELL);
    exchange(ft, L, CELL);
    exchange(fu, L, XF);
    exchange(fv, L, YF);
    exchange(fw, L, ZF);
    sync_interfaces(fu, L, XF);
    sync_interfaces(fv, L, YF);
    exchange(fu, L, XF);
    exchange(fv, L, YF);
    return c;
  }

  R mwrap_i(int gi) const { return 0; }
  int wi(int gi) const { return M.px ? wrap_mod(gi, M.nx) : gi; }
  int wj(int gj) const { return M.py ? wrap_mod(gj, M.ny) : gj; }

  // dry_pressure_gradient.cu:37-73 (vertical_gradient_at_cell)
  R vgrad(const Field<R>& f, int gi, int gj, int k, int li, int lj) const {
    const int nz = f.nz;
    if (nz == 1) return R(0);
    const std::vector<R>& z = M.zc_m;
    if (k == 0) {
      const R dz = z[M.mi(gi, gj, 1)] - z[M.mi(gi, gj, 0)];
      return (f(li, lj, 1) - f(li, lj, 0)) / std::max(dz, R(1.0e-6));
    }
    if (k + 1 == nz) {
      const R dz = z[M.mi(gi, gj, k)] - z[M.mi(gi, gj, k - 1)];
      return (f(li, lj, k) - f(li, lj, k - 1)) / std::max(dz, R(1.0e-6));
    }
    const R dz = z[M.mi(gi, gj, k + 1)] - z[M.mi(gi, gj, k - 1)];
    return (f(li, lj, k + 1) - f(li, lj, k - 1)) / std::max(dz, R(1.0e-6));
  }

  // K12/K21: x-face pressure gradient term (dry_pressure_gradient.cu:75-123,
  // dry_fast_modes.cu:160-229); returns (p_r-p_l)/dx - dzdx*dpdz_face.
  R pgf_x(const Field<R>& p, const Desc& d, int i, int j, int k) const {
    const R dx = static_cast<R>(M.dx);
    const R pl = p(i - 1, j, k), pr = p(i, j, k);
    const int gj = wj(d.j0 + j), gl = wi(d.i0 + i - 1), gr = wi(d.i0 + i);
    const R zl = M.zc_m[M.mi(gl, gj, k)], zr = M.zc_m[M.mi(gr, gj, k)];
    const R dzdx = (zr - zl) / dx;
    const R dl = vgrad(p, gl, gj, k, i - 1, j);
    const R drr = vgrad(p, gr, gj, k, i, j);
    const R dpdz = R(0.5) * (dl + drr);
    return (pr - pl) / dx - dzdx * dpdz;
  }
  R pgf_y(const Field<R>& p, const Desc& d, int i, int j, int k) const {
    const R dy = static_cast<R>(M.dy);
    const R ps = p(i, j - 1, k), pn = p(i, j, k);
    const int gi = wi(d.i0 + i), gs = wj(d.j0 + j - 1), gn = wj(d.j0 + j);
    const R zs = M.zc_m[M.mi(gi, gs, k)], zn = M.zc_m[M.mi(gi, gn, k)];
    const R dzdy = (zn - zs) / dy;
    const R ds = vgrad(p, gi, gs, k, i, j - 1);
    const R dn = vgrad(p, gi, gn, k, i, j);
    const R dpdz = R(0.5) * (ds + dn);
    return (pn - ps) / dy - dzdy * dpdz;
  }

  static R upw(R a, R ql, R qr) { return a >= R(0) ? a * ql : a * qr; }  // dry_momentum_flux.cu:75-77

  // add_dry_momentum_flux_tendencies (dry_momentum_flux.cu:36-409) on
  // halo-exchanged clones; adds into tendencies.
  void momentum_flux(const std::vector<State<R>>& C, std::vector<State<R>>& out) {
    const R dx = static_cast<R>(M.dx), dy = static_cast<R>(M.dy);
    for (size_t n = 0; n < L.size(); ++n) {
      const Desc& d = L[n];
      const Nbr nb = neighbors(L, d);
      const bool ow = nb.w < 0, oe = nb.e < 0, os = nb.s < 0, on = nb.n < 0;
      const State<R>& c = C[n];
      const int nx = d.nxl(), ny = d.nyl(), nz = d.nz, h = d.h;
      Field<R> uc(nx, ny, nz, h), vc(nx, ny, nz, h), wc(nx, ny, nz, h);
      // K16 over cells including the halo (dry_momentum_flux.cu:36-73)
#pragma omp parallel for schedule(static)
      for (int k = 0; k < nz; ++k)
        for (int j = -h; j < ny + h; ++j)
          for (int i = -h; i < nx + h; ++i) {
            const R r = std::max(c.rho(i, j, k), R(1.0e-6));
            uc(i, j, k) = R(0.5) * (c.u(i, j, k) + c.u(i + 1, j, k)) / r;
            vc(i, j, k) = R(0.5) * (c.v(i, j, k) + c.v(i, j + 1, k)) / r;
            wc(i, j, k) = R(0.5) * (c.w(i, j, k) + c.w(i, j, k + 1)) / r;
          }
      State<R>& t = out[n];
      // K17 (dry_momentum_flux.cu:79-156)
#pragma omp parallel for schedule(static)
      for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
          for (int i = 0; i <= nx; ++i) {
            if ((i == 0 && ow) || (i == nx && oe)) continue;
            const Field<R>& q = c.u;
            const R qw = q(i - 1, j, k), qc = q(i, j, k), qe = q(i + 1, j, k);
            const R qs = q(i, j - 1, k), qn = q(i, j + 1, k);
            const R qb = k > 0 ? q(i, j, k - 1) : qc;
            const R qt = k + 1 < nz ? q(i, j, k + 1) : qc;
            const R fw = upw(uc(i - 1, j, k), qw, qc);
            const R fe = upw(uc(i, j, k), qc, qe);
            R fs = 0, fn = 0, fb = 0, ft = 0;
            if (!(j == 0 && os)) fs = upw(R(0.5) * (vc(i - 1, j - 1, k) + vc(i, j - 1, k)), qs, qc);
            if (!(j + 1 == ny && on)) fn = upw(R(0.5) * (vc(i - 1, j, k) + vc(i, j, k)), qc, qn);
            if (k > 0) fb = upw(R(0.5) * (wc(i - 1, j, k - 1) + wc(i, j, k - 1)), qb, qc);
            if (k + 1 < nz) ft = upw(R(0.5) * (wc(i - 1, j, k) + wc(i, j, k)), qc, qt);
            const int gj = wj(d.j0 + j), gl = wi(d.i0 + i - 1), gr = wi(d.i0 + i);
            const R idz = R(0.5) * (M.idz_c[M.mi(gl, gj, k)] + M.idz_c[M.mi(gr, gj, k)]);
            t.u(i, j, k) += -((fe - fw) / dx + (fn - fs) / dy + (ft - fb) * idz);
          }
      // K18 (dry_momentum_flux.cu:158-235)
#pragma omp parallel for schedule(static)
      for (int k = 0; k < nz; ++k)
        for (int j = 0; j <= ny; ++j)
          for (int i = 0; i < nx; ++i) {
            if ((j == 0 && os) || (j == ny && on)) continue;
            const Field<R>& q = c.v;
            const R qw = q(i - 1, j, k), qc = q(i, j, k), qe = q(i + 1, j, k);
            const R qs = q(i, j - 1, k), qn = q(i, j + 1, k);
            const R qb = k > 0 ? q(i, j, k - 1) : qc;
            const R qt = k + 1 < nz ? q(i, j, k + 1) : qc;
            R fw = 0, fe = 0, fb = 0, ft = 0;
            if (!(i == 0 && ow)) fw = upw(R(0.5) * (uc(i - 1, j - 1, k) + uc(i - 1, j, k)), qw, qc);
            if (!(i + 1 == nx && oe)) fe = upw(R(0.5) * (uc(i, j - 1, k) + uc(i, j, k)), qc, qe);
            const R fs = upw(vc(i, j - 1, k), qs, qc);
            const R fn = upw(vc(i, j, k), qc, qn);
            if (k > 0) fb = upw(R(0.5) * (wc(i, j - 1, k - 1) + wc(i, j, k - 1)), qb, qc);
            if (k + 1 < nz) ft = upw(R(0.5) * (wc(i, j - 1, k) + wc(i, j, k)), qc, qt);
            const int gi = wi(d.i0 + i), gs = wj(d.j0 + j - 1), gn = wj(d.j0 + j);
            const R idz = R(0.5) * (M.idz_c[M.mi(gi, gs, k)] + M.idz_c[M.mi(gi, gn, k)]);
            t.v(i, j, k) += -((fe - fw) / dx + (fn - fs) / dy + (ft - fb) * idz);
          }
      // K19 (dry_momentum_flux.cu:237-312)
#pragma omp parallel for schedule(static)
      for (int k = 1; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
          for (int i = 0; i < nx; ++i) {
            const Field<R>& q = c.w;
            const int kl = k - 1, ku = k;
            const R qw = q(i - 1, j, k), qc = q(i, j, k), qe = q(i + 1, j, k);
            const R qs = q(i, j - 1, k), qn = q(i, j + 1, k);
            const R qb = q(i, j, k - 1), qt = q(i, j, k + 1);
            R fw = 0, fe = 0, fs = 0, fn = 0;
            if (!(i == 0 && ow)) fw = upw(R(0.5) * (uc(i - 1, j, kl) + uc(i - 1, j, ku)), qw, qc);
            if (!(i + 1 == nx && oe)) fe = upw(R(0.5) * (uc(i, j, kl) + uc(i, j, ku)), qc, qe);
            if (!(j == 0 && os)) fs = upw(R(0.5) * (vc(i, j - 1, kl) + vc(i, j - 1, ku)), qs, qc);
            if (!(j + 1 == ny && on)) fn = upw(R(0.5) * (vc(i, j, kl) + vc(i, j, ku)), qc, qn);
            const R fb = upw(wc(i, j, kl), qb, qc);
            const R ft = upw(wc(i, j, ku), qc, qt);
            const R idz = M.idz_f[M.mi(wi(d.i0 + i), wj(d.j0 + j), k)];
            t.w(i, j, k) += -((fe - fw) / dx + (fn - fs) / dy + (ft - fb) * idz);
          }
    }
  }

  // add_horizontal_pressure_gradient_tendencies (dry_pressure_gradient.cu:204-240)
  void pressure_gradient(const std::vector<Field<R>>& PF, std::vector<State<R>>& out) {
    for (size_t n = 0; n < L.size(); ++n) {
      const Desc& d = L[n];
      const Nbr nb = neighbors(L, d);
      const int nx = d.nxl(), ny = d.nyl(), nz = d.nz;
#pragma omp parallel for schedule(static)
      for (int k = 0; k < nz; ++k) {
        for (int j = 0; j < ny; ++j)
          for (int i = 0; i <= nx; ++i) {
            if ((i == 0 && nb.w < 0) || (i == nx && nb.e < 0)) continue;
            out[n].u(i, j, k) += -pgf_x(PF[n], d, i, j, k);
          }
        for (int j = 0; j <= ny; ++j)
          for (int i = 0; i < nx; ++i) {
            if ((j == 0 && nb.s < 0) || (j == ny && nb.n < 0)) continue;
            out[n].v(i, j, k) += -pgf_y(PF[n], d, i, j, k);
          }
      }
    }
  }

  // compute_slow_tendencies (dry_core.cu:400-539)
  void slow_tendencies(const std::vector<State<R>>& in, const orc_step_cfg& cfg,
                       std::vector<State<R>>& out) {
    std::vector<State<R>> C = halo_clones(in);
    std::vector<Field<R>> PF;
    for (size_t n = 0; n < L.size(); ++n) {
      PF.emplace_back(C[n].rho.nx, C[n].rho.ny, C[n].rho.nz, C[n].rho.h);
      eos_total(C[n].rho, C[n].rt, PF[n]);
    }
    // theta_ref: FP64 sum of FP-R theta over owned cells (dry_core.cu:462-490)
    const int nz = M.nz;
    std::vector<double> tsum(nz, 0.0);
    std::vector<unsigned long long> tcnt(nz, 0);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k)
      for (size_t n = 0; n < L.size(); ++n)
        for (int j = 0; j < in[n].rho.ny; ++j)
          for (int i = 0; i < in[n].rho.nx; ++i) {
            const R r = in[n].rho(i, j, k);
            const R th = in[n].rt(i, j, k) / std::max(r, R(1.0e-6));
            tsum[k] += static_cast<double>(th);
            tcnt[k] += 1;
          }
    std::vector<R> tref(nz);
    for (int k = 0; k < nz; ++k)
      tref[k] = static_cast<R>(tsum[k] / static_cast<double>(std::max<unsigned long long>(1, tcnt[k])));
    const R dx = static_cast<R>(M.dx), dy = static_cast<R>(M.dy);
    const R g = static_cast<R>(cfg.gravity);
    for (size_t n = 0; n < L.size(); ++n) {
      out[n].rho.fill(0);
      out[n].rt.fill(0);
      out[n].u.fill(0);
      out[n].v.fill(0);
      out[n].w.fill(0);
      const Desc& d = L[n];
      const State<R>& c = C[n];
      // K2 (dry_core.cu:62-144)
#pragma omp parallel for schedule(static)
      for (int k = 0; k < d.nz; ++k)
        for (int j = 0; j < d.nyl(); ++j)
          for (int i = 0; i < d.nxl(); ++i) {
            const R mw = c.u(i, j, k), me = c.u(i + 1, j, k);
            const R ms = c.v(i, j, k), mn = c.v(i, j + 1, k);
            const R mb = c.w(i, j, k), mt = c.w(i, j, k + 1);
            const R idz = M.idz_c[M.mi(d.i0 + i, d.j0 + j, k)];
            const R div = (me - mw) / dx + (mn - ms) / dy + (mt - mb) * idz;
            out[n].rho(i, j, k) = -div;
            const R thc = theta_of(c.rho(i, j, k), c.rt(i, j, k));
            const R thw = theta_of(c.rho(i - 1, j, k), c.rt(i - 1, j, k));
            const R the = theta_of(c.rho(i + 1, j, k), c.rt(i + 1, j, k));
            const R ths = theta_of(c.rho(i, j - 1, k), c.rt(i, j - 1, k));
            const R thn = theta_of(c.rho(i, j + 1, k), c.rt(i, j + 1, k));
            const R thb = k > 0 ? theta_of(c.rho(i, j, k - 1), c.rt(i, j, k - 1)) : thc;
            const R tht = k + 1 < d.nz ? theta_of(c.rho(i, j, k + 1), c.rt(i, j, k + 1)) : thc;
            const R fw = mw * (mw >= R(0) ? thw : thc);
            const R fe = me * (me >= R(0) ? thc : the);
            const R fs = ms * (ms >= R(0) ? ths : thc);
            const R fn = mn * (mn >= R(0) ? thc : thn);
            const R fb = mb * (mb >= R(0) ? thb : thc);
            const R ft = mt * (mt >= R(0) ? thc : tht);
            out[n].rt(i, j, k) = -((fe - fw) / dx + (fn - fs) / dy + (ft - fb) * idz);
          }
    }
    momentum_flux(C, out);
    pressure_gradient(PF, out);
    // K3 buoyancy (dry_core.cu:146-187)
    for (size_t n = 0; n < L.size(); ++n) {
      const Desc& d = L[n];
      const State<R>& c = C[n];
#pragma omp parallel for schedule(static)
      for (int k = 1; k < d.nz; ++k)
        for (int j = 0; j < d.nyl(); ++j)
          for (int i = 0; i < d.nxl(); ++i) {
            const R rl = c.rho(i, j, k - 1), ru = c.rho(i, j, k);
            const R tl = c.rt(i, j, k - 1) / std::max(rl, R(1.0e-6));
            const R tu = c.rt(i, j, k) / std::max(ru, R(1.0e-6));
            const R tf = R(0.5) * (tl + tu);
            const R trf = R(0.5) * (tref[k - 1] + tref[k]);
            const R rf = R(0.5) * (rl + ru);
            const R b = g * (tf - trf) / std::max(trf, R(1.0e-6));
            out[n].w(i, j, k) += rf * b;
          }
    }
  }

  void slow(const orc_step_cfg& cfg) override { slow_tendencies(S, cfg, T); }
  void mflux_only() override {
    // add_dry_momentum_flux_tendencies on the current fields, which the caller
    // has already halo-exchanged (test_dry_momentum_flux.cpp:41-82 usage).
    momentum_flux(S, T);
  }
  void pgf_only() override { pressure_gradient(P, T); }

  // apply_local_split_explicit_fast_modes (dry_fast_modes.cu:400-656)
  void fast_modes(std::vector<State<R>>& st, const orc_step_cfg& cfg) {
    const int ns = std::max(1, cfg.fast_substeps);
    const R dtf = static_cast<R>(cfg.dt) / static_cast<R>(ns);
    const R g = static_cast<R>(cfg.gravity);
    const R dx = static_cast<R>(M.dx), dy = static_cast<R>(M.dy);
    const size_t nr = L.size();
    std::vector<Field<R>> th(nr), rtf(nr), rh(nr), pf(nr), rref(nr), pref(nr);
    std::vector<State<R>> base(nr);
    for (size_t n = 0; n < nr; ++n) {
      th[n] = st[n].rt;
      for (size_t a = 0; a < th[n].total(); ++a) th[n].d[a] = theta_of(st[n].rho.d[a], st[n].rt.d[a]);  // K9
      rtf[n] = Field<R>(st[n].rho.nx, st[n].rho.ny, st[n].rho.nz, st[n].rho.h);
      pf[n] = rtf[n];
      rref[n] = rtf[n];
      pref[n] = rtf[n];
      base[n].u = st[n].u;
      base[n].v = st[n].v;
      base[n].w = st[n].w;
    }
    {
      auto pt = ptrs([&](size_t n) -> Field<R>& { return th[n]; });
      exchange(pt, L, CELL);
      auto bu = ptrs([&](size_t n) -> Field<R>& { return base[n].u; });
      auto bv = ptrs([&](size_t n) -> Field<R>& { return base[n].v; });
      auto bw = ptrs([&](size_t n) -> Field<R>& { return base[n].w; });
      exchange(bu, L, XF);
      exchange(bv, L, YF);
      exchange(bw, L, ZF);
      sync_interfaces(bu, L, XF);
      sync_interfaces(bv, L, YF);
      exchange(bu, L, XF);
      exchange(bv, L, YF);
    }
    long hcount = 0;
    for (const auto& d : L) hcount += static_cast<long>(d.nxl()) * d.nyl();
    const R inv_count = R(1) / static_cast<R>(std::max<long>(1, hcount));
    const int nz = M.nz;
    std::vector<R> rlev(nz), plev(nz);
    for (int s = 0; s < ns; ++s) {
      for (size_t n = 0; n < nr; ++n) rh[n] = st[n].rho;
      {
        auto pr = ptrs([&](size_t n) -> Field<R>& { return rh[n]; });
        exchange(pr, L, CELL);
      }
      for (size_t n = 0; n < nr; ++n) {
        for (size_t a = 0; a < rh[n].total(); ++a) rtf[n].d[a] = rh[n].d[a] * th[n].d[a];  // K10
        eos_total(rh[n], rtf[n], pf[n]);                                                   // K20
      }
      if (!M.terrain_active) {
        // K6/K7, deterministic FP64 accumulation (see file header)
#pragma omp parallel for schedule(static)
        for (int k = 0; k < nz; ++k) {
          double sr = 0.0, sp = 0.0;
          for (size_t n = 0; n < nr; ++n)
            for (int j = 0; j < rh[n].ny; ++j)
              for (int i = 0; i < rh[n].nx; ++i) {
                sr += static_cast<double>(rh[n](i, j, k));
                sp += static_cast<double>(pf[n](i, j, k));
              }
          rlev[k] = static_cast<R>(sr) * inv_count;
          plev[k] = static_cast<R>(sp) * inv_count;
        }
      }
      for (size_t n = 0; n < nr; ++n) {
        const Desc& d = L[n];
        const Nbr nb = neighbors(L, d);
        const int nx = d.nxl(), ny = d.nyl();
        if (M.terrain_active) {
          // K11 column hydrostatic reference (dry_fast_modes.cu:108-158)
#pragma omp parallel for schedule(static)
          for (int j = 0; j < ny; ++j) {
            std::vector<R> thc(nz), zcc(nz), pc(nz), rc(nz);
            for (int i = 0; i < nx; ++i) {
              const int gi = wi(d.i0 + i), gj = wj(d.j0 + j);
              for (int k = 0; k < nz; ++k) {
                thc[k] = theta_of(rh[n](i, j, k), rtf[n](i, j, k));
                zcc[k] = M.zc_m[M.mi(gi, gj, k)];
              }
              column_reference<R>(nz, thc.data(), zcc.data(), g, pf[n](i, j, 0), pc.data(),
                                  rc.data());
              for (int k = 0; k < nz; ++k) {
                pref[n](i, j, k) = pc[k];
                rref[n](i, j, k) = rc[k];
              }
            }
          }
        } else {
          // K8
          for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
              for (int i = 0; i < nx; ++i) {
                rref[n](i, j, k) = rlev[k];
                pref[n](i, j, k) = plev[k];
              }
        }
        if (cfg.fast_update_horizontal_momentum) {
          // K12, K13
#pragma omp parallel for schedule(static)
          for (int k = 0; k < nz; ++k) {
            for (int j = 0; j < ny; ++j)
              for (int i = 0; i <= nx; ++i) {
                if ((i == 0 && nb.w < 0) || (i == nx && nb.e < 0)) continue;
                st[n].u(i, j, k) += -dtf * pgf_x(pf[n], d, i, j, k);
              }
            for (int j = 0; j <= ny; ++j)
              for (int i = 0; i < nx; ++i) {
                if ((j == 0 && nb.s < 0) || (j == ny && nb.n < 0)) continue;
                st[n].v(i, j, k) += -dtf * pgf_y(pf[n], d, i, j, k);
              }
          }
        }
        // K14 (dry_fast_modes.cu:303-348)
#pragma omp parallel for schedule(static)
        for (int k = 1; k < nz; ++k)
          for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
              const R ppl = pf[n](i, j, k - 1) - pref[n](i, j, k - 1);
              const R ppu = pf[n](i, j, k) - pref[n](i, j, k);
              const R rpf = R(0.5) * ((rh[n](i, j, k - 1) - rref[n](i, j, k - 1)) +
                                      (rh[n](i, j, k) - rref[n](i, j, k)));
              const R idz = M.idz_f[M.mi(wi(d.i0 + i), wj(d.j0 + j), k)];
              const R tend = -(ppu - ppl) * idz - g * rpf;
              st[n].w(i, j, k) += dtf * tend;
            }
      }
      {
        auto fu = ptrs([&](size_t n) -> Field<R>& { return st[n].u; });
        auto fv = ptrs([&](size_t n) -> Field<R>& { return st[n].v; });
        auto fw = ptrs([&](size_t n) -> Field<R>& { return st[n].w; });
        exchange(fu, L, XF);
        exchange(fv, L, YF);
        exchange(fw, L, ZF);
        sync_interfaces(fu, L, XF);
        sync_interfaces(fv, L, YF);
        exchange(fu, L, XF);
        exchange(fv, L, YF);
      }
      for (size_t n = 0; n < nr; ++n) {
        const Desc& d = L[n];
        if (cfg.fast_update_density) {
          // K15 (dry_fast_modes.cu:350-394)
#pragma omp parallel for schedule(static)
          for (int k = 0; k < nz; ++k)
            for (int j = 0; j < d.nyl(); ++j)
              for (int i = 0; i < d.nxl(); ++i) {
                const State<R>& s = st[n];
                const State<R>& b = base[n];
                const R div =
                    ((s.u(i + 1, j, k) - b.u(i + 1, j, k)) - (s.u(i, j, k) - b.u(i, j, k))) / dx +
                    ((s.v(i, j + 1, k) - b.v(i, j + 1, k)) - (s.v(i, j, k) - b.v(i, j, k))) / dy +
                    ((s.w(i, j, k + 1) - b.w(i, j, k + 1)) - (s.w(i, j, k) - b.w(i, j, k))) *
                        M.idz_c[M.mi(wi(d.i0 + i), wj(d.j0 + j), k)];
                st[n].rho(i, j, k) = std::max(R(1.0e-6), st[n].rho(i, j, k) - dtf * div);
              }
        }
        for (size_t a = 0; a < st[n].rho.total(); ++a) st[n].rt.d[a] = st[n].rho.d[a] * th[n].d[a];  // K10
      }
    }
  }
  void fast(const orc_step_cfg& cfg) override { fast_modes(S, cfg); }

  // Boundary strips (boundary_conditions.cpp:9-161) and relaxation (extension).
  void set_boundary() override {
    B = S;
    QB = Q;
    have_boundary = true;
  }
  void boundary_now(int mode, double W) override {
    apply_boundary(S, mode, W);
    if (mode == ORC_BOUNDARY_STRIPS)
      for (int t = 0; t < n_tracers; ++t) {
        std::vector<Field<R>> f(L.size());
        for (size_t n = 0; n < L.size(); ++n) f[n] = Q[n][t];
        apply_tracer_boundary(f, t, mode);
        for (size_t n = 0; n < L.size(); ++n) Q[n][t] = f[n];
      }
  }
  static void strip_field(Field<R>& t, const Field<R>& s, bool w, bool e, bool so, bool no) {
    if (w)
      for (int k = 0; k < t.nz; ++k)
        for (int j = 0; j < t.ny; ++j) t(0, j, k) = s(0, j, k);
    if (e)
      for (int k = 0; k < t.nz; ++k)
        for (int j = 0; j < t.ny; ++j) t(t.nx - 1, j, k) = s(t.nx - 1, j, k);
    if (so)
      for (int k = 0; k < t.nz; ++k)
        for (int i = 0; i < t.nx; ++i) t(i, 0, k) = s(i, 0, k);
    if (no)
      for (int k = 0; k < t.nz; ++k)
        for (int i = 0; i < t.nx; ++i) t(i, t.ny - 1, k) = s(i, t.ny - 1, k);
  }
  // Extension: blend toward the boundary state with weight
  // a(d) = max(0, 1 - d/W)^2 where d is the cell (or face) distance from the
  // outermost owned strip; a(0) = 1 reproduces the strip overwrite.
  static void relax_field(Field<R>& t, const Field<R>& s, bool w, bool e, bool so, bool no,
                          double W) {
    for (int k = 0; k < t.nz; ++k)
      for (int j = 0; j < t.ny; ++j)
        for (int i = 0; i < t.nx; ++i) {
          double dmin = 1e30;
          if (w) dmin = std::min(dmin, static_cast<double>(i));
          if (e) dmin = std::min(dmin, static_cast<double>(t.nx - 1 - i));
          if (so) dmin = std::min(dmin, static_cast<double>(j));
          if (no) dmin = std::min(dmin, static_cast<double>(t.ny - 1 - j));
          if (dmin >= W) continue;
          const double a0 = 1.0 - dmin / W;
          const R a = static_cast<R>(a0 * a0);
          t(i, j, k) = t(i, j, k) + a * (s(i, j, k) - t(i, j, k));
        }
  }
  bool tw(const Desc& d) const { return !d.px && d.cx == 0; }  // boundary_conditions.cpp:95-109
  bool te(const Desc& d) const { return !d.px && d.cx == d.rx - 1; }
  bool ts(const Desc& d) const { return !d.py && d.cy == 0; }
  bool tn(const Desc& d) const { return !d.py && d.cy == d.ry - 1; }
  void apply_boundary(std::vector<State<R>>& st, int mode, double W) {
    if (mode == ORC_BOUNDARY_NONE) return;
    require(have_boundary, "boundary state not set");
    for (size_t n = 0; n < L.size(); ++n) {
      const Desc& d = L[n];
      const bool w = tw(d), e = te(d), s = ts(d), no = tn(d);
      Field<R>* tf[5] = {&st[n].rho, &st[n].rt, &st[n].u, &st[n].v, &st[n].w};
      const Field<R>* sf[5] = {&B[n].rho, &B[n].rt, &B[n].u, &B[n].v, &B[n].w};
      for (int f = 0; f < 5; ++f) {
        if (mode == ORC_BOUNDARY_STRIPS) strip_field(*tf[f], *sf[f], w, e, s, no);
        else relax_field(*tf[f], *sf[f], w, e, s, no, W);
      }
    }
  }

  // advance_dry_state_ssprk3 (dry_core.cu:541-651)
  void ssprk3(const orc_step_cfg& cfg, int bmode, double W, double /*tau*/) override {
    const R dt = static_cast<R>(cfg.dt);
    const size_t nr = L.size();
    std::vector<State<R>> q0 = S, q1(nr), q2(nr), tmp(nr), tend(nr);
    for (size_t n = 0; n < nr; ++n) {
      q1[n] = make_state(L[n]);
      q2[n] = make_state(L[n]);
      tmp[n] = make_state(L[n]);
      tend[n] = make_state(L[n]);
    }
    auto update = [&](State<R>& out, const State<R>& base, const State<R>& t) {
      auto up = [&](Field<R>& o, const Field<R>& b, const Field<R>& tt, R floor) {
        const long nn = static_cast<long>(o.total());
#pragma omp parallel for schedule(static)
        for (long a = 0; a < nn; ++a) o.d[a] = std::max(floor, b.d[a] + dt * tt.d[a]);
      };
      up(out.rho, base.rho, t.rho, R(1.0e-6));
      up(out.rt, base.rt, t.rt, R(1.0e-6));
      up(out.u, base.u, t.u, K<R>::mom_floor);
      up(out.v, base.v, t.v, K<R>::mom_floor);
      up(out.w, base.w, t.w, K<R>::mom_floor);
    };
    auto blend = [&](State<R>& out, const State<R>& a, const State<R>& b, R wa, R wb) {
      auto bl = [&](Field<R>& o, const Field<R>& x, const Field<R>& y) {
        const long nn = static_cast<long>(o.total());
#pragma omp parallel for schedule(static)
        for (long i = 0; i < nn; ++i) o.d[i] = wa * x.d[i] + wb * y.d[i];
      };
      bl(out.rho, a.rho, b.rho);
      bl(out.rt, a.rt, b.rt);
      bl(out.u, a.u, b.u);
      bl(out.v, a.v, b.v);
      bl(out.w, a.w, b.w);
    };
    apply_boundary(S, bmode, W);
    slow_tendencies(S, cfg, tend);
    for (size_t n = 0; n < nr; ++n) update(q1[n], S[n], tend[n]);
    fast_modes(q1, cfg);
    apply_boundary(q1, bmode, W);
    slow_tendencies(q1, cfg, tend);
    for (size_t n = 0; n < nr; ++n) {
      update(tmp[n], q1[n], tend[n]);
      blend(q2[n], q0[n], tmp[n], R(0.75), R(0.25));
    }
    fast_modes(q2, cfg);
    apply_boundary(q2, bmode, W);
    slow_tendencies(q2, cfg, tend);
    for (size_t n = 0; n < nr; ++n) {
      update(tmp[n], q2[n], tend[n]);
      blend(S[n], q0[n], tmp[n], R(1) / R(3), R(2) / R(3));
    }
    fast_modes(S, cfg);
  }

  // advance_passive_tracers_ssprk3 (dry_transport.cu:440-518), K23 (49-125)
  void tracer_stage(const Field<R>& qin, const State<R>& c, const Desc& d, R dt, Field<R>& out) {
    const R dx = static_cast<R>(M.dx), dy = static_cast<R>(M.dy);
    const int nx = d.nxl(), ny = d.nyl(), nz = d.nz;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          auto q = [&](int ii, int jj, int kk) {
            return qin(ii, jj, kk) / std::max(c.rho(ii, jj, kk), R(1.0e-6));
          };
          const R qc = q(i, j, k), qw = q(i - 1, j, k), qe = q(i + 1, j, k);
          const R qs = q(i, j - 1, k), qn = q(i, j + 1, k);
          const R qb = k > 0 ? q(i, j, k - 1) : qc;
          const R qt = k + 1 < nz ? q(i, j, k + 1) : qc;
          const R mw = c.u(i, j, k), me = c.u(i + 1, j, k);
          const R ms = c.v(i, j, k), mn = c.v(i, j + 1, k);
          const R mb = c.w(i, j, k), mt = c.w(i, j, k + 1);
          const R idz = M.idz_c[M.mi(d.i0 + i, d.j0 + j, k)];
          const R fw = mw * (mw >= R(0) ? qw : qc);
          const R fe = me * (me >= R(0) ? qc : qe);
          const R fs = ms * (ms >= R(0) ? qs : qc);
          const R fn = mn * (mn >= R(0) ? qc : qn);
          const R fb = mb * (mb >= R(0) ? qb : qc);
          const R ft = mt * (mt >= R(0) ? qc : qt);
          const R div = (fe - fw) / dx + (fn - fs) / dy + (ft - fb) * idz;
          out(i, j, k) = std::max(R(0), qin(i, j, k) - dt * div);
        }
  }
  void apply_tracer_boundary(std::vector<Field<R>>& f, int t, int mode) {
    if (mode == ORC_BOUNDARY_NONE) return;
    // Tracers outside the mask keep their strip values: the prepared-case
    // tracer updater overwrites only the tracers its boundary snapshot
    // carries (prepared_case_init.cpp:567-638).
    if (t < 32 && !((tracer_strip_mask >> t) & 1u)) return;
    require(have_boundary, "boundary state not set");
    for (size_t n = 0; n < L.size(); ++n) {
      const Desc& d = L[n];
      strip_field(f[n], QB[n][t], tw(d), te(d), ts(d), tn(d));
    }
  }
  void tracers(const orc_step_cfg& cfg, int bmode) override {
    if (n_tracers == 0) return;
    const R dt = static_cast<R>(cfg.dt);
    const size_t nr = L.size();
    std::vector<State<R>> C = halo_clones(S);  // prepare_dry_transport_halos
    for (int t = 0; t < n_tracers; ++t) {
      std::vector<Field<R>> q0(nr), q1(nr), q2(nr), tmp(nr), stage(nr);
      for (size_t n = 0; n < nr; ++n) {
        q0[n] = Q[n][t];
        stage[n] = Q[n][t];
        q1[n] = make_field<R>(L[n], CELL);
        q2[n] = make_field<R>(L[n], CELL);
        tmp[n] = make_field<R>(L[n], CELL);
      }
      auto ex = [&](std::vector<Field<R>>& f) {
        auto p = ptrs([&](size_t n) -> Field<R>& { return f[n]; });
        exchange(p, L, CELL);
      };
      auto bl = [&](Field<R>& o, const Field<R>& x, const Field<R>& y, R wa, R wb) {
        for (size_t i = 0; i < o.total(); ++i) o.d[i] = wa * x.d[i] + wb * y.d[i];
      };
      apply_tracer_boundary(stage, t, bmode);
      ex(stage);
      for (size_t n = 0; n < nr; ++n) tracer_stage(stage[n], C[n], L[n], dt, q1[n]);
      apply_tracer_boundary(q1, t, bmode);
      ex(q1);
      for (size_t n = 0; n < nr; ++n) {
        tracer_stage(q1[n], C[n], L[n], dt, tmp[n]);
        bl(q2[n], q0[n], tmp[n], R(0.75), R(0.25));
      }
      apply_tracer_boundary(q2, t, bmode);
      ex(q2);
      for (size_t n = 0; n < nr; ++n) {
        tracer_stage(q2[n], C[n], L[n], dt, tmp[n]);
        bl(Q[n][t], q0[n], tmp[n], R(1) / R(3), R(2) / R(3));
      }
    }
  }

  // apply_warm_rain_microphysics (warm_rain.cu:36-120 K24, 122-191 K25/K26,
  // 214-353 drivers).  Extension: accretion qc -> qr at rate k_a*qc*qr*dt
  // (k_a = 0 reproduces the reference).
  void warm_rain(const orc_warm_rain_cfg& cfgd, int metric) override {
    require(n_tracers >= 3, "warm rain requires 3 tracers (qv, qc, qr)");
    const R dt = static_cast<R>(cfgd.dt);
    const R relax = static_cast<R>(cfgd.condensation_relaxation);
    const R thr = static_cast<R>(cfgd.cloud_autoconversion_threshold);
    const R krate = static_cast<R>(cfgd.cloud_autoconversion_rate);
    const R erate = static_cast<R>(cfgd.rain_evaporation_rate);
    const R vt = static_cast<R>(cfgd.rain_terminal_velocity);
    const R depth = static_cast<R>(cfgd.sedimentation_layer_depth_m);
    const R kacc = static_cast<R>(cfgd.accretion_rate);
    const bool heat = cfgd.enable_latent_heating != 0;
    const R L_cp = K<R>::Lv;
    for (size_t n = 0; n < L.size(); ++n) {
      const Desc& d = L[n];
      State<R>& s = S[n];
      Field<R>& qvf = Q[n][0];
      Field<R>& qcf = Q[n][1];
      Field<R>& qrf = Q[n][2];
#pragma omp parallel for schedule(static)
      for (int k = 0; k < d.nz; ++k)
        for (int j = 0; j < d.nyl(); ++j)
          for (int i = 0; i < d.nxl(); ++i) {
            const R rho = std::max(s.rho(i, j, k), R(1.0e-6));
            const R rt0 = s.rt(i, j, k);
            R th = rt0 / rho;
            R p = pressure_of(rho, rt0);
            R T = temperature_of(rho, rt0);
            const R ex = std::max(T / std::max(th, R(1.0e-6)), R(1.0e-6));
            R qv = std::max(qvf(i, j, k) / rho, R(0));
            R qc = std::max(qcf(i, j, k) / rho, R(0));
            R qr = std::max(qrf(i, j, k) / rho, R(0));
            R qs = qsat(T, p);
            const R lfac = L_cp / (cp<R>() * std::max(ex, R(1.0e-6)));
            if (qv > qs) {
              const R c = std::min((qv - qs) * std::max(relax, R(0)), qv);
              qv -= c;
              qc += c;
              if (heat && c > R(0)) th += lfac * c;
            } else if (qv < qs && qc > R(0)) {
              const R e = std::min((qs - qv) * std::max(relax, R(0)), qc);
              qv += e;
              qc -= e;
              if (heat && e > R(0)) th -= lfac * e;
            }
            const R exc = std::max(qc - thr, R(0));
            if (exc > R(0) && krate > R(0)) {
              const R cv = std::min(exc, dt * krate * exc);
              qc -= cv;
              qr += cv;
            }
            if (kacc > R(0) && qc > R(0) && qr > R(0)) {  // extension
              const R a = std::min(qc, dt * kacc * qc * qr);
              qc -= a;
              qr += a;
            }
            p = pressure_of(rho, rho * th);
            T = th * std::pow(std::max(p / K<R>::p0, R(1.0e-6)), K<R>::kappa);
            qs = qsat(T, p);
            if (qr > R(0) && qv < qs && erate > R(0)) {
              const R e = std::min(qr, dt * erate * std::min(qs - qv, qr));
              qr -= e;
              qv += e;
              if (heat && e > R(0)) th -= lfac * e;
            }
            qv = std::max(qv, R(0));
            qc = std::max(qc, R(0));
            qr = std::max(qr, R(0));
            qvf(i, j, k) = rho * qv;
            qcf(i, j, k) = rho * qc;
            qrf(i, j, k) = rho * qr;
            s.rt(i, j, k) = rho * th;
          }
      if (metric == 0) {
        // K25 nominal fallout (warm_rain.cu:122-154), runs when v_t > 0 and
        // a layer depth is set (warm_rain.cu:256-268)
        if (!(vt > R(0) && depth > R(0))) continue;
        const R dz = std::max(depth, R(1.0e-3));
        const R frac = std::min(std::max(vt * dt / dz, R(0)), R(1));
        if (frac <= R(0)) continue;
        for (int j = 0; j < d.nyl(); ++j)
          for (int i = 0; i < d.nxl(); ++i) {
            R inc = 0;
            for (int k = d.nz - 1; k >= 0; --k) {
              const R old = std::max(qrf(i, j, k), R(0)) * dz;
              const R out = frac * old;
              qrf(i, j, k) = (old - out + inc) / dz;
              inc = out;
            }
            precip[n](i, j, 0) += inc;
          }
      } else {
        // K26 metric fallout (warm_rain.cu:156-191), runs when v_t > 0
        if (!(vt > R(0))) continue;
        for (int j = 0; j < d.nyl(); ++j)
          for (int i = 0; i < d.nxl(); ++i) {
            R inc = 0;
            for (int k = d.nz - 1; k >= 0; --k) {
              const R idz = std::max(M.idz_c[M.mi(d.i0 + i, d.j0 + j, k)], R(1.0e-6));
              const R dz = R(1) / idz;
              const R frac = std::min(std::max(vt * dt * idz, R(0)), R(1));
              const R old = std::max(qrf(i, j, k), R(0)) * dz;
              const R out = frac * old;
              qrf(i, j, k) = (old - out + inc) * idz;
              inc = out;
            }
            precip[n](i, j, 0) += inc;
          }
      }
    }
  }

  // extract_dry_plan_view / extract_runtime_plan_view
  // (src/io/plan_view_output.cpp:601-839): derived fields at the slice
  // level (theta, p, T, cell-centred velocities, wind speed, q, RH,
  // dewpoint, qc, qr, condensate) and column maps (cloud and rain water
  // paths, condensate, rain fraction, maximum synthetic reflectivity,
  // surface bucket and its mean rate), [field][j][i] in the reference's
  // field order.  Returns the field count: 10, 13, 21 or 23.
  int plan_view(int slice_k, int steps, double dt, int with_precip, double* out) override {
    const int nx = M.nx, ny = M.ny;
    const int k = std::clamp(slice_k, 0, M.nz - 1);
    const size_t plane = static_cast<size_t>(nx) * ny;
    const int nf = n_tracers == 0 ? 10 : (n_tracers < 3 ? 13 : (with_precip ? 23 : 21));
    const R min_rho = R(1.0e-6), min_p = R(1);
    auto put = [&](int f, int gi, int gj, double v) {
      out[static_cast<size_t>(f) * plane + static_cast<size_t>(gj) * nx + gi] = v;
    };
    // moist_thermo.hpp:38-74
    auto rh_of = [](R q, R T, R p) {
      const R qs = qsat(T, p);
      return std::min(std::max(qs > R(0) ? R(100) * q / qs : R(0), R(0)), R(100));
    };
    auto dewpoint_of = [](R q_in, R p) {
      const R q = std::min(std::max(q_in, R(0)), R(1));
      const R e = p * q / std::max(K<R>::eps + (R(1) - K<R>::eps) * q, R(1.0e-6));
      const R e_hpa = std::max(e * R(0.01), R(1.0e-6));
      const R lt = std::log(e_hpa / R(6.112));
      return (R(243.5) * lt) / (R(17.67) - lt) + R(273.15);
    };
    auto dbz_of = [](R rho, R qr_in) {
      const R qr = std::max(qr_in, R(0));
      if (qr <= R(1.0e-8) || rho <= R(1.0e-8)) return R(-20);
      const R z = R(3.63e9) * std::pow(rho * qr, R(1.75));
      return R(10) * std::log10(std::max(z, R(1.0e-6)));
    };
    const double hours = std::max(static_cast<double>(steps) * dt / 3600.0, 1.0e-9);
    for (size_t n = 0; n < L.size(); ++n) {
      const State<R>& st = S[n];
      const Desc& d = L[n];
      for (int j = 0; j < d.nyl(); ++j)
        for (int i = 0; i < d.nxl(); ++i) {
          const int gi = d.i0 + i, gj = d.j0 + j;
          const R rho = st.rho(i, j, k), rt = st.rt(i, j, k);
          const R rr = std::max(rho, min_rho);
          const R theta = rt / rr;
          const R p = pressure_of(rho, rt);
          const R T = theta * std::pow(std::max(p, min_p) / K<R>::p0, K<R>::kappa);
          const R u = R(0.5) * (st.u(i, j, k) + st.u(i + 1, j, k)) / rr;
          const R v = R(0.5) * (st.v(i, j, k) + st.v(i, j + 1, k)) / rr;
          const R w = R(0.5) * (st.w(i, j, k) + st.w(i, j, k + 1)) / rr;
          put(0, gi, gj, static_cast<double>(M.terrain[static_cast<size_t>(gj) * nx + gi]));
          put(1, gi, gj, static_cast<double>(M.zc_m[M.mi(gi, gj, k)]));
          put(2, gi, gj, static_cast<double>(rho));
          put(3, gi, gj, static_cast<double>(theta));
          put(4, gi, gj, static_cast<double>(p));
          put(5, gi, gj, static_cast<double>(T));
          put(6, gi, gj, static_cast<double>(u));
          put(7, gi, gj, static_cast<double>(v));
          put(8, gi, gj, static_cast<double>(std::sqrt(u * u + v * v)));
          put(9, gi, gj, static_cast<double>(w));
          if (n_tracers == 0) continue;
          const R q = Q[n][0](i, j, k) / rr;
          put(10, gi, gj, static_cast<double>(q));
          put(11, gi, gj, static_cast<double>(rh_of(q, T, p)));
          put(12, gi, gj, static_cast<double>(dewpoint_of(q, p)));
          if (n_tracers < 3) continue;
          const R qc = Q[n][1](i, j, k) / rr, qr = Q[n][2](i, j, k) / rr;
          put(13, gi, gj, static_cast<double>(qc));
          put(14, gi, gj, static_cast<double>(qr));
          put(15, gi, gj, static_cast<double>(qc + qr));
          double col_c = 0.0, col_r = 0.0;
          R dbz = R(-20);
          for (int kk = 0; kk < d.nz; ++kk) {
            const double dz = static_cast<double>(R(1) / M.idz_c[M.mi(gi, gj, kk)]);
            col_c += static_cast<double>(Q[n][1](i, j, kk)) * dz;
            col_r += static_cast<double>(Q[n][2](i, j, kk)) * dz;
            const R rk = st.rho(i, j, kk);
            dbz = std::max(dbz, dbz_of(rk, Q[n][2](i, j, kk) / std::max(rk, min_rho)));
          }
          const double total = col_c + col_r, tpos = std::max(total, 0.0);
          put(16, gi, gj, col_c);
          put(17, gi, gj, col_r);
          put(18, gi, gj, total);
          put(19, gi, gj, tpos <= 0.0 ? 0.0 : std::clamp(col_r / tpos, 0.0, 1.0));
          put(20, gi, gj, static_cast<double>(dbz));
          if (!with_precip) continue;
          const double acc = static_cast<double>(precip[n](i, j, 0));
          put(21, gi, gj, acc);
          put(22, gi, gj, acc / hours);
        }
    }
    return nf;
  }

  // summarize_dry_states (dry_diagnostics.cpp:11-96) and
  // summarize_runtime_state (runtime_summary.cpp:27-178).
  void summary(double* o) override {
    for (int a = 0; a < ORC_SUMMARY_N; ++a) o[a] = 0.0;
    double mn[5], mx[5];
    for (int a = 0; a < 5; ++a) {
      mn[a] = 1e300;
      mx[a] = -1e300;
    }
    double qmn[3] = {1e300, 1e300, 1e300}, qmx[3] = {-1e300, -1e300, -1e300};
    long cells = 0, wet = 0;
    double psum = 0.0, pmax = 0.0;
    for (size_t n = 0; n < L.size(); ++n) {
      const State<R>& s = S[n];
      const Desc& d = L[n];
      o[ORC_S_TOTAL_DRY_MASS] += s.rho.owned_sum();
      o[ORC_S_TOTAL_RHO_THETA] += s.rt.owned_sum();
      o[ORC_S_TOTAL_MOM_X] += s.u.owned_sum();
      o[ORC_S_TOTAL_MOM_Y] += s.v.owned_sum();
      o[ORC_S_TOTAL_MOM_Z] += s.w.owned_sum();
      for (int k = 0; k < s.rho.nz; ++k)
        for (int j = 0; j < s.rho.ny; ++j)
          for (int i = 0; i < s.rho.nx; ++i) {
            const double r = static_cast<double>(s.rho(i, j, k));
            const double th = static_cast<double>(s.rt(i, j, k)) / std::max(r, 1.0e-12);
            mn[0] = std::min(mn[0], r);
            mx[0] = std::max(mx[0], r);
            mn[1] = std::min(mn[1], th);
            mx[1] = std::max(mx[1], th);
            for (int t = 0; t < std::min(n_tracers, 3); ++t) {
              const double q = static_cast<double>(Q[n][t](i, j, k)) / std::max(r, 1.0e-12);
              qmn[t] = std::min(qmn[t], q);
              qmx[t] = std::max(qmx[t], q);
            }
          }
      const Field<R>* ff[3] = {&s.u, &s.v, &s.w};
      for (int f = 0; f < 3; ++f) {
        const Field<R>& x = *ff[f];
        for (int k = 0; k < x.nz; ++k)
          for (int j = 0; j < x.ny; ++j)
            for (int i = 0; i < x.nx; ++i) {
              const double v = static_cast<double>(x(i, j, k));
              mn[2 + f] = std::min(mn[2 + f], v);
              mx[2 + f] = std::max(mx[2 + f], v);
            }
      }
      for (int t = 0; t < std::min(n_tracers, 3); ++t) {
        double tot = 0.0;
        for (int k = 0; k < d.nz; ++k)
          for (int j = 0; j < d.nyl(); ++j)
            for (int i = 0; i < d.nxl(); ++i) {
              const double dz = 1.0 / static_cast<double>(M.idz_c[M.mi(d.i0 + i, d.j0 + j, k)]);
              tot += static_cast<double>(Q[n][t](i, j, k)) * dz;
            }
        o[ORC_S_VAPOR_PATH + t] += tot;
      }
      for (int j = 0; j < d.nyl(); ++j)
        for (int i = 0; i < d.nxl(); ++i) {
          const double v = static_cast<double>(precip[n](i, j, 0));
          psum += v;
          pmax = std::max(pmax, v);
          ++cells;
          if (v > 0.0) ++wet;
        }
    }
    o[ORC_S_RHO_MIN] = mn[0];
    o[ORC_S_RHO_MAX] = mx[0];
    o[ORC_S_THETA_MIN] = mn[1];
    o[ORC_S_THETA_MAX] = mx[1];
    o[ORC_S_U_MIN] = mn[2];
    o[ORC_S_U_MAX] = mx[2];
    o[ORC_S_V_MIN] = mn[3];
    o[ORC_S_V_MAX] = mx[3];
    o[ORC_S_W_MIN] = mn[4];
    o[ORC_S_W_MAX] = mx[4];
    o[ORC_S_PRECIP_SUM] = psum;
    o[ORC_S_PRECIP_MEAN] = cells > 0 ? psum / static_cast<double>(cells) : 0.0;
    o[ORC_S_PRECIP_MAX] = pmax;
    o[ORC_S_PRECIP_CELLS] = static_cast<double>(cells);
    o[ORC_S_PRECIPITATING_CELLS] = static_cast<double>(wet);
    for (int t = 0; t < 3; ++t) {
      o[ORC_S_TRACER_QMIN0 + 2 * t] = t < n_tracers ? qmn[t] : 0.0;
      o[ORC_S_TRACER_QMIN0 + 2 * t + 1] = t < n_tracers ? qmx[t] : 0.0;
    }
  }

  // Initial states.
  // DryState::fill_constant (dry_core.cu:299-306), make_constant_dry_state (348-361)
  void init_constant(double rho, double th, double u, double v, double w) override {
    const R r = static_cast<R>(rho);
    for (auto& s : S) {
      s.rho.fill(r);
      s.rt.fill(r * static_cast<R>(th));
      s.u.fill(r * static_cast<R>(u));
      s.v.fill(r * static_cast<R>(v));
      s.w.fill(r * static_cast<R>(w));
    }
  }
  // hydrostatic_rho_at_height (dry_core.cu:22-35 == idealized_cases.cpp:22-34)
  static R hydro_rho(R rho_s, R th, R z) {
    const R ps = pressure_of(rho_s, rho_s * th);
    const R exs = std::pow(std::max(ps / K<R>::p0, R(1.0e-6)), K<R>::kappa);
    const R ex = std::max(R(1.0e-4), exs - K<R>::g_init * z / (cp<R>() * th));
    const R p = K<R>::p0 * std::pow(ex, R(1) / K<R>::kappa);
    return rho_from_p_theta(p, th);
  }
  // make_hydrostatic_rest_state (dry_core.cu:363-398)
  void init_hydro(double rho_s, double th) override {
    init_constant(0.0, th, 0, 0, 0);
    for (size_t n = 0; n < L.size(); ++n) {
      const Desc& d = L[n];
      for (int k = 0; k < d.nz; ++k)
        for (int j = 0; j < d.nyl(); ++j)
          for (int i = 0; i < d.nxl(); ++i) {
            const R r = hydro_rho(static_cast<R>(rho_s), static_cast<R>(th),
                                  M.zc_m[M.mi(d.i0 + i, d.j0 + j, k)]);
            S[n].rho(i, j, k) = r;
            S[n].rt(i, j, k) = r * static_cast<R>(th);
          }
      S[n].u.fill(0);
      S[n].v.fill(0);
      S[n].w.fill(0);
    }
  }
  // make_bubble_like_state (idealized_cases.cpp:105-162), bubble_weight (14-20).
  // p = rho_bg, theta_bg, u, v, w, cx, cy, cz, rx, ry, rz, dtheta.
  // Extension: rho_bg <= 0 means "hydrostatic background" (hydrostatic rest
  // at rho_surface = -rho_bg, the bubble applied to rho*theta at fixed rho).
  void init_bubble(const double* p, int sign) override {
    const bool hydro = p[0] <= 0.0;
    if (hydro) init_hydro(-p[0], p[1]);
    else init_constant(p[0], p[1], p[2], p[3], p[4]);
    const R dxr = static_cast<R>(M.dx), dyr = static_cast<R>(M.dy);
    const R x0 = static_cast<R>(p[5]) * static_cast<R>(M.nx - 1) * dxr;
    const R y0 = static_cast<R>(p[6]) * static_cast<R>(M.ny - 1) * dyr;
    const R z0 = static_cast<R>(p[7]) * static_cast<R>(M.z_top);
    const R rx = std::max(static_cast<R>(p[8]) * static_cast<R>(M.nx) * dxr, R(1));
    const R ry = std::max(static_cast<R>(p[9]) * static_cast<R>(M.ny) * dyr, R(1));
    const R rz = std::max(static_cast<R>(p[10]) * static_cast<R>(M.z_top), R(1));
    const R dth = static_cast<R>(sign >= 0 ? std::fabs(p[11]) : -std::fabs(p[11]));
    const R thb = static_cast<R>(p[1]);
    for (size_t n = 0; n < L.size(); ++n) {
      const Desc& d = L[n];
      for (int k = 0; k < d.nz; ++k)
        for (int j = 0; j < d.nyl(); ++j) {
          const R y = static_cast<R>(d.j0 + j) * dyr;
          for (int i = 0; i < d.nxl(); ++i) {
            const R x = static_cast<R>(d.i0 + i) * dxr;
            const R z = M.zc_m[M.mi(d.i0 + i, d.j0 + j, k)];
            const R wx = (x - x0) / rx, wy = (y - y0) / ry, wz = (z - z0) / rz;
            const R r2 = wx * wx + wy * wy + wz * wz;
            const R wgt = r2 >= R(1) ? R(0) : R(0.5) * (R(1) + std::cos(K<R>::pi * std::sqrt(r2)));
            const R th = thb + dth * wgt;
            const R rb = hydro ? S[n].rho(i, j, k) : static_cast<R>(p[0]);
            S[n].rt(i, j, k) = rb * th;
          }
        }
    }
  }
  // make_mountain_wave_background_state (idealized_cases.cpp:182-223),
  // set_background_face_momenta (36-103)
  void init_mountain(double rho_s, double th, double u, double v, double w) override {
    init_constant(0.0, th, u, v, w);
    const R U = static_cast<R>(u), V = static_cast<R>(v), W = static_cast<R>(w);
    for (size_t n = 0; n < L.size(); ++n) {
      const Desc& d = L[n];
      State<R>& s = S[n];
      for (int k = 0; k < d.nz; ++k)
        for (int j = 0; j < d.nyl(); ++j)
          for (int i = 0; i < d.nxl(); ++i) {
            const double terr = static_cast<double>(M.terrain[static_cast<size_t>(d.j0 + j) * M.nx + d.i0 + i]);
            const double z = M.z_c[k] + M.w_c[k] * terr;
            const R r = hydro_rho(static_cast<R>(rho_s), static_cast<R>(th), static_cast<R>(z));
            s.rho(i, j, k) = r;
            s.rt(i, j, k) = r * static_cast<R>(th);
          }
      const int nx = s.rho.nx, ny = s.rho.ny, nz = s.rho.nz;
      auto fdx = [&](int i, int j, int k) {
        if (nx == 1 || i <= 0) return s.rho(0, j, k);
        if (i >= nx) return s.rho(nx - 1, j, k);
        return R(0.5) * (s.rho(i - 1, j, k) + s.rho(i, j, k));
      };
      auto fdy = [&](int i, int j, int k) {
        if (ny == 1 || j <= 0) return s.rho(i, 0, k);
        if (j >= ny) return s.rho(i, ny - 1, k);
        return R(0.5) * (s.rho(i, j - 1, k) + s.rho(i, j, k));
      };
      auto fdz = [&](int i, int j, int k) {
        if (nz == 1 || k <= 0) return s.rho(i, j, 0);
        if (k >= nz) return s.rho(i, j, nz - 1);
        return R(0.5) * (s.rho(i, j, k - 1) + s.rho(i, j, k));
      };
      for (int k = 0; k < s.u.nz; ++k)
        for (int j = 0; j < s.u.ny; ++j)
          for (int i = 0; i < s.u.nx; ++i) s.u(i, j, k) = fdx(i, j, k) * U;
      for (int k = 0; k < s.v.nz; ++k)
        for (int j = 0; j < s.v.ny; ++j)
          for (int i = 0; i < s.v.nx; ++i) s.v(i, j, k) = fdy(i, j, k) * V;
      for (int k = 0; k < s.w.nz; ++k)
        for (int j = 0; j < s.w.ny; ++j)
          for (int i = 0; i < s.w.nx; ++i) s.w(i, j, k) = fdz(i, j, k) * W;
    }
  }
};

template <class F>
int guard(F f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

WorldBase* W(void* p) {
  require(p != nullptr, "null world");
  return static_cast<WorldBase*>(p);
}


template <class R>
double thermo_eval(int fn, double a, double b, double c) {
  const R x = static_cast<R>(a), y = static_cast<R>(b), z = static_cast<R>(c);
  switch (fn) {
    case 0: return static_cast<double>(theta_of(x, y));
    case 1: return static_cast<double>(pressure_of(x, y));
    case 2: return static_cast<double>(rho_from_p_theta(x, y));
    case 3: return static_cast<double>(esat_pa(x));
    case 4: return static_cast<double>(qsat(x, y));
    case 5: return static_cast<double>(temperature_of(x, y));
    case 6: return static_cast<double>(World<R>::hydro_rho(x, y, z));
    case 7: return static_cast<double>(cp<R>());
    case 8: return static_cast<double>(K<R>::eps);
    default: throw std::runtime_error("bad thermo function");
  }
}

}  // namespace orc

using namespace orc;

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

void* orc_world_create(int precision, int nx, int ny, int nz, int halo, int ranks_x,
                       int ranks_y, int periodic_x, int periodic_y, double dx, double dy,
                       double z_top, const double* z_interfaces, const float* terrain,
                       double terrain_taper_eta, int n_tracers) {
  try {
    if (precision == 4)
      return new World<float>(4, nx, ny, nz, halo, ranks_x, ranks_y, periodic_x != 0,
                              periodic_y != 0, dx, dy, z_top, z_interfaces, terrain,
                              terrain_taper_eta, n_tracers);
    if (precision == 8)
      return new World<double>(8, nx, ny, nz, halo, ranks_x, ranks_y, periodic_x != 0,
                               periodic_y != 0, dx, dy, z_top, z_interfaces, terrain,
                               terrain_taper_eta, n_tracers);
    g_err = "precision must be 4 or 8";
  } catch (const std::exception& e) {
    g_err = e.what();
  }
  return nullptr;
}

void orc_world_destroy(void* world) { delete static_cast<WorldBase*>(world); }

int orc_num_ranks(void* world) { return static_cast<int>(W(world)->L.size()); }

int orc_rank_desc(void* world, int rank, int* o) {
  return guard([&] {
    const Desc& d = W(world)->L.at(rank);
    const int v[13] = {d.rank, d.cx, d.cy, d.rx, d.ry, d.gnx, d.gny, d.nz, d.h, d.i0, d.i1, d.j0, d.j1};
    std::memcpy(o, v, sizeof(v));
  });
}
int orc_neighbors(void* world, int rank, int* o) {
  return guard([&] {
    const auto& L = W(world)->L;
    const Nbr n = neighbors(L, L.at(rank));
    o[0] = n.w;
    o[1] = n.e;
    o[2] = n.s;
    o[3] = n.n;
  });
}
long orc_rank_storage_size(void* world, int rank, int field, int* dims4) {
  long r = -1;
  guard([&] { r = W(world)->storage_size(rank, field, dims4); });
  return r;
}
int orc_get_rank_storage(void* w, int rank, int f, void* out) {
  return guard([&] { W(w)->get_rank(rank, f, out); });
}
int orc_set_rank_storage(void* w, int rank, int f, const void* in) {
  return guard([&] { W(w)->set_rank(rank, f, in); });
}
int orc_get_global(void* w, int f, void* out) {
  return guard([&] { W(w)->get_global(f, out); });
}
int orc_set_global(void* w, int f, const void* in) {
  return guard([&] { W(w)->set_global(f, in); });
}
int orc_fill(void* w, int f, double v) {
  return guard([&] { W(w)->fill(f, v); });
}
int orc_get_metric(void* w, int which, void* out) {
  return guard([&] { W(w)->get_metric(which, out); });
}
int orc_exchange(void* w, int f) {
  return guard([&] { W(w)->exch(f); });
}
int orc_sync_interfaces(void* w, int f) {
  return guard([&] { W(w)->sync(f); });
}
int orc_pressure(void* w) {
  return guard([&] { W(w)->pressure(); });
}
int orc_slow_tendencies(void* w, const orc_step_cfg* c) {
  return guard([&] { W(w)->slow(*c); });
}
int orc_momentum_flux_only(void* w) {
  return guard([&] { W(w)->mflux_only(); });
}
int orc_pressure_gradient_only(void* w) {
  return guard([&] { W(w)->pgf_only(); });
}
int orc_fast_modes(void* w, const orc_step_cfg* c) {
  return guard([&] { W(w)->fast(*c); });
}
int orc_set_boundary_from_current(void* w) {
  return guard([&] { W(w)->set_boundary(); });
}
int orc_set_tracer_strip_mask(void* w, unsigned mask) {
  return guard([&] { W(w)->tracer_strip_mask = mask; });
}
int orc_apply_boundary(void* w, int mode, double width) {
  return guard([&] { W(w)->boundary_now(mode, width); });
}

// K11 of one column: theta and centre heights in, p_ref and rho_ref out
// (working precision `precision`, FP64 arrays at the boundary).
int orc_column_reference(int precision, int nz, const double* th, const double* zc, double g,
                         double p_bottom, double* pref, double* rref) {
  return guard([&] {
    require(nz >= 1 && th && zc && pref && rref, "bad column arguments");
    auto run = [&](auto r) {
      using R = decltype(r);
      std::vector<R> t(th, th + nz), z(zc, zc + nz), p(nz), q(nz);
      column_reference<R>(nz, t.data(), z.data(), static_cast<R>(g), static_cast<R>(p_bottom),
                          p.data(), q.data());
      for (int k = 0; k < nz; ++k) {
        pref[k] = static_cast<double>(p[k]);
        rref[k] = static_cast<double>(q[k]);
      }
    };
    if (precision == 4) run(float{});
    else run(double{});
  });
}

double orc_thermo(int precision, int fn, double a, double b, double c) {
  double r = 0.0;
  const int rc = guard([&] { r = precision == 4 ? thermo_eval<float>(fn, a, b, c) : thermo_eval<double>(fn, a, b, c); });
  return rc == 0 ? r : std::nan("");
}

int orc_ssprk3_step(void* w, const orc_step_cfg* c, int bmode, double width, double tau) {
  return guard([&] { W(w)->ssprk3(*c, bmode, width, tau); });
}
int orc_tracers_ssprk3(void* w, const orc_step_cfg* c, int bmode) {
  return guard([&] { W(w)->tracers(*c, bmode); });
}
int orc_warm_rain(void* w, const orc_warm_rain_cfg* c, int metric) {
  return guard([&] { W(w)->warm_rain(*c, metric); });
}
int orc_summary(void* w, double* out) {
  return guard([&] { W(w)->summary(out); });
}
int orc_init_constant(void* w, double rho, double th, double u, double v, double wv) {
  return guard([&] { W(w)->init_constant(rho, th, u, v, wv); });
}
int orc_init_hydrostatic_rest(void* w, double rho_s, double th) {
  return guard([&] { W(w)->init_hydro(rho_s, th); });
}
int orc_init_bubble(void* w, const double* p, int sign) {
  return guard([&] { W(w)->init_bubble(p, sign); });
}
int orc_init_mountain_wave(void* w, double rho_s, double th, double u, double v, double wv) {
  return guard([&] { W(w)->init_mountain(rho_s, th, u, v, wv); });
}

int orc_plan_view(void* w, int slice_k, int steps, double dt, int with_precip, double* out) {
  int nf = -1;
  guard([&] { nf = W(w)->plan_view(slice_k, steps, dt, with_precip, out); });
  return nf;
}

double orc_time_steps(void* w, const orc_step_cfg* c, const orc_warm_rain_cfg* wr, int moist,
                      int steps, int threads) {
  double secs = -1.0;
  guard([&] {
    if (threads > 0) omp_set_num_threads(threads);
    WorldBase* wb = W(w);
    const auto t0 = std::chrono::steady_clock::now();
    for (int s = 0; s < steps; ++s) {
      wb->ssprk3(*c, ORC_BOUNDARY_NONE, 0.0, 0.0);
      if (moist) {
        wb->tracers(*c, ORC_BOUNDARY_NONE);
        wb->warm_rain(*wr, 1);
      }
    }
    const auto t1 = std::chrono::steady_clock::now();
    secs = std::chrono::duration<double>(t1 - t0).count();
  });
  return secs;
}

}  // extern "C"
