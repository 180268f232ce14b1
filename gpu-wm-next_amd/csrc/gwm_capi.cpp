// gwm_capi.cpp -- host side of the C-ABI (include/gwm_hip.h): contexts, grids,
// device states, the SSPRK3 / fast-mode / tracer / warm-rain orchestration,
// initial states and compensated budgets.  All compute runs in the HIP
// kernels of gwm_kernels.hip; this file only allocates, launches and copies.
#include "../../include/gwm_hip.h"
#include "gwm_surrogate.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <limits>
#include <map>
#include <mutex>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "gwm_kernels.hpp"

using gwm::DMetrics;
using gwm::DState;
using gwm::Geo;

namespace {

thread_local std::string g_err;

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void require(bool c, const std::string& m) {
  if (!c) throw Error(GWM_E_INVALID, m);
}

void hipchk(hipError_t e, const char* what, const char* file, int line) {
  if (e != hipSuccess)
    throw Error(GWM_E_HIP, std::string("HIP error at ") + file + ":" + std::to_string(line) +
                               " for " + what + ": " + hipGetErrorString(e));
}
#define HIPCHK(x) hipchk((x), #x, __FILE__, __LINE__)

template <class F>
int guard(F f) {
  try {
    f();
    return GWM_OK;
  } catch (const Error& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_err = "out of memory";
    return GWM_E_NOMEM;
  } catch (const std::exception& e) {
    g_err = e.what();
    return GWM_E_INVALID;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
struct gwm_ctx {
  int device = 0;
  int prec = GWM_F32;
  hipStream_t stream = nullptr;
  bool timing = false;
  struct Stat {
    long launches = 0;
    double ms = 0.0;
  };
  struct Pending {
    std::string name;
    hipEvent_t a, b;
  };
  std::map<std::string, Stat> stats;
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  double* partials = nullptr;  // device scratch for reductions
  size_t partials_n = 0;
  // This device's copy of the static scratch of gwm_budget_sums_noalloc
  // (resolved at context creation, once per device).
  double* noalloc = nullptr;
  int refs = 0;            // live grids and states
  bool closing = false;    // gwm_ctx_destroy called while refs > 0
  // Level chunks per column of the column kernels: chosen per (kernel,
  // shape, precision) by timing the candidates at first use (autotuner), or
  // fixed by gwm_ctx_set_chunks / GWM_CHUNKS.
  int force_chunks = 0;
  std::map<std::string, int> tuned;

  size_t esize() const { return prec == GWM_F64 ? 8 : 4; }
  hipEvent_t ev() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    return e;
  }
  void resolve() {
    for (auto& p : pending) {
      HIPCHK(hipEventSynchronize(p.b));
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, p.a, p.b));
      auto& s = stats[p.name];
      s.launches += 1;
      s.ms += ms;
      pool.push_back(p.a);
      pool.push_back(p.b);
    }
    pending.clear();
  }
  double* scratch(size_t n) {
    if (n > partials_n) {
      if (partials) HIPCHK(hipFree(partials));
      HIPCHK(hipMalloc(&partials, n * sizeof(double)));
      partials_n = n;
    }
    return partials;
  }
};

namespace {

// Wraps one kernel launch with HIP events on the context stream when
// per-kernel timing is on.
struct Timed {
  gwm_ctx* c;
  hipEvent_t a = nullptr;
  const char* name;
  Timed(gwm_ctx* ctx, const char* n) : c(ctx), name(n) {
    if (c->timing) {
      a = c->ev();
      HIPCHK(hipEventRecord(a, c->stream));
    }
  }
  ~Timed() noexcept(false) {
    if (std::uncaught_exceptions() > 0) return;
    HIPCHK(hipGetLastError());
    if (c->timing) {
      hipEvent_t b = c->ev();
      HIPCHK(hipEventRecord(b, c->stream));
      c->pending.push_back({name, a, b});
    }
  }
};

void* dalloc(size_t bytes) {
  void* p = nullptr;
  if (bytes == 0) return nullptr;
  HIPCHK(hipMalloc(&p, bytes));
  return p;
}

// Host <-> device copy ordered on a context stream (non-blocking, so a
// null-stream hipMemcpy would not wait for its kernels), complete on return.
void scopy(hipStream_t st, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  HIPCHK(hipMemcpyAsync(dst, src, bytes, kind, st));
  HIPCHK(hipStreamSynchronize(st));
}

}  // namespace

// ---------------------------------------------------------------------------
// Grid / metrics
// ---------------------------------------------------------------------------
namespace {
void ctx_free(gwm_ctx* ctx);
void ctx_release(gwm_ctx* ctx) {
  if (!ctx) return;
  ctx->refs -= 1;
  if (ctx->closing && ctx->refs <= 0) ctx_free(ctx);
}
}  // namespace

struct gwm_grid {
  gwm_ctx* ctx = nullptr;
  Geo geo{};
  bool terrain_nonzero = false;
  bool terrain_active = false;  // |h| > 1e-6 somewhere (dry_fast_modes.cu:484-492)
  std::vector<double> z_if, z_c, w_if, w_c;
  void *zif_m = nullptr, *zc_m = nullptr, *idz_c = nullptr, *idz_f = nullptr;
  void *sx = nullptr, *sy = nullptr, *terr = nullptr;
  void *lev_idz_c = nullptr, *lev_idz_f = nullptr, *lev_zc = nullptr;
  ~gwm_grid() {
    for (void* p : {zif_m, zc_m, idz_c, idz_f, sx, sy, terr, lev_idz_c, lev_idz_f, lev_zc})
      if (p) (void)hipFree(p);
    ctx_release(ctx);
  }
  template <class R>
  DMetrics<R> dm() const {
    return DMetrics<R>{static_cast<const R*>(zc_m), static_cast<const R*>(idz_c),
                       static_cast<const R*>(idz_f), static_cast<const R*>(lev_idz_c),
                       static_cast<const R*>(lev_idz_f), static_cast<const R*>(lev_zc)};
  }
};

namespace {

// Terrain taper weight (grid_metrics.hpp:126-144).
double taper_weight(double eta, double taper) {
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
void build_grid(gwm_grid* G, const double* zin, const float* terrain) {
  const Geo& g = G->geo;
  const int nz = g.nz;
  (void)zin;  // the flat columns were filled by gwm_grid_create
  if (terrain) {
    for (size_t a = 0; a < static_cast<size_t>(g.nx) * g.ny; ++a) {
      if (terrain[a] != 0.0f) G->terrain_nonzero = true;
      if (std::fabs(static_cast<double>(terrain[a])) > 1.0e-6) G->terrain_active = true;
    }
  }
  const size_t n2 = static_cast<size_t>(g.nx) * g.ny;
  const size_t es = sizeof(R);
  G->zif_m = dalloc(n2 * (nz + 1) * es);
  G->zc_m = dalloc(n2 * nz * es);
  G->idz_c = dalloc(n2 * nz * es);
  G->idz_f = dalloc(n2 * (nz + 1) * es);
  G->sx = dalloc(n2 * es);
  G->sy = dalloc(n2 * es);
  G->terr = dalloc(n2 * es);
  double *d_zif = nullptr, *d_zc = nullptr, *d_wif = nullptr, *d_wc = nullptr;
  float* d_terr = nullptr;
  d_zif = static_cast<double*>(dalloc((nz + 1) * sizeof(double)));
  d_zc = static_cast<double*>(dalloc(nz * sizeof(double)));
  d_wif = static_cast<double*>(dalloc((nz + 1) * sizeof(double)));
  d_wc = static_cast<double*>(dalloc(nz * sizeof(double)));
  hipStream_t st = G->ctx->stream;
  HIPCHK(hipMemcpyAsync(d_zif, G->z_if.data(), (nz + 1) * sizeof(double), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(d_zc, G->z_c.data(), nz * sizeof(double), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(d_wif, G->w_if.data(), (nz + 1) * sizeof(double), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(d_wc, G->w_c.data(), nz * sizeof(double), hipMemcpyHostToDevice, st));
  if (terrain) {
    d_terr = static_cast<float*>(dalloc(n2 * sizeof(float)));
    HIPCHK(hipMemcpyAsync(d_terr, terrain, n2 * sizeof(float), hipMemcpyHostToDevice, st));
  }
  {
    Timed t(G->ctx, "metrics");
    gwm::launch_metrics<R>(st, g, d_zif, d_zc, d_wif, d_wc, d_terr, static_cast<R*>(G->zif_m),
                           static_cast<R*>(G->zc_m), static_cast<R*>(G->idz_c),
                           static_cast<R*>(G->idz_f), static_cast<R*>(G->sx),
                           static_cast<R*>(G->sy), static_cast<R*>(G->terr));
  }
  // Per-level columns for flat terrain, taken from column (0,0) of the
  // device build so they are identical to the 3-D fields.
  G->lev_idz_c = dalloc(nz * es);
  G->lev_idz_f = dalloc((nz + 1) * es);
  G->lev_zc = dalloc(nz * es);
  HIPCHK(hipMemcpy2DAsync(G->lev_idz_c, es, G->idz_c, n2 * es, es, nz, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpy2DAsync(G->lev_idz_f, es, G->idz_f, n2 * es, es, nz + 1, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpy2DAsync(G->lev_zc, es, G->zc_m, n2 * es, es, nz, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  for (void* p : {static_cast<void*>(d_zif), static_cast<void*>(d_zc), static_cast<void*>(d_wif),
                  static_cast<void*>(d_wc), static_cast<void*>(d_terr)})
    if (p) HIPCHK(hipFree(p));
}

}  // namespace

// ---------------------------------------------------------------------------
// State
// ---------------------------------------------------------------------------
struct gwm_state {
  gwm_ctx* ctx = nullptr;
  const gwm_grid* grid = nullptr;
  int ntr = 0;
  bool view = false;        // stage view handed to a boundary callback
  gwm_state* owner = nullptr;
  int nsets = 0;
  void* set[5][5] = {};     // buffer sets: rho, rho_theta (or theta), u, v, w
  int cur = 0;              // canonical set index
  void* tend[5] = {};
  void* pressure = nullptr;
  std::vector<void*> tr;             // canonical tracers
  std::vector<void*> tr_scr[3];      // tracer stage buffers
  void* precip = nullptr;
  void* bref[5] = {};
  std::vector<void*> bref_tr;
  bool have_boundary = false;
  unsigned tracer_strip_mask = ~0u;  // tracers the strip boundary overwrites
  void* theta_ref = nullptr;  // nz
  void* rref_lev = nullptr;   // nz
  void* pref_lev = nullptr;   // nz
  void* theta_scr = nullptr;  // cell field
  // FP64 level partial sums of the step (theta_ref, fast-mode level means):
  // owned by the state and sized once for its grid, so the step graphs
  // captured from it keep a valid address.
  double* step_part = nullptr;
  // Network-cell queue and segment counts of the surrogate warm-rain call
  // (allocated at its first call, sized for the grid).
  unsigned* sur_queue = nullptr;
  // Perturbation form (FP32; flat and doubly periodic, or terrain): rho and rho*theta are
  // stored as departures from per-level bases, the FP64 level means of the
  // state they were built from; the device holds the bases in working
  // precision (rho_b, (rho theta)_b, theta_b, p_b).
  //
  // On terrain (pert3) the base is a 3-D field: the FP64 state it was built
  // from, cell by cell (base_rho64 / base_rt64 over the cell storage); the
  // device holds its level means as above, the per-cell deviations from
  // them (b3_dev) and the base's pressure-gradient and gravity forcing per
  // face (bf_dev), built for the gravity `forcing_g`.
  bool pert = false;
  bool pert3 = false;
  std::vector<double> base_rho64, base_rt64;
  void* b_dev = nullptr;   // 4 * nz values: rho_b, rt_b, theta_b, p_b
  void* b3_dev = nullptr;  // 3 cell fields: drho, dtheta, dp
  void* bf_dev = nullptr;  // fu (x faces), fv (y faces), fw (z faces)
  double forcing_g = 9.81;
  template <class R>
  gwm::Base<R> base() const {
    if (!pert) return gwm::Base<R>{nullptr, nullptr, nullptr, nullptr};
    const gwm_state* o = view ? owner : this;
    const R* b = static_cast<const R*>(o->b_dev);
    const int nz = grid->geo.nz;
    gwm::Base<R> r{b, b + nz, b + 2 * nz, b + 3 * nz};
    if (o->pert3) {
      const Geo& g = grid->geo;
      const R* d = static_cast<const R*>(o->b3_dev);
      const R* f = static_cast<const R*>(o->bf_dev);
      r.drho = d;
      r.dth = d + g.cell_total();
      r.dp = d + 2 * g.cell_total();
      r.fu = f;
      r.fv = f + g.xf_total();
      r.fw = f + g.xf_total() + g.yf_total();
    }
    return r;
  }
  // index of a cell storage position in the FP64 bases
  size_t bidx(size_t a) const {
    const Geo& g = grid->geo;
    return pert3 ? a : a / (static_cast<size_t>(g.pxc) * g.pyc);
  }
  // Captured two-step moist-step graphs, one per buffer parity (the step
  // rotates the buffer sets and tracer buffers with period two).
  struct Graph {
    hipGraphExec_t exec = nullptr;
    int cur = -1;
    void* tr0 = nullptr;
    std::vector<double> key;
    std::vector<const void*> bufs;  // device buffers at capture (buffers())
  };
  // Every device buffer of the state, in a fixed order.  A captured step
  // graph bakes these addresses in; a replay requires each buffer that
  // existed at capture to be the same allocation still.
  std::vector<const void*> buffers() const {
    std::vector<const void*> b;
    for (const auto& st : set)
      for (const void* q : st) b.push_back(q);
    for (const void* q : tend) b.push_back(q);
    for (const void* q : bref) b.push_back(q);
    for (const auto* v : {&tr, &tr_scr[0], &tr_scr[1], &tr_scr[2], &bref_tr}) {
      b.push_back(reinterpret_cast<const void*>(v->size()));
      for (const void* q : *v) b.push_back(q);
    }
    for (const void* q : {pressure, precip, theta_ref, rref_lev, pref_lev, theta_scr, b_dev, b3_dev,
                          bf_dev, static_cast<void*>(step_part), static_cast<void*>(sur_queue)})
      b.push_back(q);
    return b;
  }
  Graph graphs[2];
  // configuration (graph_key) of the last eager warm-up step of a batched
  // ensemble: every scratch buffer allocated and every kernel tuned for it
  std::vector<double> batch_key;
  bool faces_synced = false;
  // theta_ref partials of the state (rho_set) that a subcycle's last substep
  // wrote, ready for the next stage's theta_ref (no pass over the state)
  const void* theta_parts_rho = nullptr;
  int theta_nparts = 0;
  // view of a stage set; a tracer stage view also points at the stage's
  // tracer buffers (empty: the owner's canonical tracers)
  void* vset[5] = {};
  std::vector<void*> vtr;

  const Geo& g() const { return grid->geo; }
  size_t es() const { return ctx->esize(); }
  size_t fsize(int f) const {
    const Geo& G = g();
    switch (f) {
      case GWM_MOM_U:
      case GWM_TEND_U:
        return G.xf_total();
      case GWM_MOM_V:
      case GWM_TEND_V:
        return G.yf_total();
      case GWM_MOM_W:
      case GWM_TEND_W:
        return G.zf_total();
      case GWM_PRECIP:
        return static_cast<size_t>(G.nx) * G.ny;
      default:
        return G.cell_total();
    }
  }
  void* alloc_set_field(int f) { return dalloc(fsize(f) * es()); }
  void ensure_sets(int n) {
    for (int s = nsets; s < n; ++s)
      for (int f = 0; f < 5; ++f) set[s][f] = alloc_set_field(f);
    nsets = std::max(nsets, n);
  }
  void** cur_set() { return view ? vset : set[cur]; }
  void* field(int f) {
    if (f >= GWM_RHO && f <= GWM_MOM_W) return cur_set()[f];
    if (f == GWM_PRESSURE) {
      if (!pressure) pressure = dalloc(g().cell_total() * es());
      return pressure;
    }
    if (f >= GWM_TEND_RHO && f <= GWM_TEND_W) {
      const int a = f - GWM_TEND_RHO;
      if (!tend[a]) tend[a] = dalloc(fsize(f) * es());
      return tend[a];
    }
    if (f == GWM_PRECIP) return precip;
    const int t = f - GWM_TRACER0;
    require(t >= 0 && t < ntr, "bad field id");
    if (view) return vtr.empty() ? owner->tr[t] : vtr[t];
    return tr[t];
  }
  ~gwm_state() {
    if (view) return;
    for (auto& gr : graphs)
      if (gr.exec) (void)hipGraphExecDestroy(gr.exec);
    struct Release {
      gwm_ctx* c;
      ~Release() { ctx_release(c); }
    } release{ctx};
    for (auto& s : set)
      for (void* p : s)
        if (p) (void)hipFree(p);
    for (void* p : tend)
      if (p) (void)hipFree(p);
    for (void* p : bref)
      if (p) (void)hipFree(p);
    for (auto* v : {&tr, &tr_scr[0], &tr_scr[1], &tr_scr[2], &bref_tr})
      for (void* p : *v)
        if (p) (void)hipFree(p);
    for (void* p : {pressure, precip, theta_ref, rref_lev, pref_lev, theta_scr, b_dev, b3_dev, bf_dev,
                    static_cast<void*>(step_part), static_cast<void*>(sur_queue)})
      if (p) (void)hipFree(p);
  }
  unsigned* surrogate_queue() {
    if (!sur_queue)
      sur_queue = static_cast<unsigned*>(dalloc(gwm::sur_queue_words(g()) * sizeof(unsigned)));
    return sur_queue;
  }
  // Partial-sum buffer of the step's level means (both uses fit).
  double* level_part() {
    if (!step_part) {
      const Geo& gg = g();
      const size_t per = std::max<size_t>(gwm::fast_column_parts(gg.nx, gg.ny),
                                          gwm::level_partials_chunks(gg.ny));
      step_part = static_cast<double*>(dalloc(static_cast<size_t>(gg.nz) * per * 2 * sizeof(double)));
    }
    return step_part;
  }
};

namespace {

// Uploads the perturbation form's bases (defined with the conversions below).
void push_base(gwm_state* S);

gwm_state* S_(gwm_state* s) {
  require(s != nullptr, "null state");
  return s;
}

template <class R>
DState<R> dstate(void* const* f) {
  return DState<R>{static_cast<R*>(f[0]), static_cast<R*>(f[1]), static_cast<R*>(f[2]),
                   static_cast<R*>(f[3]), static_cast<R*>(f[4])};
}

// Wrap-face synchronisation of the canonical faces (makes the reference's
// owned-interface averaging an identity inside the step).
template <class R>
void ensure_synced(gwm_state* S, void** fs) {
  if (S->faces_synced) return;
  const Geo& g = S->g();
  if (g.per_x) {
    Timed t(S->ctx, "sync_interfaces");
    gwm::launch_sync_interfaces<R>(S->ctx->stream, g, 1, static_cast<R*>(fs[GWM_MOM_U]));
  }
  if (g.per_y) {
    Timed t(S->ctx, "sync_interfaces");
    gwm::launch_sync_interfaces<R>(S->ctx->stream, g, 2, static_cast<R*>(fs[GWM_MOM_V]));
  }
  S->faces_synced = true;
}

template <class R>
void level_theta_ref(gwm_state* S, const R* rho, const R* rt) {
  const Geo& g = S->g();
  if (!S->theta_ref) S->theta_ref = dalloc(g.nz * sizeof(R));
  double* part = S->level_part();
  int nch = 0;
  if (S->theta_parts_rho == rho) {  // emitted by the subcycle that wrote (rho, rt)
    nch = S->theta_nparts;
  } else {
    Timed t(S->ctx, "level_partials");
    gwm::launch_level_partials<R>(S->ctx->stream, g, rho, rt, 0, part, &nch, S->base<R>());
  }
  S->theta_parts_rho = nullptr;
  {
    Timed t(S->ctx, "level_finalize");
    gwm::launch_level_finalize<R>(S->ctx->stream, g.nz, nch, part, 0,
                                  static_cast<double>(g.nx) * g.ny,
                                  static_cast<R*>(S->theta_ref), nullptr);
  }
}

// Autotuner for the launch shape of a column kernel: the level-chunk count
// and, where the kernel has more than one (`shapes`), the tile height.
// `launch(code)` runs the kernel with code = chunks | 16 * shape; the kernels
// are out-of-place and their results do not depend on the shape, so the
// candidates run on the live arguments and every candidate writes identical
// values.  Each candidate is timed over two launches after a warm-up.
// `fine` adds the chunk counts 3, 6 and 12 (the slow and tracer stages,
// whose resident work-groups per CU make the round count of a launch
// matter); the fast substep keeps 1, 2, 4, 8, where the extra counts won
// back-to-back pairs while timing slower launch by launch.
template <class Launch>
int pick_chunks(gwm_ctx* ctx, const char* kname, const Geo& g, int variant, Launch&& launch,
                int shapes = 1, bool whole_columns = false, bool fine = false) {
  if (ctx->force_chunks > 0) return ctx->force_chunks;
  char key[160];
  std::snprintf(key, sizeof key, "%s:%dx%dx%d:%s:%d", kname, g.nx, g.ny, g.nz,
                ctx->prec == GWM_F64 ? "f64" : "f32", variant);
  auto it = ctx->tuned.find(key);
  if (it != ctx->tuned.end()) return it->second;
  // a recorded step launches nothing, so it cannot time candidates
  if (gwm::refuse_record("an untuned column kernel")) return 1;
  hipEvent_t e0 = ctx->ev(), e1 = ctx->ev();
  // Three rounds over the candidates, each timing two launches after a warm
  // one; a candidate's time is its fastest round, so clock and cache drift
  // between rounds does not decide the choice.
  std::vector<int> cands;
  for (int sh = 0; sh < shapes; ++sh)
    for (int n : {1, 2, 3, 4, 6, 8, 12})
      if (n == 1 || (!whole_columns && g.nz / n >= 4 && (fine || (n & (n - 1)) == 0)))
        cands.push_back(n | (sh << 4));
  std::vector<float> t_min(cands.size(), 1e30f);
  for (int round = 0; round < 3; ++round)
    for (size_t c = 0; c < cands.size(); ++c) {
      launch(cands[c]);
      HIPCHK(hipEventRecord(e0, ctx->stream));
      launch(cands[c]);
      launch(cands[c]);
      HIPCHK(hipEventRecord(e1, ctx->stream));
      HIPCHK(hipEventSynchronize(e1));
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, e0, e1));
      t_min[c] = std::min(t_min[c], ms);
    }
  size_t best = 0;
  for (size_t c = 1; c < cands.size(); ++c)
    if (t_min[c] < t_min[best]) best = c;
  ctx->pool.push_back(e0);
  ctx->pool.push_back(e1);
  ctx->tuned[key] = cands[best];
  return cands[best];
}

// One fast subcycle (apply_local_split_explicit_fast_modes) from rho_in and
// the entry faces (ub, vb, wb), which are also the base faces of K15, with
// fixed theta; returns the set index holding the result (rho, rho*theta,
// faces).  Between substeps the out sets z0 / z1 hold face departures from
// the entry faces.
template <class R>
int subcycle(gwm_state* S, const gwm_step_config& c, const R* rho_in, const R* theta,
             const R* ub, const R* vb, const R* wb, int z0, int z1, bool theta_parts = false) {
  const Geo& g = S->g();
  const gwm_grid* G = S->grid;
  const int ns = std::max(1, c.fast_substeps);
  const R dtf = static_cast<R>(c.dt) / static_cast<R>(ns);
  // Terrain runs use the column reference (K11), other runs the level
  // references; the perturbation form on terrain neither (K14 on the
  // departures plus the base forcing, built for the step's gravity).
  const bool tp = S->pert3;
  const bool colref = G->terrain_active && !tp;
  const bool flat = !G->terrain_active;
  if (tp && S->forcing_g != c.gravity) {
    S->forcing_g = c.gravity;
    push_base(S);
  }
  if (flat) {
    if (!S->rref_lev) S->rref_lev = dalloc(g.nz * sizeof(R));
    if (!S->pref_lev) S->pref_lev = dalloc(g.nz * sizeof(R));
  }
  const int outs[2] = {z0, z1};
  const R* rho = rho_in;
  const R *u = ub, *v = vb, *w = wb;  // departures after the first substep
  int last = z0;
  // Level partials: substep 0 from a pass over the entry state, later
  // substeps from the previous substep's column kernel.
  double* part = flat ? S->level_part() : nullptr;
  int nparts = 0;
  for (int s = 0; s < ns; ++s) {
    const int o = outs[s % 2];
    if (flat) {
      if (s == 0) {
        Timed t(S->ctx, "level_partials");
        gwm::launch_level_partials<R>(S->ctx->stream, g, rho, theta, 1, part, &nparts,
                                      S->base<R>());
      }
      Timed t(S->ctx, "level_finalize");
      gwm::launch_level_finalize<R>(S->ctx->stream, g.nz, nparts, part, 1,
                                    static_cast<double>(g.nx) * g.ny,
                                    static_cast<R*>(S->rref_lev), static_cast<R*>(S->pref_lev));
    }
    gwm::FastArgs<R> a{};
    a.g = g;
    a.rho = rho;
    a.theta = theta;
    a.u = u;
    a.v = v;
    a.w = w;
    a.ub = ub;
    a.vb = vb;
    a.wb = wb;
    a.rho_out = static_cast<R*>(S->set[o][0]);
    a.rt_out = s == ns - 1 ? static_cast<R*>(S->set[o][1]) : nullptr;
    a.u_out = static_cast<R*>(S->set[o][2]);
    a.v_out = static_cast<R*>(S->set[o][3]);
    a.w_out = static_cast<R*>(S->set[o][4]);
    a.rref_lev = static_cast<const R*>(S->rref_lev);
    a.pref_lev = static_cast<const R*>(S->pref_lev);
    a.m = G->dm<R>();
    a.dtf = dtf;
    a.gravity = static_cast<R>(c.gravity);
    a.horizontal = c.fast_update_horizontal_momentum;
    a.density = c.fast_update_density;
    a.col_ref = colref;
    a.first = s == 0;
    a.last = s == ns - 1;
    // Every other substep marches its columns from the top down, so it
    // reads first the levels the substep before wrote last (still in the
    // 256 MB last-level cache); the last substep marches down, so the
    // upward-marching stage kernels after it do too.  Results do not change.
    a.down = flat && ((ns - 1 - s) % 2 == 0);
    a.base = S->base<R>();
    // level sums for the next substep, or (last substep, on request) the
    // theta_ref sums of the written state for the next stage
    const bool emit_th = flat && theta_parts && s + 1 == ns;
    const bool emit = (flat && s + 1 < ns) || emit_th;
    a.part_theta = emit_th ? 1 : 0;
    auto run = [&](int n) {
      gwm::launch_fast_column<R>(S->ctx->stream, a, G->terrain_nonzero, n,
                                 emit ? part : nullptr, emit ? &nparts : nullptr);
    };
    // flat launches choose level chunks and the tile height (8 or 4 rows);
    // column-reference launches keep whole columns and choose the height
    const int nkc =
        colref ? pick_chunks(S->ctx, "fast_column", g, 2, run, 2, /*whole_columns=*/true)
               : pick_chunks(S->ctx, "fast_column", g, G->terrain_nonzero ? (tp ? 3 : 1) : 0, run,
                             G->terrain_nonzero ? 1 : 2);
    Timed t(S->ctx, "fast_column");
    run(nkc);
    rho = a.rho_out;
    u = a.u_out;
    v = a.v_out;
    w = a.w_out;
    last = o;
    if (emit_th) {
      S->theta_parts_rho = a.rho_out;
      S->theta_nparts = nparts;
    }
  }
  return last;
}

template <class R>
void slow_stage(gwm_state* S, const gwm_step_config& c, void* const* in, void* const* q0,
                void* const* out, int mode, R wa, R wb, int out_theta) {
  const Geo& g = S->g();
  level_theta_ref<R>(S, static_cast<const R*>(in[0]), static_cast<const R*>(in[1]));
  gwm::SlowArgs<R> a{};
  a.g = g;
  a.s = dstate<R>(in);
  if (q0) a.q0 = dstate<R>(q0);
  a.out = dstate<R>(out);
  a.m = S->grid->dm<R>();
  a.theta_ref = static_cast<const R*>(S->theta_ref);
  a.gravity = static_cast<R>(c.gravity);
  a.dt = static_cast<R>(c.dt);
  a.wa = wa;
  a.wb = wb;
  a.mode = mode;
  a.out_theta = out_theta;
  a.base = S->base<R>();
  auto run = [&](int n) {
    gwm::launch_slow_column<R>(S->ctx->stream, a, S->grid->terrain_nonzero, n);
  };
  const int nkc = pick_chunks(S->ctx, "slow_column", g,
                              (S->grid->terrain_nonzero ? 1 : 0) + 2 * mode + (S->pert3 ? 16 : 0),
                              run, 1, false, true);
  Timed t(S->ctx, "slow_stage");
  run(nkc);
}

int orient_of(int f) {
  switch (f) {
    case GWM_MOM_U:
    case GWM_TEND_U:
      return 1;
    case GWM_MOM_V:
    case GWM_TEND_V:
      return 2;
    case GWM_MOM_W:
    case GWM_TEND_W:
      return 3;
    default:
      return 0;
  }
}

template <class R>
void apply_boundary_set(gwm_state* S, void* const* fs, int mode, double width) {
  if (mode == GWM_BOUNDARY_NONE) return;
  require(S->have_boundary, "boundary reference not captured (gwm_state_capture_boundary)");
  const Geo& g = S->g();
  for (int f = 0; f < 5; ++f) {
    Timed t(S->ctx, "boundary");
    gwm::launch_boundary<R>(S->ctx->stream, g, orient_of(f), static_cast<R*>(fs[f]),
                            static_cast<const R*>(S->bref[f]), mode, width);
  }
}

// The stage's lateral boundary: first the host callback (which may refresh
// the boundary reference with gwm_state_set_boundary, or edit the stage),
// then `mode` against the boundary reference.
template <class R>
void run_boundary(gwm_state* S, void** fs, const gwm_boundary_config* b, double sim_time) {
  if (!b) return;
  if (b->callback) {
    gwm_state view;
    view.ctx = S->ctx;
    view.grid = S->grid;
    view.ntr = S->ntr;
    view.view = true;
    view.owner = S;
    view.pert = S->pert;  // reads and writes convert with the owner's bases
    view.pert3 = S->pert3;
    view.faces_synced = true;
    for (int f = 0; f < 5; ++f) view.vset[f] = fs[f];
    HIPCHK(hipStreamSynchronize(S->ctx->stream));
    const int rc = b->callback(&view, sim_time, b->user);
    require(rc == 0, "boundary callback failed");
  }
  apply_boundary_set<R>(S, fs, b->mode, b->relax_width);
}

// advance_dry_state_ssprk3 (dry_core.cu:541-651).  fast_fn != null replaces
// the device fast modes with a host plug-in run on each stage state.
template <class R>
void ssprk3(gwm_state* S, const gwm_step_config& c, const gwm_boundary_config* b,
            gwm_boundary_fn fast_fn = nullptr, void* fast_user = nullptr) {
  const bool bnd = b && (b->mode != GWM_BOUNDARY_NONE || b->callback);
  S->ensure_sets(bnd ? 5 : 4);
  S->theta_parts_rho = nullptr;
  const int X = S->cur;
  int others[4], n = 0;
  for (int s = 0; s < S->nsets && n < 4; ++s)
    if (s != X) others[n++] = s;
  const int Y = others[0], Z0 = others[1], Z1 = others[2];
  ensure_synced<R>(S, S->set[X]);
  void** in = S->set[X];
  if (bnd) {
    const int Wb = others[3];
    for (int f = 0; f < 5; ++f) {
      Timed t(S->ctx, "copy");
      gwm::launch_copy<R>(S->ctx->stream, static_cast<const R*>(S->set[X][f]),
                          static_cast<R*>(S->set[Wb][f]), S->fsize(f));
    }
    in = S->set[Wb];
    run_boundary<R>(S, in, b, 0.0);
  }
  // theta_parts: the subcycle's last substep also emits the next stage's
  // theta_ref sums; only without boundaries, which change the state between.
  auto sub = [&](int y, bool theta_parts) {
    void** ys = S->set[y];
    if (fast_fn) {
      // Host plug-in: hand it the stage state (rho, rho*theta, faces) in Z0.
      void** zs = S->set[Z0];
      const Geo& g = S->g();
      for (int f = 0; f < 5; ++f) {
        Timed t(S->ctx, "copy");
        gwm::launch_copy<R>(S->ctx->stream, static_cast<const R*>(ys[f]), static_cast<R*>(zs[f]),
                            S->fsize(f));
      }
      {
        // ys[1] holds theta; rebuild rho*theta = rho*theta (K10) for the view.
        Timed t(S->ctx, "theta");
        gwm::launch_rebuild_rt<R>(S->ctx->stream, g, static_cast<const R*>(ys[0]),
                                  static_cast<const R*>(ys[1]), static_cast<R*>(zs[1]),
                                  S->base<R>());
      }
      gwm_state view;
      view.ctx = S->ctx;
      view.grid = S->grid;
      view.ntr = S->ntr;
      view.view = true;
      view.owner = S;
      view.pert = S->pert;
      view.pert3 = S->pert3;
      view.faces_synced = true;
      for (int f = 0; f < 5; ++f) view.vset[f] = zs[f];
      HIPCHK(hipStreamSynchronize(S->ctx->stream));
      require(fast_fn(&view, 0.0, fast_user) == 0, "fast-mode callback failed");
      return Z0;
    }
    return subcycle<R>(S, c, static_cast<const R*>(ys[0]), static_cast<const R*>(ys[1]),
                       static_cast<const R*>(ys[2]), static_cast<const R*>(ys[3]),
                       static_cast<const R*>(ys[4]), Z0, Z1, theta_parts);
  };
  // stage 1
  slow_stage<R>(S, c, in, nullptr, S->set[Y], 1, R(0), R(1), 1);
  int r = sub(Y, !bnd);
  run_boundary<R>(S, S->set[r], b, c.dt);
  // stage 2
  slow_stage<R>(S, c, S->set[r], S->set[X], S->set[Y], 2, R(0.75), R(0.25), 1);
  r = sub(Y, !bnd);
  run_boundary<R>(S, S->set[r], b, 0.5 * c.dt);
  // stage 3
  slow_stage<R>(S, c, S->set[r], S->set[X], S->set[Y], 2, R(1) / R(3), R(2) / R(3), 1);
  r = sub(Y, false);
  S->cur = r;
}

// apply_local_split_explicit_fast_modes on the canonical state.
template <class R>
void fast_modes(gwm_state* S, const gwm_step_config& c) {
  S->ensure_sets(4);
  const int X = S->cur;
  int others[3], n = 0;
  for (int s = 0; s < S->nsets && n < 3; ++s)
    if (s != X) others[n++] = s;
  ensure_synced<R>(S, S->set[X]);
  const Geo& g = S->g();
  if (!S->theta_scr) S->theta_scr = dalloc(g.cell_total() * sizeof(R));
  {
    Timed t(S->ctx, "theta");
    gwm::launch_theta<R>(S->ctx->stream, g, static_cast<const R*>(S->set[X][0]),
                         static_cast<const R*>(S->set[X][1]), static_cast<R*>(S->theta_scr),
                         S->base<R>());
  }
  void** xs = S->set[X];
  const int r = subcycle<R>(S, c, static_cast<const R*>(xs[0]),
                            static_cast<const R*>(S->theta_scr), static_cast<const R*>(xs[2]),
                            static_cast<const R*>(xs[3]), static_cast<const R*>(xs[4]),
                            others[0], others[1]);
  S->cur = r;
}

template <class R>
void slow_tendencies(gwm_state* S, const gwm_step_config& c) {
  ensure_synced<R>(S, S->cur_set());
  void* out[5];
  for (int f = 0; f < 5; ++f) out[f] = S->field(GWM_TEND_RHO + f);
  slow_stage<R>(S, c, S->cur_set(), nullptr, out, 0, R(0), R(1), 0);
}

// advance_passive_tracers_ssprk3 (dry_transport.cu:440-518).  The lateral
// boundary runs on each stage input (strips from the captured reference,
// then the host callback `fn` on a stage view at the step-relative times 0,
// dt, dt/2, as apply_tracer_boundaries_to_fields, dry_transport.cu:254-276).
template <class R>
void tracers_ssprk3(gwm_state* S, const gwm_step_config& c, int bmode,
                    gwm_boundary_fn fn = nullptr, void* user = nullptr) {
  if (S->ntr == 0) return;
  const Geo& g = S->g();
  ensure_synced<R>(S, S->set[S->cur]);
  for (int k = 0; k < 3; ++k)
    if (S->tr_scr[k].empty())
      for (int t = 0; t < S->ntr; ++t) S->tr_scr[k].push_back(dalloc(g.cell_total() * sizeof(R)));
  void** ds = S->set[S->cur];
  gwm::TracerArgs<R> a{};
  a.g = g;
  a.rho = static_cast<const R*>(ds[0]);
  a.u = static_cast<const R*>(ds[2]);
  a.v = static_cast<const R*>(ds[3]);
  a.w = static_cast<const R*>(ds[4]);
  a.idz_c = static_cast<const R*>(S->grid->idz_c);
  a.lev_idz_c = static_cast<const R*>(S->grid->lev_idz_c);
  a.ntr = S->ntr;
  a.dt = static_cast<R>(c.dt);
  // callback first (it may refresh the tracer boundary reference), then
  // the strips of the tracers in the strip mask
  auto strips = [&](std::vector<void*>& f, double sim_time) {
    if (fn) {
      gwm_state view;
      view.ctx = S->ctx;
      view.grid = S->grid;
      view.ntr = S->ntr;
      view.view = true;
      view.owner = S;
      view.pert = S->pert;
      view.pert3 = S->pert3;
      view.faces_synced = true;
      for (int q = 0; q < 5; ++q) view.vset[q] = ds[q];
      view.vtr = f;
      HIPCHK(hipStreamSynchronize(S->ctx->stream));
      require(fn(&view, sim_time, user) == 0, "tracer boundary callback failed");
    }
    if (bmode != GWM_BOUNDARY_NONE) {
      require(S->have_boundary, "boundary reference not captured");
      for (int t = 0; t < S->ntr; ++t) {
        if (t < 32 && !((S->tracer_strip_mask >> t) & 1u)) continue;
        Timed tt(S->ctx, "boundary");
        gwm::launch_boundary<R>(S->ctx->stream, g, 0, static_cast<R*>(f[t]),
                                static_cast<const R*>(S->bref_tr[t]), GWM_BOUNDARY_STRIPS, 0.0);
      }
    }
  };
  const bool bnd = bmode != GWM_BOUNDARY_NONE || fn;
  std::vector<void*>& A = S->tr_scr[0];
  std::vector<void*>& B = S->tr_scr[1];
  std::vector<void*>& C = S->tr_scr[2];
  std::vector<void*> in1 = S->tr;
  if (bnd) {
    for (int t = 0; t < S->ntr; ++t) {
      Timed tt(S->ctx, "copy");
      gwm::launch_copy<R>(S->ctx->stream, static_cast<const R*>(S->tr[t]), static_cast<R*>(C[t]),
                          g.cell_total());
    }
    strips(C, 0.0);
    in1 = C;
  }
  auto stage = [&](const std::vector<void*>& in, std::vector<void*>& out, int mode, R wa, R wb) {
    for (int t = 0; t < S->ntr; ++t) {
      a.in[t] = static_cast<const R*>(in[t]);
      a.q0[t] = static_cast<const R*>(S->tr[t]);
      a.out[t] = static_cast<R*>(out[t]);
    }
    a.mode = mode;
    a.wa = wa;
    a.wb = wb;
    a.base_rho = S->base<R>().rho;
    auto run = [&](int n) {
      gwm::launch_tracer_column<R>(S->ctx->stream, a, S->grid->terrain_nonzero, n);
    };
    const int nkc = pick_chunks(S->ctx, "tracer_column", g,
                                (S->grid->terrain_nonzero ? 1 : 0) + 2 * S->ntr, run, 1, false,
                                true);
    Timed tt(S->ctx, "tracer_stage");
    run(nkc);
  };
  stage(in1, A, 1, R(0), R(1));
  if (bnd) strips(A, c.dt);
  stage(A, B, 2, R(0.75), R(0.25));
  if (bnd) strips(B, 0.5 * c.dt);
  stage(B, A, 2, R(1) / R(3), R(2) / R(3));
  std::swap(S->tr, S->tr_scr[0]);
}

template <class R>
void warm_rain(gwm_state* S, const gwm_warm_rain_config& c, int metric) {
  require(S->ntr >= 3, "warm rain requires the qv, qc, qr tracers");
  const Geo& g = S->g();
  void** ds = S->cur_set();
  gwm::WarmRainArgs<R> a{};
  a.g = g;
  a.rho = static_cast<const R*>(ds[0]);
  a.rt = static_cast<R*>(ds[1]);
  a.qv = static_cast<R*>(S->field(GWM_TRACER0 + 0));
  a.qc = static_cast<R*>(S->field(GWM_TRACER0 + 1));
  a.qr = static_cast<R*>(S->field(GWM_TRACER0 + 2));
  a.precip = static_cast<R*>(S->view ? S->owner->precip : S->precip);
  a.idz_c = static_cast<const R*>(S->grid->idz_c);
  a.lev_idz_c = static_cast<const R*>(S->grid->lev_idz_c);
  a.dt = static_cast<R>(c.dt);
  a.relax = static_cast<R>(c.condensation_relaxation);
  a.thr = static_cast<R>(c.cloud_autoconversion_threshold);
  a.krate = static_cast<R>(c.cloud_autoconversion_rate);
  a.erate = static_cast<R>(c.rain_evaporation_rate);
  a.vt = static_cast<R>(c.rain_terminal_velocity);
  a.depth = static_cast<R>(c.sedimentation_layer_depth_m);
  a.kacc = static_cast<R>(c.accretion_rate);
  a.heat = c.enable_latent_heating;
  if (metric) a.fallout = a.vt > R(0) ? 2 : 0;
  else a.fallout = (a.vt > R(0) && a.depth > R(0)) ? 1 : 0;
  a.base = S->base<R>();
  Timed t(S->ctx, "warm_rain");
  gwm::launch_warm_rain<R>(S->ctx->stream, a, S->grid->terrain_nonzero);
}

// Where a reduction keeps its block partials: by default the context's
// scratch (grown on demand) and a host vector; a caller may pass both.
struct ReduceBufs {
  double* dev = nullptr;   // nb * kReduceParts doubles, or null: context scratch
  double* host = nullptr;  // the same count, or null: a host vector
};

constexpr int kReduceBlocks = 512;
static_assert(kReduceBlocks * gwm::kReduceParts <= gwm::kNoAllocScratchDoubles,
              "the budget partials must fit the static scratch");

// Compensated reduction of one field; returns {sum, min, max, non-finite
// count} (sum and extremes over the finite values).
template <class R>
void reduce_field(gwm_state* S, const R* f, int fx, int fy, int fz, int kind, const R* rho,
                  double out[4], const R* fbase = nullptr, const R* rbase = nullptr,
                  const ReduceBufs& rb = ReduceBufs{}, int halo = -1) {
  const int nb = kReduceBlocks;
  const int np = gwm::kReduceParts;
  double* part = rb.dev ? rb.dev : S->ctx->scratch(static_cast<size_t>(nb) * np);
  gwm::ReduceArgs<R> a{};
  a.f = f;
  a.rho = rho;
  a.fbase = fbase;
  a.rbase = rbase;
  a.cb = S->base<R>();
  a.idz3 = S->grid->terrain_nonzero ? static_cast<const R*>(S->grid->idz_c) : nullptr;
  a.idz_lev = static_cast<const R*>(S->grid->lev_idz_c);
  a.fx = fx;
  a.fy = fy;
  a.fz = fz;
  a.h = halo >= 0 ? halo : S->g().h;
  a.kind = kind;
  a.partials = part;
  if (rb.dev) {
    gwm::launch_reduce<R>(S->ctx->stream, a, nb);
    HIPCHK(hipGetLastError());
  } else {
    Timed t(S->ctx, "reduce");
    gwm::launch_reduce<R>(S->ctx->stream, a, nb);
  }
  std::vector<double> hv;
  double* h = rb.host;
  if (!h) {
    hv.resize(static_cast<size_t>(nb) * np);
    h = hv.data();
  }
  HIPCHK(hipMemcpyAsync(h, part, static_cast<size_t>(nb) * np * sizeof(double),
                        hipMemcpyDeviceToHost, S->ctx->stream));
  HIPCHK(hipStreamSynchronize(S->ctx->stream));
  double sum = 0.0, comp = 0.0, mn = 1e300, mx = -1e300, bad = 0.0;
  for (int b = 0; b < nb; ++b) {  // Neumaier over block partials, fixed order
    for (int q = 0; q < 2; ++q) {
      const double v = h[b * np + q];
      const double s2 = sum + v;
      comp += std::fabs(sum) >= std::fabs(v) ? (sum - s2) + v : (v - s2) + sum;
      sum = s2;
    }
    mn = std::min(mn, h[b * np + 2]);
    mx = std::max(mx, h[b * np + 3]);
    bad += h[b * np + 4];
  }
  out[0] = sum + comp;
  out[1] = mn;
  out[2] = mx;
  out[3] = bad;
}

// rb.dev set: the allocation-free form (static scratch, no host vectors;
// the precipitation bucket is reduced on the device too).
template <class R>
void budget(gwm_state* S, double* o, const ReduceBufs& rb = ReduceBufs{}) {
  const Geo& g = S->g();
  for (int a = 0; a < GWM_SUMMARY_N; ++a) o[a] = 0.0;
  void** ds = S->cur_set();
  const R* rho = static_cast<const R*>(ds[0]);
  const gwm::Base<R> B = S->base<R>();
  double r[4];
  // Non-finite values over every prognostic field (the extremes and sums
  // above are over the finite values only): a field with any NaN or
  // infinity reports NaN extremes, so no check of them passes by accident.
  double nonfinite = 0.0;
  auto extremes = [&](int lo, int hi) {
    nonfinite += r[3];
    if (r[3] > 0.0) {
      o[lo] = std::numeric_limits<double>::quiet_NaN();
      o[hi] = std::numeric_limits<double>::quiet_NaN();
    } else {
      o[lo] = r[1];
      o[hi] = r[2];
    }
  };
  // Perturbation form: sums of the departures plus the FP64 base sums, so
  // the mass budget keeps the departures' precision.
  double base_rho = 0.0, base_rt = 0.0;
  const gwm_state* bo = S->view ? S->owner : S;
  if (bo->pert3) {
    for (int k = 0; k < g.nz; ++k)
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) {
          base_rho += bo->base_rho64[g.cell(i, j, k)];
          base_rt += bo->base_rt64[g.cell(i, j, k)];
        }
  } else if (S->pert) {
    for (int k = 0; k < g.nz; ++k) {
      base_rho += static_cast<double>(g.nx) * g.ny * bo->base_rho64[static_cast<size_t>(k)];
      base_rt += static_cast<double>(g.nx) * g.ny * bo->base_rt64[static_cast<size_t>(k)];
    }
  }
  reduce_field<R>(S, rho, g.nx, g.ny, g.nz, 0, nullptr, r, B.rho, nullptr, rb);
  o[GWM_S_TOTAL_DRY_MASS] = r[0] + base_rho;
  extremes(GWM_S_RHO_MIN, GWM_S_RHO_MAX);
  reduce_field<R>(S, static_cast<const R*>(ds[1]), g.nx, g.ny, g.nz, 0, nullptr, r, B.rt,
                  nullptr, rb);
  o[GWM_S_TOTAL_RHO_THETA] = r[0] + base_rt;
  nonfinite += r[3];
  reduce_field<R>(S, static_cast<const R*>(ds[1]), g.nx, g.ny, g.nz, 2, rho, r, B.rt, B.rho,
                  rb);
  const double rt_bad = nonfinite;
  extremes(GWM_S_THETA_MIN, GWM_S_THETA_MAX);
  nonfinite = rt_bad;  // theta's non-finite cells are rho's and rho*theta's
  reduce_field<R>(S, static_cast<const R*>(ds[2]), g.nx + 1, g.ny, g.nz, 0, nullptr, r, nullptr,
                  nullptr, rb);
  o[GWM_S_TOTAL_MOM_X] = r[0];
  extremes(GWM_S_U_MIN, GWM_S_U_MAX);
  reduce_field<R>(S, static_cast<const R*>(ds[3]), g.nx, g.ny + 1, g.nz, 0, nullptr, r, nullptr,
                  nullptr, rb);
  o[GWM_S_TOTAL_MOM_Y] = r[0];
  extremes(GWM_S_V_MIN, GWM_S_V_MAX);
  reduce_field<R>(S, static_cast<const R*>(ds[4]), g.nx, g.ny, g.nz + 1, 0, nullptr, r, nullptr,
                  nullptr, rb);
  o[GWM_S_TOTAL_MOM_Z] = r[0];
  extremes(GWM_S_W_MIN, GWM_S_W_MAX);
  for (int t = 0; t < std::min(S->ntr, 3); ++t) {
    const R* q = static_cast<const R*>(S->field(GWM_TRACER0 + t));
    reduce_field<R>(S, q, g.nx, g.ny, g.nz, 1, nullptr, r, nullptr, nullptr, rb);
    o[GWM_S_VAPOR_PATH + t] = r[0];
    nonfinite += r[3];
    const double q_bad = nonfinite;
    reduce_field<R>(S, q, g.nx, g.ny, g.nz, 2, rho, r, nullptr, B.rho, rb);
    extremes(GWM_S_TRACER_QMIN0 + 2 * t, GWM_S_TRACER_QMIN0 + 2 * t + 1);
    nonfinite = q_bad;
  }
  const size_t n2 = static_cast<size_t>(g.nx) * g.ny;
  const R* bucket = static_cast<const R*>(S->view ? S->owner->precip : S->precip);
  if (rb.dev) {
    // the bucket (ny, nx, no halo) reduced on the device: sum and maximum,
    // then the count of precipitating columns
    reduce_field<R>(S, bucket, g.nx, g.ny, 1, 0, nullptr, r, nullptr, nullptr, rb, 0);
    const double psum = r[0], pmax = r[2];
    nonfinite += r[3];
    reduce_field<R>(S, bucket, g.nx, g.ny, 1, 3, nullptr, r, nullptr, nullptr, rb, 0);
    o[GWM_S_PRECIP_SUM] = psum;
    o[GWM_S_PRECIP_MEAN] = n2 ? psum / static_cast<double>(n2) : 0.0;
    o[GWM_S_PRECIP_MAX] = std::max(0.0, pmax);
    o[GWM_S_PRECIP_CELLS] = static_cast<double>(n2);
    o[GWM_S_PRECIPITATING_CELLS] = r[0];
    o[GWM_S_NONFINITE] = nonfinite;
    return;
  }
  std::vector<R> p(n2);
  HIPCHK(hipMemcpyAsync(p.data(), bucket, n2 * sizeof(R),
                        hipMemcpyDeviceToHost, S->ctx->stream));
  HIPCHK(hipStreamSynchronize(S->ctx->stream));
  double ps = 0.0, pc = 0.0, pm = 0.0;
  long wet = 0;
  for (size_t a = 0; a < n2; ++a) {
    const double v = static_cast<double>(p[a]);
    if (!std::isfinite(v)) {
      nonfinite += 1.0;
      continue;
    }
    const double s2 = ps + v;
    pc += std::fabs(ps) >= std::fabs(v) ? (ps - s2) + v : (v - s2) + ps;
    ps = s2;
    pm = std::max(pm, v);
    if (v > 0.0) ++wet;
  }
  o[GWM_S_PRECIP_SUM] = ps + pc;
  o[GWM_S_PRECIP_MEAN] = n2 ? (ps + pc) / static_cast<double>(n2) : 0.0;
  o[GWM_S_PRECIP_MAX] = pm;
  o[GWM_S_PRECIP_CELLS] = static_cast<double>(n2);
  o[GWM_S_PRECIPITATING_CELLS] = static_cast<double>(wet);
  o[GWM_S_NONFINITE] = nonfinite;
}

// ---------------------------------------------------------------------------
// Host initial states.  Built in the working precision on the host with the
// reference's formulas, then uploaded.
// ---------------------------------------------------------------------------
template <class R>
struct HostThermo;
template <>
struct HostThermo<float> {
  static constexpr float Rd = 287.05f, p0 = 100000.0f, kappa = 0.2854f, g = 9.81f;
  static constexpr float pi = 3.14159265358979323846f;
};
template <>
struct HostThermo<double> {
  static constexpr double Rd = 287.05, p0 = 100000.0, kappa = 0.2854, g = 9.81;
  static constexpr double pi = 3.14159265358979323846;
};

template <class R>
R h_pressure(R rho, R rt) {
  using T = HostThermo<R>;
  const R th = rt / std::max(rho, R(1.0e-6));
  const R ratio = std::max(rho * T::Rd * th / T::p0, R(1.0e-6));
  return T::p0 * std::pow(ratio, R(1) / (R(1) - T::kappa));
}
template <class R>
R h_rho_from_p_theta(R p, R th) {
  using T = HostThermo<R>;
  const R ex = std::pow(std::max(p / T::p0, R(1.0e-6)), T::kappa);
  return p / std::max(T::Rd * th * ex, R(1.0e-6));
}
// hydrostatic_rho_at_height (dry_core.cu:22-35)
template <class R>
R h_hydro_rho(R rho_s, R th, R z) {
  using T = HostThermo<R>;
  const R ps = h_pressure(rho_s, rho_s * th);
  const R exs = std::pow(std::max(ps / T::p0, R(1.0e-6)), T::kappa);
  const R ex = std::max(R(1.0e-4), exs - T::g * z / ((T::Rd / T::kappa) * th));
  const R p = T::p0 * std::pow(ex, R(1) / T::kappa);
  return h_rho_from_p_theta(p, th);
}

// ---- perturbation form: bases and conversions (host, FP64) ----------------
double eos64(double rho, double rt) {  // dry_thermo.hpp:26-33 in FP64
  const double Rd = 287.05, p0 = 1.0e5, kap = 0.2854;
  const double th = rt / std::max(rho, 1.0e-6);
  return p0 * std::pow(std::max(rho * Rd * th / p0, 1.0e-6), 1.0 / (1.0 - kap));
}

// FP64 level means of the owned cells of a full cell field (storage layout).
template <class T>
std::vector<double> level_means(const gwm_state* S, const T* full) {
  const Geo& g = S->g();
  std::vector<double> m(static_cast<size_t>(g.nz), 0.0);
  for (int k = 0; k < g.nz; ++k) {
    double s = 0.0;
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) s += static_cast<double>(full[g.cell(i, j, k)]);
    m[static_cast<size_t>(k)] = s / (static_cast<double>(g.nx) * g.ny);
  }
  return m;
}

// Terrain form: the base of a field is the field itself (FP64), over the
// whole cell storage; halo positions take the value of the owned cell they
// stand for (periodic wrap, or the nearest owned cell across an open side),
// so every stored position holds a valid base.
template <class T>
std::vector<double> cell_base(const gwm_state* S, const T* full) {
  const Geo& g = S->g();
  std::vector<double> b(g.cell_total());
  for (int k = 0; k < g.nz; ++k)
    for (int j = -g.h; j < g.ny + g.h; ++j)
      for (int i = -g.h; i < g.nx + g.h; ++i) {
        const int oi = g.per_x ? ((i % g.nx) + g.nx) % g.nx : std::clamp(i, 0, g.nx - 1);
        const int oj = g.per_y ? ((j % g.ny) + g.ny) % g.ny : std::clamp(j, 0, g.ny - 1);
        b[g.cell(i, j, k)] = static_cast<double>(full[g.cell(oi, oj, k)]);
      }
  return b;
}

// The base of a field from full values: level means (flat) or the values
// themselves (terrain).
template <class T>
std::vector<double> field_base(const gwm_state* S, const T* full) {
  return S->pert3 ? cell_base(S, full) : level_means(S, full);
}

// Terrain form: level values, per-cell deviations and the base forcing,
// from the FP64 bases.  The forcing is the base state's share of the fast
// and slow pressure-gradient terms, in FP64 on the FP64 metrics (the
// device metric formulas, grid_metrics.hpp:184-236): the terrain-following
// horizontal gradient of p_b (dry_pressure_gradient.cu:37-73) on the x and
// y faces, and -dp_b/dz - g rho_b on the interior z faces (K14 without a
// reference state).
void push_base3(gwm_state* S) {
  const Geo& g = S->g();
  const int nz = g.nz;
  const size_t nc = g.cell_total();
  const std::vector<double>& br = S->base_rho64;
  const std::vector<double>& bt = S->base_rt64;
  // level means over the owned cells
  std::vector<double> lr(static_cast<size_t>(nz), 0.0), lt(static_cast<size_t>(nz), 0.0);
  for (int k = 0; k < nz; ++k) {
    double sr = 0.0, st = 0.0;
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        sr += br[g.cell(i, j, k)];
        st += bt[g.cell(i, j, k)];
      }
    lr[static_cast<size_t>(k)] = sr / (static_cast<double>(g.nx) * g.ny);
    lt[static_cast<size_t>(k)] = st / (static_cast<double>(g.nx) * g.ny);
  }
  std::vector<float> lev(4 * static_cast<size_t>(nz));
  for (int k = 0; k < nz; ++k) {
    const double r = lr[static_cast<size_t>(k)], t = lt[static_cast<size_t>(k)];
    lev[k] = static_cast<float>(r);
    lev[nz + k] = static_cast<float>(t);
    lev[2 * nz + k] = static_cast<float>(t / r);
    lev[3 * nz + k] = static_cast<float>(eos64(r, t));
  }
  // per-cell base pressure and the deviations
  std::vector<double> pb(nc);
  std::vector<float> dev(3 * nc);
  const size_t plane = static_cast<size_t>(g.pxc) * g.pyc;
  for (size_t a = 0; a < nc; ++a) {
    const size_t k = a / plane;
    pb[a] = eos64(br[a], bt[a]);
    dev[a] = static_cast<float>(br[a] - static_cast<double>(lev[k]));
    dev[nc + a] = static_cast<float>(bt[a] / br[a] - static_cast<double>(lev[2 * nz + k]));
    dev[2 * nc + a] = static_cast<float>(pb[a] - static_cast<double>(lev[3 * nz + k]));
  }
  // FP64 metrics: z of the cell centres, zc[k] + wc[k] h(i, j)
  const size_t n2 = static_cast<size_t>(g.nx) * g.ny;
  std::vector<float> ter(n2, 0.0f);
  scopy(S->ctx->stream, ter.data(), S->grid->terr, n2 * sizeof(float), hipMemcpyDeviceToHost);
  auto zc = [&](int i, int j, int k) {
    const double th = static_cast<double>(ter[static_cast<size_t>(j) * g.nx + i]);
    const double prod = S->grid->w_c[static_cast<size_t>(k)] * th;
    return S->grid->z_c[static_cast<size_t>(k)] + prod;
  };
  auto wrap_i = [&](int i) { return g.per_x ? ((i % g.nx) + g.nx) % g.nx : std::clamp(i, 0, g.nx - 1); };
  auto wrap_j = [&](int j) { return g.per_y ? ((j % g.ny) + g.ny) % g.ny : std::clamp(j, 0, g.ny - 1); };
  auto P = [&](int i, int j, int k) { return pb[g.cell(wrap_i(i), wrap_j(j), k)]; };
  auto vgrad = [&](int i, int j, int k) -> double {  // dp_b/dz of column (i, j)
    if (nz == 1) return 0.0;
    const int ii = wrap_i(i), jj = wrap_j(j);
    if (k == 0) return (P(ii, jj, 1) - P(ii, jj, 0)) / std::max(zc(ii, jj, 1) - zc(ii, jj, 0), 1.0e-6);
    if (k + 1 == nz)
      return (P(ii, jj, k) - P(ii, jj, k - 1)) / std::max(zc(ii, jj, k) - zc(ii, jj, k - 1), 1.0e-6);
    return (P(ii, jj, k + 1) - P(ii, jj, k - 1)) / std::max(zc(ii, jj, k + 1) - zc(ii, jj, k - 1), 1.0e-6);
  };
  auto pgf = [&](int il, int jl, int ir, int jr, int k, double d) -> double {
    const double dzd = (zc(wrap_i(ir), wrap_j(jr), k) - zc(wrap_i(il), wrap_j(jl), k)) / d;
    return (P(ir, jr, k) - P(il, jl, k)) / d - dzd * (0.5 * (vgrad(il, jl, k) + vgrad(ir, jr, k)));
  };
  const size_t nxf = g.xf_total(), nyf = g.yf_total(), nzf = g.zf_total();
  std::vector<float> frc(nxf + nyf + nzf, 0.0f);
  const double grav = S->forcing_g;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i <= g.nx; ++i) {
        if (!g.per_x && (i == 0 || i == g.nx)) continue;  // closed faces of open sides
        frc[g.xf(i, j, k)] = static_cast<float>(pgf(i - 1, j, i, j, k, g.dx));
      }
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j <= g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        if (!g.per_y && (j == 0 || j == g.ny)) continue;
        frc[nxf + g.yf(i, j, k)] = static_cast<float>(pgf(i, j - 1, i, j, k, g.dy));
      }
  // z faces: K14 of the base with the column reference, as the FP64 step
  // computes it, is the plain -dp_b/dz - g rho_b plus the reference's share
  // (p_ref,u - p_ref,l)/dz + g (rho_ref,l + rho_ref,u)/2, which its
  // recurrence (four fixed-point iterations) leaves at the last iteration's
  // change; that share, a smooth function of height, is kept here from the
  // base column, so the departures' K14 needs no reference state.
  const double Rd = 287.05, p0 = 1.0e5, kap = 0.2854;
  auto rho_of = [&](double p, double th) {  // dry_thermo.hpp:35-40
    const double ex = std::pow(std::max(p / p0, 1.0e-6), kap);
    return p / std::max(Rd * th * ex, 1.0e-6);
  };
  for (int j = 0; j < g.ny; ++j)
    for (int i = 0; i < g.nx; ++i) {
      const size_t c0 = g.cell(i, j, 0);
      double pref = eos64(br[c0], bt[c0]);
      double rref = rho_of(pref, bt[c0] / std::max(br[c0], 1.0e-6));
      for (int k = 1; k < nz; ++k) {
        const size_t lo = g.cell(i, j, k - 1), hi = g.cell(i, j, k);
        const double thk = bt[hi] / std::max(br[hi], 1.0e-6);
        const double dz = zc(i, j, k) - zc(i, j, k - 1);
        double pn = std::max(1.0, pref - grav * rref * dz);
        for (int it = 0; it < 4; ++it) {
          const double rn = rho_of(pn, thk);
          pn = std::max(pref - grav * 0.5 * (rref + rn) * dz, 1.0);
        }
        const double rref1 = rho_of(pn, thk);
        const double idz = 1.0 / std::max(dz, 1.0e-6);
        const double share = (pn - pref) * idz + grav * 0.5 * (rref + rref1);
        frc[nxf + nyf + g.zf(i, j, k)] = static_cast<float>(
            -(pb[hi] - pb[lo]) * idz - grav * (0.5 * (br[lo] + br[hi])) + share);
        pref = pn;
        rref = rref1;
      }
    }
  if (!S->b_dev) S->b_dev = dalloc(lev.size() * sizeof(float));
  if (!S->b3_dev) S->b3_dev = dalloc(dev.size() * sizeof(float));
  if (!S->bf_dev) S->bf_dev = dalloc(frc.size() * sizeof(float));
  scopy(S->ctx->stream, S->b_dev, lev.data(), lev.size() * sizeof(float), hipMemcpyHostToDevice);
  scopy(S->ctx->stream, S->b3_dev, dev.data(), dev.size() * sizeof(float), hipMemcpyHostToDevice);
  scopy(S->ctx->stream, S->bf_dev, frc.data(), frc.size() * sizeof(float), hipMemcpyHostToDevice);
}

void push_base(gwm_state* S) {
  if (S->pert3) {
    push_base3(S);
    return;
  }
  const int nz = S->g().nz;
  std::vector<float> b(4 * static_cast<size_t>(nz));
  for (int k = 0; k < nz; ++k) {
    const double r = S->base_rho64[static_cast<size_t>(k)];
    const double t = S->base_rt64[static_cast<size_t>(k)];
    b[k] = static_cast<float>(r);
    b[nz + k] = static_cast<float>(t);
    b[2 * nz + k] = static_cast<float>(t / r);
    b[3 * nz + k] = static_cast<float>(eos64(r, t));
  }
  if (!S->b_dev) S->b_dev = dalloc(b.size() * sizeof(float));
  scopy(S->ctx->stream, S->b_dev, b.data(), b.size() * sizeof(float), hipMemcpyHostToDevice);
}

// Full cell field -> departures from `base` (its level means, or on
// terrain the 3-D base), FP32, device.
template <class T>
void upload_departure(gwm_state* S, void* dst, const T* full, const std::vector<double>& base) {
  const Geo& g = S->g();
  const size_t n = g.cell_total();
  std::vector<float> d(n);
  for (size_t a = 0; a < n; ++a)
    d[a] = static_cast<float>(static_cast<double>(full[a]) - base[S->bidx(a)]);
  scopy(S->ctx->stream, dst, d.data(), n * sizeof(float), hipMemcpyHostToDevice);
}

// Whole state in the perturbation form: bases from the FP64 level means of
// rho and rho*theta, departures and the other fields in FP32.
template <class T>
void upload_pert_fields(gwm_state* S, const T* rho, const T* rt, const T* u, const T* v,
                        const T* w, const std::vector<std::vector<T>>& q) {
  S->base_rho64 = field_base(S, rho);
  S->base_rt64 = field_base(S, rt);
  push_base(S);
  upload_departure(S, S->cur_set()[0], rho, S->base_rho64);
  upload_departure(S, S->cur_set()[1], rt, S->base_rt64);
  const T* f[3] = {u, v, w};
  for (int a = 2; a < 5; ++a) {
    const size_t n = S->fsize(a);
    std::vector<float> h(n);
    for (size_t e = 0; e < n; ++e) h[e] = static_cast<float>(f[a - 2][e]);
    scopy(S->ctx->stream, S->cur_set()[a], h.data(), n * sizeof(float), hipMemcpyHostToDevice);
  }
  for (size_t t = 0; t < q.size() && t < static_cast<size_t>(S->ntr); ++t) {
    std::vector<float> h(q[t].size());
    for (size_t e = 0; e < h.size(); ++e) h[e] = static_cast<float>(q[t][e]);
    scopy(S->ctx->stream, S->field(GWM_TRACER0 + static_cast<int>(t)), h.data(),
                     h.size() * sizeof(float), hipMemcpyHostToDevice);
  }
  S->faces_synced = false;
}

template <class R>
struct HostState {
  std::vector<R> rho, rt, u, v, w;
  std::vector<std::vector<R>> q;
  HostState(const gwm_state* S) {
    const Geo& g = S->g();
    rho.assign(g.cell_total(), R(0));
    rt.assign(g.cell_total(), R(0));
    u.assign(g.xf_total(), R(0));
    v.assign(g.yf_total(), R(0));
    w.assign(g.zf_total(), R(0));
    q.assign(S->ntr, std::vector<R>(g.cell_total(), R(0)));
  }
  void upload(gwm_state* S) {
    if (S->pert) {
      upload_pert_fields(S, rho.data(), rt.data(), u.data(), v.data(), w.data(), q);
      return;
    }
    std::vector<R>* f[5] = {&rho, &rt, &u, &v, &w};
    for (int a = 0; a < 5; ++a)
      scopy(S->ctx->stream, S->cur_set()[a], f[a]->data(), f[a]->size() * sizeof(R),
            hipMemcpyHostToDevice);
    for (int t = 0; t < S->ntr; ++t)
      scopy(S->ctx->stream, S->field(GWM_TRACER0 + t), q[t].data(), q[t].size() * sizeof(R),
            hipMemcpyHostToDevice);
    S->faces_synced = false;
  }
};

template <class R>
void fill_constant(const gwm_state* S, HostState<R>& H, R rho, R th, R u, R v, R w) {
  (void)S;
  std::fill(H.rho.begin(), H.rho.end(), rho);
  std::fill(H.rt.begin(), H.rt.end(), rho * th);
  std::fill(H.u.begin(), H.u.end(), rho * u);
  std::fill(H.v.begin(), H.v.end(), rho * v);
  std::fill(H.w.begin(), H.w.end(), rho * w);
}

template <class R>
std::vector<R> host_metric(const gwm_state* S, int which) {
  const Geo& g = S->g();
  const size_t n2 = static_cast<size_t>(g.nx) * g.ny;
  const size_t n = which == 0 || which == 3 ? n2 * (g.nz + 1) : n2 * g.nz;
  std::vector<R> out(n);
  const void* src = which == 0 ? S->grid->zif_m : which == 1 ? S->grid->zc_m
                  : which == 2 ? S->grid->idz_c : S->grid->idz_f;
  if (S->ctx->esize() == sizeof(R)) {
    scopy(S->ctx->stream, out.data(), src, n * sizeof(R), hipMemcpyDeviceToHost);
  } else if (S->pert3) {
    // FP64 host build of a perturbation-form state on terrain: the FP64
    // metrics themselves (the device formulas, k_metrics), so its initial
    // state, which is also its base, is the FP64 context's bit for bit.
    std::vector<float> ter(n2);
    scopy(S->ctx->stream, ter.data(), S->grid->terr, n2 * sizeof(float), hipMemcpyDeviceToHost);
    const gwm_grid* G = S->grid;
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        const double thd = static_cast<double>(ter[static_cast<size_t>(j) * g.nx + i]);
        auto zif = [&](int k) { const double prod = G->w_if[k] * thd; return G->z_if[k] + prod; };
        auto zcm = [&](int k) { const double prod = G->w_c[k] * thd; return G->z_c[k] + prod; };
        const int kn = which == 0 || which == 3 ? g.nz + 1 : g.nz;
        for (int k = 0; k < kn; ++k) {
          double v;
          if (which == 0) {
            v = zif(k);
          } else if (which == 1) {
            v = zcm(k);
          } else if (which == 2) {
            v = 1.0 / std::fmax(zif(k + 1) - zif(k), 1.0e-6);
          } else {
            v = (k == 0 || k == g.nz) ? 0.0 : 1.0 / std::fmax(zcm(k) - zcm(k - 1), 1.0e-6);
          }
          out[g.metric(i, j, k)] = static_cast<R>(v);
        }
      }
  } else {  // FP64 host build of a perturbation-form (FP32) state
    std::vector<float> tmp(n);
    scopy(S->ctx->stream, tmp.data(), src, n * sizeof(float), hipMemcpyDeviceToHost);
    for (size_t a = 0; a < n; ++a) out[a] = static_cast<R>(tmp[a]);
  }
  return out;
}

template <class R>
void init_hydro(gwm_state* S, double rho_s, double th, HostState<R>& H) {
  const Geo& g = S->g();
  fill_constant<R>(S, H, R(0), static_cast<R>(th), R(0), R(0), R(0));
  const std::vector<R> zc = host_metric<R>(S, 1);
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        const R r = h_hydro_rho(static_cast<R>(rho_s), static_cast<R>(th), zc[g.metric(i, j, k)]);
        H.rho[g.cell(i, j, k)] = r;
        H.rt[g.cell(i, j, k)] = r * static_cast<R>(th);
      }
  std::fill(H.u.begin(), H.u.end(), R(0));
  std::fill(H.v.begin(), H.v.end(), R(0));
  std::fill(H.w.begin(), H.w.end(), R(0));
}

// make_bubble_like_state (idealized_cases.cpp:105-162)
template <class R>
void init_bubble(gwm_state* S, const double* p, int sign) {
  using T = HostThermo<R>;
  const Geo& g = S->g();
  HostState<R> H(S);
  const bool hydro = p[0] <= 0.0;
  if (hydro) init_hydro<R>(S, -p[0], p[1], H);
  else fill_constant<R>(S, H, static_cast<R>(p[0]), static_cast<R>(p[1]), static_cast<R>(p[2]),
                        static_cast<R>(p[3]), static_cast<R>(p[4]));
  const std::vector<R> zc = host_metric<R>(S, 1);
  const double z_top = S->grid->z_if.back();
  const R dxr = static_cast<R>(g.dx), dyr = static_cast<R>(g.dy);
  const R x0 = static_cast<R>(p[5]) * static_cast<R>(g.nx - 1) * dxr;
  const R y0 = static_cast<R>(p[6]) * static_cast<R>(g.ny - 1) * dyr;
  const R z0 = static_cast<R>(p[7]) * static_cast<R>(z_top);
  const R rx = std::max(static_cast<R>(p[8]) * static_cast<R>(g.nx) * dxr, R(1));
  const R ry = std::max(static_cast<R>(p[9]) * static_cast<R>(g.ny) * dyr, R(1));
  const R rz = std::max(static_cast<R>(p[10]) * static_cast<R>(z_top), R(1));
  const R dth = static_cast<R>(sign >= 0 ? std::fabs(p[11]) : -std::fabs(p[11]));
  const R thb = static_cast<R>(p[1]);
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j) {
      const R y = static_cast<R>(j) * dyr;
      for (int i = 0; i < g.nx; ++i) {
        const R x = static_cast<R>(i) * dxr;
        const R z = zc[g.metric(i, j, k)];
        const R wx = (x - x0) / rx, wy = (y - y0) / ry, wz = (z - z0) / rz;
        const R r2 = wx * wx + wy * wy + wz * wz;
        const R wgt = r2 >= R(1) ? R(0) : R(0.5) * (R(1) + std::cos(T::pi * std::sqrt(r2)));
        const R rb = hydro ? H.rho[g.cell(i, j, k)] : static_cast<R>(p[0]);
        H.rt[g.cell(i, j, k)] = rb * (thb + dth * wgt);
      }
    }
  H.upload(S);
}

// set_background_face_momenta (idealized_cases.cpp:36-103)
template <class R>
void face_momenta(gwm_state* S, HostState<R>& H, R U, R V, R W) {
  const Geo& g = S->g();
  auto rho = [&](int i, int j, int k) { return H.rho[g.cell(i, j, k)]; };
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i <= g.nx; ++i) {
        R f;
        if (g.nx == 1 || i <= 0) f = rho(0, j, k);
        else if (i >= g.nx) f = rho(g.nx - 1, j, k);
        else f = R(0.5) * (rho(i - 1, j, k) + rho(i, j, k));
        H.u[g.xf(i, j, k)] = f * U;
      }
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j <= g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        R f;
        if (g.ny == 1 || j <= 0) f = rho(i, 0, k);
        else if (j >= g.ny) f = rho(i, g.ny - 1, k);
        else f = R(0.5) * (rho(i, j - 1, k) + rho(i, j, k));
        H.v[g.yf(i, j, k)] = f * V;
      }
  for (int k = 0; k <= g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        R f;
        if (g.nz == 1 || k <= 0) f = rho(i, j, 0);
        else if (k >= g.nz) f = rho(i, j, g.nz - 1);
        else f = R(0.5) * (rho(i, j, k - 1) + rho(i, j, k));
        H.w[g.zf(i, j, k)] = f * W;
      }
}

// make_mountain_wave_background_state (idealized_cases.cpp:182-223)
template <class R>
void init_mountain(gwm_state* S, double rho_s, double th, double u, double v, double w) {
  const Geo& g = S->g();
  HostState<R> H(S);
  fill_constant<R>(S, H, R(0), static_cast<R>(th), static_cast<R>(u), static_cast<R>(v),
                   static_cast<R>(w));
  std::vector<R> terr(static_cast<size_t>(g.nx) * g.ny);
  if (S->ctx->esize() == sizeof(R)) {
    scopy(S->ctx->stream, terr.data(), S->grid->terr, terr.size() * sizeof(R), hipMemcpyDeviceToHost);
  } else {  // FP64 host build of a perturbation-form (FP32) state
    std::vector<float> tmp(terr.size());
    scopy(S->ctx->stream, tmp.data(), S->grid->terr, tmp.size() * sizeof(float), hipMemcpyDeviceToHost);
    for (size_t a = 0; a < tmp.size(); ++a) terr[a] = static_cast<R>(tmp[a]);
  }
  const auto& zc = S->grid->z_c;
  const auto& wc = S->grid->w_c;
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        const double z = zc[k] + wc[k] * static_cast<double>(terr[static_cast<size_t>(j) * g.nx + i]);
        const R r = h_hydro_rho(static_cast<R>(rho_s), static_cast<R>(th), static_cast<R>(z));
        H.rho[g.cell(i, j, k)] = r;
        H.rt[g.cell(i, j, k)] = r * static_cast<R>(th);
      }
  face_momenta<R>(S, H, static_cast<R>(u), static_cast<R>(v), static_cast<R>(w));
  H.upload(S);
}

// Extension: constant-N sounding on terrain.  theta(z) = theta_s exp(N^2 z/g);
// Exner from the closed-form hydrostatic integral
// Pi(z) = Pi_s - g^2/(c_p theta_s N^2) (1 - exp(-N^2 z / g)), in FP64.
template <class R>
void init_constant_n(gwm_state* S, double th_s, double N, double p_s, double u, double v) {
  const Geo& g = S->g();
  HostState<R> H(S);
  const double Rd = 287.05, kap = 0.2854, cp = Rd / kap, grav = 9.81, p0 = 1.0e5;
  const double Pis = std::pow(p_s / p0, kap);
  const std::vector<R> zc = host_metric<R>(S, 1);
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        const double z = static_cast<double>(zc[g.metric(i, j, k)]);
        const double th = th_s * std::exp(N * N * z / grav);
        const double Pi = Pis - grav * grav / (cp * th_s * N * N) * (1.0 - std::exp(-N * N * z / grav));
        const double p = p0 * std::pow(Pi, 1.0 / kap);
        const double r = p / (Rd * th * Pi);
        H.rho[g.cell(i, j, k)] = static_cast<R>(r);
        H.rt[g.cell(i, j, k)] = static_cast<R>(r) * static_cast<R>(th);
      }
  face_momenta<R>(S, H, static_cast<R>(u), static_cast<R>(v), R(0));
  H.upload(S);
}

// Counter-based uniform deviate in [-0.5, 0.5) keyed by (seed, i, j, k)
// (splitmix64 finaliser), identical on every host.
double noise(unsigned long long seed, int i, int j, int k) {
  uint64_t x = seed * 0x9E3779B97F4A7C15ull ^ (static_cast<uint64_t>(i) << 42) ^
               (static_cast<uint64_t>(j) << 21) ^ static_cast<uint64_t>(k);
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  x ^= x >> 31;
  return static_cast<double>(x >> 11) * (1.0 / 9007199254740992.0) - 0.5;
}

// Extension: warm-rain convection sounding (flat terrain), FP64 column
// integration of the Exner function with a fine trapezoid rule.
template <class R>
void init_convection(gwm_state* S, double p_s, unsigned long long seed, double amp, double ztop_noise) {
  const Geo& g = S->g();
  require(!S->grid->terrain_nonzero, "the convection sounding expects flat terrain");
  require(S->ntr >= 3, "the convection sounding needs the warm-rain tracers");
  HostState<R> H(S);
  const double Rd = 287.05, kap = 0.2854, cp = Rd / kap, grav = 9.81, p0 = 1.0e5, ztr = 12000.0;
  const double Rv = 461.5, eps = Rd / Rv;
  auto theta_at = [&](double z, double Ttr, double Pitr) {
    if (z <= ztr) return 300.0 + 43.0 * std::pow(z / ztr, 1.25);
    const double Pi = Pitr * std::exp(-grav * (z - ztr) / (cp * Ttr));
    return Ttr / Pi;
  };
  // Exner at 12 km by integrating dPi/dz = -g/(c_p theta)
  const double Pis = std::pow(p_s / p0, kap);
  auto integrate = [&](double z1, double Pi0) {
    const int n = 20000;
    double Pi = Pi0, zz = 0.0;
    const double h = z1 / n;
    for (int s = 0; s < n; ++s) {
      const double a = 300.0 + 43.0 * std::pow(std::min(zz, ztr) / ztr, 1.25);
      const double b = 300.0 + 43.0 * std::pow(std::min(zz + h, ztr) / ztr, 1.25);
      Pi -= 0.5 * h * grav / cp * (1.0 / a + 1.0 / b);
      zz += h;
    }
    return Pi;
  };
  const double Pitr = integrate(ztr, Pis);
  const double Ttr = 343.0 * Pitr;
  const auto& zc = S->grid->z_c;
  std::vector<double> th(g.nz), rho(g.nz), qv(g.nz);
  for (int k = 0; k < g.nz; ++k) {
    const double z = zc[k];
    const double Pi = z <= ztr ? integrate(z, Pis) : Pitr * std::exp(-grav * (z - ztr) / (cp * Ttr));
    th[k] = theta_at(z, Ttr, Pitr);
    const double p = p0 * std::pow(Pi, 1.0 / kap);
    rho[k] = p / (Rd * th[k] * Pi);
    const double T = th[k] * Pi;
    const double rh = z <= ztr ? 1.0 - 0.75 * std::pow(z / ztr, 1.25) : 0.25;
    const double tc = T - 273.15;
    const double es = 6.112 * std::exp(17.67 * tc / (tc + 243.5)) * 100.0;
    const double qs = std::min(std::max(eps * es / std::max(p - (1.0 - eps) * es, 1000.0), 0.0), 0.05);
    qv[k] = std::min(0.014, rh * qs);
  }
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        const double t = th[k] + (zc[k] < ztop_noise ? 2.0 * amp * noise(seed, i, j, k) : 0.0);
        const R r = static_cast<R>(rho[k]);
        H.rho[g.cell(i, j, k)] = r;
        H.rt[g.cell(i, j, k)] = r * static_cast<R>(t);
        H.q[0][g.cell(i, j, k)] = r * static_cast<R>(qv[k]);
      }
  H.upload(S);
}

// extract_dry_plan_view / extract_runtime_plan_view
// (src/io/plan_view_output.cpp:601-839) on the device: one column pass
// writes every product, the host downloads the [field][j][i] planes.
template <class R>
int plan_view(gwm_state* S, int slice_k, int steps, double dt, int with_precip, double* host) {
  const Geo& g = S->g();
  const int k = std::clamp(slice_k, 0, g.nz - 1);
  const int nf = S->ntr == 0 ? 10 : (S->ntr < 3 ? 13 : (with_precip ? 23 : 21));
  const size_t plane = static_cast<size_t>(g.nx) * g.ny;
  struct Buf {
    double* p = nullptr;
    ~Buf() {
      if (p) (void)hipFree(p);
    }
  } buf;
  buf.p = static_cast<double*>(dalloc(plane * gwm::kPvCount * sizeof(double)));
  void** ds = S->cur_set();
  gwm::PlanViewArgs<R> a{};
  a.g = g;
  a.rho = static_cast<const R*>(ds[0]);
  a.rt = static_cast<const R*>(ds[1]);
  a.u = static_cast<const R*>(ds[2]);
  a.v = static_cast<const R*>(ds[3]);
  a.w = static_cast<const R*>(ds[4]);
  if (S->ntr >= 1) a.qv = static_cast<const R*>(S->field(GWM_TRACER0));
  if (S->ntr >= 3) {
    a.qc = static_cast<const R*>(S->field(GWM_TRACER0 + 1));
    a.qr = static_cast<const R*>(S->field(GWM_TRACER0 + 2));
    if (with_precip) a.precip = static_cast<const R*>(S->view ? S->owner->precip : S->precip);
  }
  a.idz_c = static_cast<const R*>(S->grid->idz_c);
  a.lev_idz_c = static_cast<const R*>(S->grid->lev_idz_c);
  a.zc = static_cast<const R*>(S->grid->zc_m);
  a.terrain = static_cast<const R*>(S->grid->terr);
  a.base = S->base<R>();
  a.k = k;
  a.elapsed_hours = std::max(static_cast<double>(steps) * dt / 3600.0, 1.0e-9);
  a.out = buf.p;
  {
    Timed t(S->ctx, "plan_view");
    gwm::launch_plan_view<R>(S->ctx->stream, a, S->grid->terrain_nonzero);
  }
  scopy(S->ctx->stream, host, buf.p, plane * nf * sizeof(double), hipMemcpyDeviceToHost);
  return nf;
}

template <class F>
auto dispatch(const gwm_ctx* c, F f) {
  if (c->prec == GWM_F64) return f(double{});
  return f(float{});
}

// Dispatch of the host-built initial states: FP64 host arithmetic for the
// perturbation form, the context precision otherwise.
template <class F>
auto init_dispatch(gwm_state* S, F f) {
  if (S->pert) return f(double{});
  return dispatch(S->ctx, f);
}


}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" {

const char* gwm_last_error(void) { return g_err.c_str(); }

int gwm_ctx_create(int device, int precision, gwm_ctx** out) {
  return guard([&] {
    require(out != nullptr, "null output");
    require(precision == GWM_F32 || precision == GWM_F64, "precision must be GWM_F32 or GWM_F64");
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    require(device >= 0 && device < n, "no such HIP device");
    HIPCHK(hipSetDevice(device));
    auto c = std::make_unique<gwm_ctx>();
    c->device = device;
    c->prec = precision;
    const char* ch = std::getenv("GWM_CHUNKS");
    if (ch != nullptr) c->force_chunks = std::max(0, std::atoi(ch));
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    const hipError_t se = gwm::noalloc_scratch_setup(device, &c->noalloc);
    if (se != hipSuccess) {
      (void)hipStreamDestroy(c->stream);
      require(false, std::string("hipGetSymbolAddress of the budget scratch failed: ") +
                         hipGetErrorString(se));
    }
    *out = c.release();
  });
}

namespace {
void ctx_free(gwm_ctx* ctx) {
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& p : ctx->pending) {
      (void)hipEventDestroy(p.a);
      (void)hipEventDestroy(p.b);
    }
    for (auto e : ctx->pool) (void)hipEventDestroy(e);
    if (ctx->partials) (void)hipFree(ctx->partials);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}
}  // namespace

int gwm_ctx_destroy(gwm_ctx* ctx) {
  return guard([&] {
    if (!ctx) return;
    // Grids and states still alive keep the context (and its stream) until
    // the last of them is destroyed.
    ctx->closing = true;
    if (ctx->refs <= 0) ctx_free(ctx);
  });
}

int gwm_sync(gwm_ctx* ctx) {
  return guard([&] {
    require(ctx, "null context");
    HIPCHK(hipStreamSynchronize(ctx->stream));
  });
}

int gwm_ctx_precision(const gwm_ctx* ctx) { return ctx ? ctx->prec : -1; }

int gwm_ctx_set_kernel_timing(gwm_ctx* ctx, int enable) {
  return guard([&] {
    require(ctx, "null context");
    ctx->timing = enable != 0;
  });
}

int gwm_ctx_kernel_stats(gwm_ctx* ctx, const char* name, long* launches, double* total_ms) {
  return guard([&] {
    require(ctx && name, "null argument");
    ctx->resolve();
    auto it = ctx->stats.find(name);
    if (launches) *launches = it == ctx->stats.end() ? 0 : it->second.launches;
    if (total_ms) *total_ms = it == ctx->stats.end() ? 0.0 : it->second.ms;
  });
}

int gwm_ctx_set_chunks(gwm_ctx* ctx, int chunks) {
  return guard([&] {
    require(ctx, "null context");
    require(chunks >= 0, "chunks must be >= 0");
    ctx->force_chunks = chunks;
  });
}

int gwm_fp32_math(gwm_ctx* ctx, int fn, const void* dev_x, void* dev_y, long n) {
  return guard([&] {
    require(ctx && dev_x && dev_y, "null argument");
    require(fn >= GWM_MATH_POW_EOS && fn <= GWM_MATH_POW1PM1, "unknown function");
    require(n >= 0, "n must be >= 0");
    if (n == 0) return;
    HIPCHK(hipSetDevice(ctx->device));
    gwm::launch_fp32_math(ctx->stream, fn, static_cast<const float*>(dev_x),
                          static_cast<float*>(dev_y), static_cast<size_t>(n));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
  });
}

int gwm_ctx_tuning_report(gwm_ctx* ctx, char* buf, size_t size) {
  return guard([&] {
    require(ctx && buf && size > 0, "null argument");
    std::string r;
    for (const auto& kv : ctx->tuned) {
      if (!r.empty()) r += ";";
      r += kv.first + "=" + std::to_string(kv.second);
    }
    std::snprintf(buf, size, "%s", r.c_str());
  });
}

int gwm_ctx_reset_kernel_stats(gwm_ctx* ctx) {
  return guard([&] {
    require(ctx, "null context");
    ctx->resolve();
    ctx->stats.clear();
  });
}

int gwm_grid_create(gwm_ctx* ctx, const gwm_grid_desc* d, const double* z_interfaces,
                    const float* terrain, gwm_grid** out) {
  return guard([&] {
    require(ctx && d && out, "null argument");
    require(d->nx > 0 && d->ny > 0 && d->nz > 0, "grid dimensions must be positive");
    // the column kernels stage per-level values of a whole column in LDS
    require(d->nz <= 512, "at most 512 levels");
    require(d->halo >= 1, "halo must be at least 1");
    require(d->dx > 0 && d->dy > 0, "grid spacing must be positive");
    require(d->halo <= d->nx && d->halo <= d->ny, "halo wider than the domain");
    {
      const double elems = static_cast<double>(d->nx + 1 + 2 * d->halo) *
                           (d->ny + 1 + 2 * d->halo) * (d->nz + 1);
      require(elems < 2147483647.0, "grid too large: a field must hold fewer than 2^31 elements");
      // kernels address arrays by 32-bit byte offsets
      require(elems * static_cast<double>(ctx->esize()) < 4294967296.0,
              "grid too large: each field must stay below 4 GiB");
    }
    auto G = std::make_unique<gwm_grid>();
    require(!ctx->closing, "context already destroyed");
    G->ctx = ctx;
    ctx->refs += 1;
    Geo& g = G->geo;
    g.nx = d->nx;
    g.ny = d->ny;
    g.nz = d->nz;
    g.h = d->halo;
    g.per_x = d->periodic_x ? 1 : 0;
    g.per_y = d->periodic_y ? 1 : 0;
    g.pxc = d->nx + 2 * d->halo;
    g.pyc = d->ny + 2 * d->halo;
    g.pxu = d->nx + 1 + 2 * d->halo;
    g.pyv = d->ny + 1 + 2 * d->halo;
    g.dx = d->dx;
    g.dy = d->dy;
    const int nz = d->nz;
    G->z_if.resize(nz + 1);
    G->z_c.resize(nz);
    G->w_if.resize(nz + 1);
    G->w_c.resize(nz);
    if (!z_interfaces) {  // grid_metrics.hpp:146-161
      require(d->z_top > 0, "z_top must be positive");
      for (int k = 0; k <= nz; ++k) {
        const double eta = static_cast<double>(k) / static_cast<double>(nz);
        G->w_if[k] = taper_weight(eta, d->terrain_taper_eta);
        G->z_if[k] = eta * d->z_top;
        if (k < nz) {
          const double ec = (static_cast<double>(k) + 0.5) / static_cast<double>(nz);
          G->w_c[k] = taper_weight(ec, d->terrain_taper_eta);
          G->z_c[k] = ec * d->z_top;
        }
      }
    } else {  // stretched flat interfaces (extension)
      const double zt = z_interfaces[nz];
      for (int k = 0; k <= nz; ++k) {
        if (k > 0) require(z_interfaces[k] > z_interfaces[k - 1], "z_interfaces must increase");
        G->z_if[k] = z_interfaces[k];
        G->w_if[k] = taper_weight(z_interfaces[k] / zt, d->terrain_taper_eta);
      }
      for (int k = 0; k < nz; ++k) {
        G->z_c[k] = 0.5 * (z_interfaces[k] + z_interfaces[k + 1]);
        G->w_c[k] = taper_weight(G->z_c[k] / zt, d->terrain_taper_eta);
      }
    }
    dispatch(ctx, [&](auto r) {
      using R = decltype(r);
      build_grid<R>(G.get(), z_interfaces, terrain);
      return 0;
    });
    *out = G.release();
  });
}

int gwm_grid_destroy(gwm_grid* grid) {
  return guard([&] { delete grid; });
}

int gwm_grid_terrain_active(const gwm_grid* grid) { return grid && grid->terrain_active ? 1 : 0; }

int gwm_grid_download_metric(gwm_grid* G, int which, void* host) {
  return guard([&] {
    require(G && host, "null argument");
    const Geo& g = G->geo;
    const size_t n2 = static_cast<size_t>(g.nx) * g.ny;
    const size_t es = G->ctx->esize();
    switch (which) {
      case 0: scopy(G->ctx->stream, host, G->zif_m, n2 * (g.nz + 1) * es, hipMemcpyDeviceToHost); break;
      case 1: scopy(G->ctx->stream, host, G->zc_m, n2 * g.nz * es, hipMemcpyDeviceToHost); break;
      case 2: scopy(G->ctx->stream, host, G->idz_c, n2 * g.nz * es, hipMemcpyDeviceToHost); break;
      case 3: scopy(G->ctx->stream, host, G->idz_f, n2 * (g.nz + 1) * es, hipMemcpyDeviceToHost); break;
      case 4: scopy(G->ctx->stream, host, G->terr, n2 * es, hipMemcpyDeviceToHost); break;
      case 5: scopy(G->ctx->stream, host, G->sx, n2 * es, hipMemcpyDeviceToHost); break;
      case 6: scopy(G->ctx->stream, host, G->sy, n2 * es, hipMemcpyDeviceToHost); break;
      case 7: std::memcpy(host, G->z_if.data(), G->z_if.size() * sizeof(double)); break;
      case 8: std::memcpy(host, G->z_c.data(), G->z_c.size() * sizeof(double)); break;
      case 9: std::memcpy(host, G->w_if.data(), G->w_if.size() * sizeof(double)); break;
      case 10: std::memcpy(host, G->w_c.data(), G->w_c.size() * sizeof(double)); break;
      default: require(false, "bad metric id");
    }
  });
}

// idealized_domain_builder.cpp:14-35
int gwm_terrain_cosine_mountain(int nx, int ny, double cx, double cy, double hx, double hy,
                                double height, float* out) {
  return guard([&] {
    require(nx > 0 && ny > 0 && out, "bad arguments");
    const double kPi = 3.14159265358979323846;
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        float v = 0.0f;
        if (height > 0.0) {
          const double x = nx > 1 ? static_cast<double>(i) / (nx - 1) : 0.0;
          const double y = ny > 1 ? static_cast<double>(j) / (ny - 1) : 0.0;
          const double dx = (x - cx) / hx, dy = (y - cy) / hy;
          if (std::fabs(dx) < 1.0 && std::fabs(dy) < 1.0)
            v = static_cast<float>(height * (0.5 * (1.0 + std::cos(kPi * dx))) *
                                   (0.5 * (1.0 + std::cos(kPi * dy))));
        }
        out[static_cast<size_t>(j) * nx + i] = v;
      }
  });
}

int gwm_terrain_agnesi_ridge(int nx, int ny, double dx, double xc, double a, double h0, float* out) {
  return guard([&] {
    require(nx > 0 && ny > 0 && out && a > 0, "bad arguments");
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const double x = (static_cast<double>(i) + 0.5) * dx - xc;
        out[static_cast<size_t>(j) * nx + i] = static_cast<float>(h0 / (1.0 + (x / a) * (x / a)));
      }
  });
}

// dz_k = dz_b + (dz_t - dz_b) s_k^p with s_k = k/(nz-1), p chosen by
// bisection so that the column sums to z_top exactly (rescaled at the end).
int gwm_stretched_levels(int nz, double dzb, double dzt, double ztop, double* out) {
  return guard([&] {
    require(nz >= 2 && dzb > 0 && dzt > 0 && ztop > 0 && out, "bad arguments");
    auto total = [&](double p) {
      double s = 0.0;
      for (int k = 0; k < nz; ++k) s += dzb + (dzt - dzb) * std::pow(static_cast<double>(k) / (nz - 1), p);
      return s;
    };
    double lo = 1e-3, hi = 1e3, p = 1.0;
    const bool grows = dzt > dzb;
    for (int it = 0; it < 200; ++it) {
      p = std::sqrt(lo * hi);
      const double t = total(p);
      // larger p -> smaller total when dz grows with height
      if ((t > ztop) == grows) lo = p;
      else hi = p;
    }
    double z = 0.0;
    out[0] = 0.0;
    for (int k = 0; k < nz; ++k) {
      z += dzb + (dzt - dzb) * std::pow(static_cast<double>(k) / (nz - 1), p);
      out[k + 1] = z;
    }
    const double scale = ztop / z;
    for (int k = 0; k <= nz; ++k) out[k] *= scale;
    out[nz] = ztop;
  });
}

int gwm_state_create(gwm_ctx* ctx, const gwm_grid* grid, int ntr, gwm_state** out) {
  return guard([&] {
    require(ctx && grid && out, "null argument");
    require(ntr >= 0 && ntr <= gwm::kMaxTracers, "tracer count out of range");
    require(!ctx->closing, "context already destroyed");
    auto S = std::make_unique<gwm_state>();
    S->ctx = ctx;
    ctx->refs += 1;
    S->grid = grid;
    S->ntr = ntr;
    S->ensure_sets(1);
    for (int t = 0; t < ntr; ++t) S->tr.push_back(dalloc(S->g().cell_total() * ctx->esize()));
    S->precip = dalloc(static_cast<size_t>(S->g().nx) * S->g().ny * ctx->esize());
    for (int f = 0; f < 5; ++f)
      HIPCHK(hipMemsetAsync(S->set[0][f], 0, S->fsize(f) * ctx->esize(), ctx->stream));
    for (void* p : S->tr)
      HIPCHK(hipMemsetAsync(p, 0, S->g().cell_total() * ctx->esize(), ctx->stream));
    HIPCHK(hipMemsetAsync(S->precip, 0, static_cast<size_t>(S->g().nx) * S->g().ny * ctx->esize(),
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    S->faces_synced = false;
    *out = S.release();
  });
}

int gwm_state_create_perturbation(gwm_ctx* ctx, const gwm_grid* grid, int ntr,
                                  gwm_state** out) {
  return guard([&] {
    require(ctx && grid && out, "null argument");
    require(ctx->prec == GWM_F32, "the perturbation form is an FP32 state form");
    const bool terrain = grid->terrain_nonzero;
    require(!terrain || grid->terrain_active,
            "the perturbation form on terrain needs terrain above 1e-6 m somewhere");
    require(terrain || (grid->geo.per_x && grid->geo.per_y),
            "the perturbation form on flat terrain needs a doubly periodic domain");
    require(!terrain || ntr == 0,
            "the perturbation form on terrain carries the dry state only (no tracers)");
    gwm_state* s = nullptr;
    const int rc = gwm_state_create(ctx, grid, ntr, &s);
    if (rc != GWM_OK) throw Error(rc, gwm_last_error());
    s->pert = true;
    s->pert3 = terrain;
    // Until a state is loaded, the departures are zero against a nominal
    // base (rho_b = 1, (rho theta)_b = 300).
    const size_t nb = terrain ? grid->geo.cell_total() : static_cast<size_t>(grid->geo.nz);
    s->base_rho64.assign(nb, 1.0);
    s->base_rt64.assign(nb, 300.0);
    try {
      push_base(s);
    } catch (...) {
      delete s;
      throw;
    }
    *out = s;
  });
}

int gwm_state_is_perturbation(const gwm_state* s) { return s && s->pert ? 1 : 0; }

int gwm_state_destroy(gwm_state* s) {
  return guard([&] {
    if (s) (void)hipStreamSynchronize(s->ctx->stream);
    delete s;
  });
}

long gwm_state_field_size(const gwm_state* s, int f) {
  if (!s) return -1;
  if (f >= GWM_TRACER0 && f - GWM_TRACER0 >= s->ntr) return -1;
  return static_cast<long>(s->fsize(f));
}

int gwm_state_upload(gwm_state* s, int f, const void* host) {
  return guard([&] {
    S_(s);
    require(host, "null host pointer");
    void* d = s->field(f);
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
    if (s->pert && (f == GWM_RHO || f == GWM_RHO_THETA)) {
      // full values in; the base of the field becomes their level means
      // (on terrain: the values themselves)
      gwm_state* o = s->view ? s->owner : s;
      const float* h = static_cast<const float*>(host);
      if (!s->view) (f == GWM_RHO ? o->base_rho64 : o->base_rt64) = field_base(s, h);
      push_base(o);
      upload_departure(s, d, h, f == GWM_RHO ? o->base_rho64 : o->base_rt64);
      return;
    }
    scopy(s->ctx->stream, d, host, s->fsize(f) * s->es(), hipMemcpyHostToDevice);
    if (f == GWM_MOM_U || f == GWM_MOM_V) s->faces_synced = false;
  });
}

int gwm_state_download(gwm_state* s, int f, void* host) {
  return guard([&] {
    S_(s);
    require(host, "null host pointer");
    void* d = s->field(f);
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
    scopy(s->ctx->stream, host, d, s->fsize(f) * s->es(), hipMemcpyDeviceToHost);
    if (s->pert && (f == GWM_RHO || f == GWM_RHO_THETA)) {
      // departures -> full values (base added in FP64, rounded once)
      const gwm_state* o = s->view ? s->owner : s;
      const auto& b = f == GWM_RHO ? o->base_rho64 : o->base_rt64;
      float* h = static_cast<float*>(host);
      for (size_t a = 0; a < s->fsize(f); ++a)
        h[a] = static_cast<float>(static_cast<double>(h[a]) + b[o->bidx(a)]);
    }
  });
}

int gwm_state_download_f64(gwm_state* s, int f, double* host) {
  return guard([&] {
    S_(s);
    require(host, "null host pointer");
    void* d = s->field(f);
    const size_t n = s->fsize(f);
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
    if (s->es() == 8) {
      scopy(s->ctx->stream, host, d, n * sizeof(double), hipMemcpyDeviceToHost);
      return;
    }
    std::vector<float> h(n);
    scopy(s->ctx->stream, h.data(), d, n * sizeof(float), hipMemcpyDeviceToHost);
    const bool dep = s->pert && (f == GWM_RHO || f == GWM_RHO_THETA);
    const gwm_state* o = s->view ? s->owner : s;
    for (size_t a = 0; a < n; ++a)
      host[a] = static_cast<double>(h[a]) +
                (dep ? (f == GWM_RHO ? o->base_rho64 : o->base_rt64)[o->bidx(a)] : 0.0);
  });
}

int gwm_state_fill(gwm_state* s, int f, double v) {
  return guard([&] {
    S_(s);
    void* d = s->field(f);
    if (s->pert && (f == GWM_RHO || f == GWM_RHO_THETA)) {
      const std::vector<float> h(s->fsize(f), static_cast<float>(v));
      const int rc = gwm_state_upload(s, f, h.data());
      if (rc != GWM_OK) throw Error(rc, gwm_last_error());
      return;
    }
    dispatch(s->ctx, [&](auto r) {
      using R = decltype(r);
      Timed t(s->ctx, "fill");
      gwm::launch_fill<R>(s->ctx->stream, static_cast<R*>(d), s->fsize(f), static_cast<R>(v));
      return 0;
    });
    if (f == GWM_MOM_U || f == GWM_MOM_V) s->faces_synced = false;
  });
}

int gwm_state_copy(gwm_state* dst, const gwm_state* src) {
  return guard([&] {
    require(dst && src && dst->grid == src->grid && dst->ntr == src->ntr, "state mismatch");
    require(dst->pert == src->pert, "state form mismatch (full / perturbation)");
    gwm_state* s = const_cast<gwm_state*>(src);
    if (src->pert) {
      dst->base_rho64 = src->base_rho64;
      dst->base_rt64 = src->base_rt64;
      push_base(dst);
    }
    for (int f = 0; f < 5; ++f)
      HIPCHK(hipMemcpyAsync(dst->field(f), s->field(f), s->fsize(f) * s->es(),
                            hipMemcpyDeviceToDevice, dst->ctx->stream));
    for (int t = 0; t < s->ntr; ++t)
      HIPCHK(hipMemcpyAsync(dst->field(GWM_TRACER0 + t), s->field(GWM_TRACER0 + t),
                            s->fsize(GWM_RHO) * s->es(), hipMemcpyDeviceToDevice, dst->ctx->stream));
    HIPCHK(hipMemcpyAsync(dst->precip, s->precip, s->fsize(GWM_PRECIP) * s->es(),
                          hipMemcpyDeviceToDevice, dst->ctx->stream));
    HIPCHK(hipStreamSynchronize(dst->ctx->stream));
    dst->faces_synced = s->faces_synced;
  });
}

int gwm_state_capture_boundary(gwm_state* s) {
  return guard([&] {
    S_(s);
    require(!s->view, "cannot capture on a stage view");
    for (int f = 0; f < 5; ++f) {
      if (!s->bref[f]) s->bref[f] = s->alloc_set_field(f);
      HIPCHK(hipMemcpyAsync(s->bref[f], s->field(f), s->fsize(f) * s->es(), hipMemcpyDeviceToDevice,
                            s->ctx->stream));
    }
    while (static_cast<int>(s->bref_tr.size()) < s->ntr)
      s->bref_tr.push_back(dalloc(s->g().cell_total() * s->es()));
    for (int t = 0; t < s->ntr; ++t)
      HIPCHK(hipMemcpyAsync(s->bref_tr[t], s->tr[t], s->g().cell_total() * s->es(),
                            hipMemcpyDeviceToDevice, s->ctx->stream));
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
    s->have_boundary = true;
  });
}

int gwm_state_set_boundary(gwm_state* s, int f, const void* host) {
  return guard([&] {
    S_(s);
    require(host, "null host pointer");
    gwm_state* o = s->view ? s->owner : s;
    require(o->have_boundary, "boundary reference not captured (gwm_state_capture_boundary)");
    void* d = nullptr;
    if (f >= GWM_RHO && f <= GWM_MOM_W) {
      d = o->bref[f];
    } else {
      const int t = f - GWM_TRACER0;
      require(t >= 0 && t < o->ntr && t < static_cast<int>(o->bref_tr.size()), "bad field id");
      d = o->bref_tr[static_cast<size_t>(t)];
    }
    HIPCHK(hipStreamSynchronize(o->ctx->stream));
    if (o->pert && (f == GWM_RHO || f == GWM_RHO_THETA)) {
      // full values in, stored as departures from the state's bases
      upload_departure(o, d, static_cast<const float*>(host),
                       f == GWM_RHO ? o->base_rho64 : o->base_rt64);
      return;
    }
    scopy(o->ctx->stream, d, host, o->fsize(f) * o->es(), hipMemcpyHostToDevice);
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Restart dump.  The reference fixes only the artifact contract of a native
// checkpoint (tools/casebuilder/prepare_case.py:376-392: schema
// "gwm-next-checkpoint-stub/v1", restart family "gwm-native-checkpoint",
// cold-start groups); this is that checkpoint.  Layout: a fixed header
// (magic, version, shape, precision, form, tracer count, section count),
// a table of sections (name, byte offset, byte count, FNV-1a 64 checksum),
// then the sections: the canonical fields in storage layout (halos
// included), the tracers, the surface bucket, the perturbation form's FP64
// bases and the captured boundary reference.  A restore is bit-exact.
// ---------------------------------------------------------------------------
namespace {

constexpr char kCkptMagic[8] = {'G', 'W', 'M', 'C', 'K', 'P', 'T', '1'};
constexpr std::uint32_t kCkptVersion = 1;

struct CkptHeader {
  char magic[8];
  std::uint32_t version, precision, nx, ny, nz, halo, per_x, per_y, pert, ntr, boundary, nsec;
  double dx, dy;
};
struct CkptSection {
  char name[24];
  std::uint64_t offset, bytes, fnv;
};

std::uint64_t fnv1a(const unsigned char* p, size_t n, std::uint64_t h = 1469598103934665603ull) {
  for (size_t a = 0; a < n; ++a) {
    h ^= p[a];
    h *= 1099511628211ull;
  }
  return h;
}

// The sections of a state, in file order: name, device pointer or host
// bytes, byte count.
struct CkptItem {
  std::string name;
  void* dev = nullptr;
  std::vector<unsigned char>* host = nullptr;
  size_t bytes = 0;
};

std::vector<CkptItem> ckpt_items(gwm_state* S, std::vector<unsigned char>& bases) {
  std::vector<CkptItem> v;
  const char* fname[5] = {"rho", "rho_theta", "mom_u", "mom_v", "mom_w"};
  for (int f = 0; f < 5; ++f) v.push_back({fname[f], S->field(f), nullptr, S->fsize(f) * S->es()});
  for (int t = 0; t < S->ntr; ++t)
    v.push_back({"tracer" + std::to_string(t), S->tr[static_cast<size_t>(t)], nullptr,
                 S->g().cell_total() * S->es()});
  if (S->ntr > 0)
    v.push_back({"precip", S->precip, nullptr, static_cast<size_t>(S->g().nx) * S->g().ny * S->es()});
  if (S->pert) {
    const size_t nb = S->base_rho64.size();  // nz, or the cell storage on terrain
    bases.resize(2 * nb * sizeof(double));
    v.push_back({"bases_fp64", nullptr, &bases, bases.size()});
  }
  if (S->have_boundary) {
    for (int f = 0; f < 5; ++f)
      v.push_back({std::string("boundary_") + fname[f], S->bref[f], nullptr, S->fsize(f) * S->es()});
    for (int t = 0; t < S->ntr && t < static_cast<int>(S->bref_tr.size()); ++t)
      v.push_back({"boundary_tracer" + std::to_string(t), S->bref_tr[static_cast<size_t>(t)], nullptr,
                   S->g().cell_total() * S->es()});
  }
  return v;
}

std::string json_escape(const std::string& in) {
  std::string o;
  for (char c : in) {
    if (c == '"' || c == '\\') o += '\\';
    if (static_cast<unsigned char>(c) < 0x20) continue;
    o += c;
  }
  return o;
}

void checkpoint_write(gwm_state* S, const std::string& path, const char* meta) {
  require(!S->view, "checkpoint of a stage view");
  const Geo& g = S->g();
  std::vector<unsigned char> bases;
  auto items = ckpt_items(S, bases);
  if (S->pert) {
    const size_t nb = S->base_rho64.size();
    std::memcpy(bases.data(), S->base_rho64.data(), nb * sizeof(double));
    std::memcpy(bases.data() + nb * sizeof(double), S->base_rt64.data(), nb * sizeof(double));
  }
  CkptHeader h{};
  std::memcpy(h.magic, kCkptMagic, 8);
  h.version = kCkptVersion;
  h.precision = static_cast<std::uint32_t>(S->ctx->prec);
  h.nx = g.nx;
  h.ny = g.ny;
  h.nz = g.nz;
  h.halo = g.h;
  h.per_x = g.per_x;
  h.per_y = g.per_y;
  h.pert = S->pert ? 1 : 0;
  h.ntr = static_cast<std::uint32_t>(S->ntr);
  h.boundary = S->have_boundary ? 1 : 0;
  h.nsec = static_cast<std::uint32_t>(items.size());
  h.dx = g.dx;
  h.dy = g.dy;
  std::vector<CkptSection> sec(items.size());
  std::uint64_t off = sizeof(CkptHeader) + sec.size() * sizeof(CkptSection);
  for (size_t n = 0; n < items.size(); ++n) {
    std::memset(sec[n].name, 0, sizeof sec[n].name);
    std::strncpy(sec[n].name, items[n].name.c_str(), sizeof sec[n].name - 1);
    off = (off + 255) & ~static_cast<std::uint64_t>(255);
    sec[n].offset = off;
    sec[n].bytes = items[n].bytes;
    off += items[n].bytes;
  }
  HIPCHK(hipStreamSynchronize(S->ctx->stream));
  const std::string tmp = path + ".partial";
  std::FILE* fp = std::fopen(tmp.c_str(), "wb");
  require(fp != nullptr, "checkpoint: cannot open " + tmp + " for writing");
  struct Closer {
    std::FILE*& f;
    ~Closer() {
      if (f) std::fclose(f);
    }
  } closer{fp};
  auto put = [&](const void* p, size_t n, std::uint64_t at) {
    require(std::fseek(fp, static_cast<long>(at), SEEK_SET) == 0 && std::fwrite(p, 1, n, fp) == n,
            "checkpoint: write to " + tmp + " failed");
  };
  std::vector<unsigned char> buf;
  for (size_t n = 0; n < items.size(); ++n) {
    const auto& it = items[n];
    const unsigned char* src;
    if (it.host) {
      src = it.host->data();
    } else {
      buf.resize(it.bytes);
      scopy(S->ctx->stream, buf.data(), it.dev, it.bytes, hipMemcpyDeviceToHost);
      src = buf.data();
    }
    sec[n].fnv = fnv1a(src, it.bytes);
    put(src, it.bytes, sec[n].offset);
  }
  put(&h, sizeof h, 0);
  put(sec.data(), sec.size() * sizeof(CkptSection), sizeof h);
  require(std::fclose(fp) == 0, "checkpoint: closing " + tmp + " failed");
  fp = nullptr;
  require(std::rename(tmp.c_str(), path.c_str()) == 0, "checkpoint: cannot move " + tmp + " to " + path);
  // Manifest beside the dump: the reference's checkpoint-stub contract
  // (read by tools/verify) plus what the dump holds.
  std::FILE* mf = std::fopen((path + ".json").c_str(), "wb");
  require(mf != nullptr, "checkpoint: cannot write the manifest " + path + ".json");
  std::string m = "{\n  \"schema_version\": \"gwm-next-checkpoint-stub/v1\",\n";
  m += "  \"restart_family\": \"gwm-native-checkpoint\",\n";
  m += "  \"cold_start_contract\": {\n    \"analysis_state_schema\": \"gwm-next-analysis-state/v1\",\n";
  m += "    \"boundary_cache_schema\": \"gwm-next-boundary-cache/v1\",\n";
  m += "    \"required_groups\": [\"atmosphere\", \"surface\", \"static_surface\"]\n  },\n";
  const std::string base = path.substr(path.find_last_of('/') == std::string::npos ? 0 : path.find_last_of('/') + 1);
  char line[512];
  std::snprintf(line, sizeof line,
                "  \"checkpoint\": {\n    \"file\": \"%s\",\n    \"format\": \"gwm-next-checkpoint/v1\",\n"
                "    \"grid\": {\"nx\": %d, \"ny\": %d, \"nz\": %d, \"halo\": %d, \"dx\": %.17g, "
                "\"dy\": %.17g, \"periodic_x\": %s, \"periodic_y\": %s},\n"
                "    \"precision\": \"%s\",\n    \"perturbation_form\": %s,\n"
                "    \"n_tracers\": %d,\n    \"boundary_reference\": %s,\n    \"sections\": %zu\n  }",
                json_escape(base).c_str(), g.nx, g.ny, g.nz, g.h, g.dx, g.dy, g.per_x ? "true" : "false",
                g.per_y ? "true" : "false", S->ctx->prec == GWM_F64 ? "fp64" : "fp32",
                S->pert ? "true" : "false", S->ntr, S->have_boundary ? "true" : "false", items.size());
  m += line;
  if (meta != nullptr && *meta) m += std::string(",\n  \"metadata\": ") + meta;
  m += "\n}\n";
  const bool ok = std::fwrite(m.data(), 1, m.size(), mf) == m.size();
  const bool closed = std::fclose(mf) == 0;
  require(ok && closed, "checkpoint: writing the manifest " + path + ".json failed");
}

void checkpoint_read(gwm_state* S, const std::string& path) {
  require(!S->view, "restore into a stage view");
  const Geo& g = S->g();
  std::FILE* fp = std::fopen(path.c_str(), "rb");
  require(fp != nullptr, "checkpoint: cannot open " + path);
  struct Closer {
    std::FILE* f;
    ~Closer() { std::fclose(f); }
  } closer{fp};
  auto get = [&](void* p, size_t n, std::uint64_t at) {
    require(std::fseek(fp, static_cast<long>(at), SEEK_SET) == 0 && std::fread(p, 1, n, fp) == n,
            "checkpoint: " + path + " is truncated");
  };
  CkptHeader h{};
  get(&h, sizeof h, 0);
  require(std::memcmp(h.magic, kCkptMagic, 8) == 0, "checkpoint: " + path + " is not a gwm checkpoint");
  require(h.version == kCkptVersion, "checkpoint: unsupported version " + std::to_string(h.version));
  auto same = [&](bool c, const std::string& what) {
    require(c, "checkpoint: " + path + " holds a different " + what + " than the state");
  };
  same(h.nx == static_cast<std::uint32_t>(g.nx) && h.ny == static_cast<std::uint32_t>(g.ny) &&
           h.nz == static_cast<std::uint32_t>(g.nz) && h.halo == static_cast<std::uint32_t>(g.h),
       "grid shape (" + std::to_string(h.nx) + "x" + std::to_string(h.ny) + "x" + std::to_string(h.nz) +
           ", halo " + std::to_string(h.halo) + ")");
  same(h.per_x == static_cast<std::uint32_t>(g.per_x) && h.per_y == static_cast<std::uint32_t>(g.per_y),
       "lateral periodicity");
  same(h.precision == static_cast<std::uint32_t>(S->ctx->prec), "precision");
  same(h.pert == (S->pert ? 1u : 0u), "storage form (perturbation or full)");
  same(h.ntr == static_cast<std::uint32_t>(S->ntr), "tracer count");
  if (h.boundary && !S->have_boundary) {  // the reference arrives with the dump
    for (int f = 0; f < 5; ++f)
      if (!S->bref[f]) S->bref[f] = S->alloc_set_field(f);
    while (static_cast<int>(S->bref_tr.size()) < S->ntr)
      S->bref_tr.push_back(dalloc(g.cell_total() * S->es()));
    S->have_boundary = true;
  }
  std::vector<unsigned char> bases;
  const auto items = ckpt_items(S, bases);
  require(h.nsec == items.size() || (!h.boundary && S->have_boundary),
          "checkpoint: " + path + " has an unexpected section count");
  std::vector<CkptSection> sec(h.nsec);
  get(sec.data(), sec.size() * sizeof(CkptSection), sizeof h);
  HIPCHK(hipStreamSynchronize(S->ctx->stream));
  std::vector<unsigned char> buf;
  for (size_t n = 0; n < sec.size(); ++n) {
    const auto& it = items[n];
    require(std::string(sec[n].name, strnlen(sec[n].name, sizeof sec[n].name)) == it.name &&
                sec[n].bytes == it.bytes,
            "checkpoint: section " + std::to_string(n) + " of " + path + " does not match " + it.name);
    buf.resize(it.bytes);
    get(buf.data(), it.bytes, sec[n].offset);
    require(fnv1a(buf.data(), it.bytes) == sec[n].fnv,
            "checkpoint: section " + it.name + " of " + path + " fails its checksum");
    if (it.host)
      *it.host = buf;
    else
      scopy(S->ctx->stream, it.dev, buf.data(), it.bytes, hipMemcpyHostToDevice);
  }
  if (S->pert) {
    const size_t nb = S->base_rho64.size();
    S->base_rho64.assign(reinterpret_cast<const double*>(bases.data()),
                         reinterpret_cast<const double*>(bases.data()) + nb);
    S->base_rt64.assign(reinterpret_cast<const double*>(bases.data()) + nb,
                        reinterpret_cast<const double*>(bases.data()) + 2 * nb);
    push_base(S);
  }
  S->faces_synced = false;
  S->theta_parts_rho = nullptr;
}

}  // namespace

extern "C" {

int gwm_state_checkpoint_write(gwm_state* s, const char* path, const char* metadata_json) {
  return guard([&] {
    S_(s);
    require(path != nullptr && *path, "null path");
    checkpoint_write(s, path, metadata_json);
  });
}

int gwm_state_checkpoint_read(gwm_state* s, const char* path) {
  return guard([&] {
    S_(s);
    require(path != nullptr && *path, "null path");
    checkpoint_read(s, path);
  });
}

int gwm_state_set_tracer_strip_mask(gwm_state* s, unsigned mask) {
  return guard([&] {
    S_(s);
    (s->view ? s->owner : s)->tracer_strip_mask = mask;
  });
}

int gwm_halo_fill(gwm_state* s, int f) {
  return guard([&] {
    S_(s);
    require(f != GWM_PRECIP, "the surface bucket has no halo");
    void* d = s->field(f);
    dispatch(s->ctx, [&](auto r) {
      using R = decltype(r);
      Timed t(s->ctx, "halo_fill");
      gwm::launch_halo_fill<R>(s->ctx->stream, s->g(), orient_of(f), static_cast<R*>(d));
      return 0;
    });
  });
}

int gwm_sync_interfaces(gwm_state* s, int f) {
  return guard([&] {
    S_(s);
    const int o = orient_of(f);
    if (o != 1 && o != 2) return;
    if ((o == 1 && !s->g().per_x) || (o == 2 && !s->g().per_y)) return;
    void* d = s->field(f);
    dispatch(s->ctx, [&](auto r) {
      using R = decltype(r);
      Timed t(s->ctx, "sync_interfaces");
      gwm::launch_sync_interfaces<R>(s->ctx->stream, s->g(), o, static_cast<R*>(d));
      return 0;
    });
  });
}

int gwm_halo_exchange_layout(gwm_ctx* ctx, int nr, const int* desc13, int px, int py, int orient,
                             void* const* fields) {
  return guard([&] {
    require(ctx && desc13 && fields && nr > 0, "null argument");
    require(orient >= 0 && orient <= 3, "bad orientation");
    // neighbours (cartesian_topology.cpp:57-74)
    std::vector<int> d(static_cast<size_t>(nr) * 17);
    auto rank_at = [&](int cx, int cy) {
      const int rx = desc13[3], ry = desc13[4];
      if (px) cx = ((cx % rx) + rx) % rx;
      else if (cx < 0 || cx >= rx) return -1;
      if (py) cy = ((cy % ry) + ry) % ry;
      else if (cy < 0 || cy >= ry) return -1;
      for (int n = 0; n < nr; ++n)
        if (desc13[n * 13 + 1] == cx && desc13[n * 13 + 2] == cy) return desc13[n * 13 + 0];
      return -1;
    };
    int maxe = 0;
    for (int n = 0; n < nr; ++n) {
      const int* D = desc13 + n * 13;
      require(D[0] == n, "ranks must be listed in order");
      std::copy(D, D + 13, d.begin() + n * 17);
      d[n * 17 + 13] = rank_at(D[1] - 1, D[2]);
      d[n * 17 + 14] = rank_at(D[1] + 1, D[2]);
      d[n * 17 + 15] = rank_at(D[1], D[2] - 1);
      d[n * 17 + 16] = rank_at(D[1], D[2] + 1);
      const int h = D[8], nxl = D[10] - D[9], nyl = D[12] - D[11];
      const int fx = nxl + (orient == 1), fy = nyl + (orient == 2), fz = D[7] + (orient == 3);
      maxe = std::max(maxe, std::max(fz * fy * h, fz * h * (fx + 2 * h)));
    }
    int* dd = static_cast<int*>(dalloc(d.size() * sizeof(int)));
    void** df = static_cast<void**>(dalloc(nr * sizeof(void*)));
    HIPCHK(hipMemcpyAsync(dd, d.data(), d.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(df, fields, nr * sizeof(void*), hipMemcpyHostToDevice, ctx->stream));
    dispatch(ctx, [&](auto r) {
      using R = decltype(r);
      for (int phase = 0; phase < 2; ++phase) {
        Timed t(ctx, "exchange_layout");
        gwm::launch_exchange_layout<R>(ctx->stream, phase, orient, nr, dd,
                                       reinterpret_cast<R* const*>(df), maxe);
      }
      return 0;
    });
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipFree(dd));
    HIPCHK(hipFree(df));
  });
}

int gwm_device_alloc(gwm_ctx* ctx, size_t bytes, void** out) {
  return guard([&] {
    require(ctx && out, "null argument");
    *out = dalloc(bytes);
  });
}
int gwm_device_free(gwm_ctx* ctx, void* p) {
  return guard([&] {
    require(ctx, "null context");
    if (p) HIPCHK(hipFree(p));
  });
}
// The context stream is non-blocking (it does not order with the null
// stream), so the copies run on it: after every kernel the context queued
// before them, complete when the call returns.
int gwm_memcpy_h2d(gwm_ctx* ctx, void* dst, const void* src, size_t bytes) {
  return guard([&] {
    require(ctx, "null context");
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
  });
}
int gwm_memcpy_d2h(gwm_ctx* ctx, void* dst, const void* src, size_t bytes) {
  return guard([&] {
    require(ctx, "null context");
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
  });
}

int gwm_pressure(gwm_state* s) {
  return guard([&] {
    S_(s);
    require(!s->pert, "gwm_pressure: full-form states only (use gwm_state_download)");
    void* p = s->field(GWM_PRESSURE);
    void** fs = s->cur_set();
    dispatch(s->ctx, [&](auto r) {
      using R = decltype(r);
      Timed t(s->ctx, "pressure");
      gwm::launch_pressure<R>(s->ctx->stream, s->g(), static_cast<const R*>(fs[0]),
                              static_cast<const R*>(fs[1]), static_cast<R*>(p));
      return 0;
    });
  });
}

int gwm_slow_tendencies(gwm_state* s, const gwm_step_config* c) {
  return guard([&] {
    S_(s);
    require(c, "null config");
    require(!s->view, "operator not available on a stage view");
    dispatch(s->ctx, [&](auto r) {
      slow_tendencies<decltype(r)>(s, *c);
      return 0;
    });
  });
}

int gwm_fast_modes(gwm_state* s, const gwm_step_config* c) {
  return guard([&] {
    S_(s);
    require(c, "null config");
    require(!s->view, "operator not available on a stage view");
    dispatch(s->ctx, [&](auto r) {
      fast_modes<decltype(r)>(s, *c);
      return 0;
    });
  });
}

int gwm_ssprk3_step(gwm_state* s, const gwm_step_config* c, const gwm_boundary_config* b) {
  return guard([&] {
    S_(s);
    require(c, "null config");
    require(!s->view, "operator not available on a stage view");
    dispatch(s->ctx, [&](auto r) {
      ssprk3<decltype(r)>(s, *c, b);
      return 0;
    });
  });
}

int gwm_ssprk3_step_custom_fast(gwm_state* s, const gwm_step_config* c,
                                const gwm_boundary_config* b, gwm_boundary_fn fast_fn,
                                void* fast_user) {
  return guard([&] {
    S_(s);
    require(c && fast_fn, "null argument");
    require(!s->view, "operator not available on a stage view");
    dispatch(s->ctx, [&](auto r) {
      ssprk3<decltype(r)>(s, *c, b, fast_fn, fast_user);
      return 0;
    });
  });
}

int gwm_tracers_ssprk3(gwm_state* s, const gwm_step_config* c, int bmode) {
  return guard([&] {
    S_(s);
    require(c, "null config");
    require(!s->view, "operator not available on a stage view");
    dispatch(s->ctx, [&](auto r) {
      tracers_ssprk3<decltype(r)>(s, *c, bmode);
      return 0;
    });
  });
}

int gwm_tracers_ssprk3_callback(gwm_state* s, const gwm_step_config* c, int bmode,
                                gwm_boundary_fn fn, void* user) {
  return guard([&] {
    S_(s);
    require(c, "null config");
    require(!s->view, "operator not available on a stage view");
    require(bmode == GWM_BOUNDARY_NONE || bmode == GWM_BOUNDARY_STRIPS,
            "tracer boundary mode must be GWM_BOUNDARY_NONE or GWM_BOUNDARY_STRIPS");
    dispatch(s->ctx, [&](auto r) {
      tracers_ssprk3<decltype(r)>(s, *c, bmode, fn, user);
      return 0;
    });
  });
}

int gwm_warm_rain(gwm_state* s, const gwm_warm_rain_config* c, int metric) {
  return guard([&] {
    S_(s);
    require(c, "null config");
    dispatch(s->ctx, [&](auto r) {
      warm_rain<decltype(r)>(s, *c, metric);
      return 0;
    });
  });
}

int gwm_apply_boundary(gwm_state* s, int mode, double width) {
  return guard([&] {
    S_(s);
    gwm_state* owner = s->view ? s->owner : s;
    require(owner->have_boundary, "boundary reference not captured");
    dispatch(s->ctx, [&](auto r) {
      using R = decltype(r);
      void** fs = s->cur_set();
      gwm_state* o = owner;
      const Geo& g = s->g();
      if (mode == GWM_BOUNDARY_NONE) return 0;
      for (int f = 0; f < 5; ++f) {
        Timed t(s->ctx, "boundary");
        gwm::launch_boundary<R>(s->ctx->stream, g, orient_of(f), static_cast<R*>(fs[f]),
                                static_cast<const R*>(o->bref[f]), mode, width);
      }
      if (mode == GWM_BOUNDARY_STRIPS)
        for (int t = 0; t < o->ntr; ++t) {
          Timed tt(s->ctx, "boundary");
          gwm::launch_boundary<R>(s->ctx->stream, g, 0, static_cast<R*>(o->tr[t]),
                                  static_cast<const R*>(o->bref_tr[t]), mode, width);
        }
      return 0;
    });
    owner->faces_synced = false;
  });
}

int gwm_moist_step(gwm_state* s, const gwm_step_config* c, const gwm_warm_rain_config* wr,
                   const gwm_boundary_config* b) {
  return guard([&] {
    S_(s);
    require(c && wr, "null config");
    dispatch(s->ctx, [&](auto r) {
      using R = decltype(r);
      ssprk3<R>(s, *c, b);
      tracers_ssprk3<R>(s, *c, b ? (b->mode == GWM_BOUNDARY_STRIPS ? GWM_BOUNDARY_STRIPS : GWM_BOUNDARY_NONE) : GWM_BOUNDARY_NONE);
      warm_rain<R>(s, *wr, 1);
      return 0;
    });
  });
}

int gwm_budget_sums(gwm_state* s, double* out) {
  return guard([&] {
    S_(s);
    require(out, "null output");
    dispatch(s->ctx, [&](auto r) {
      budget<decltype(r)>(s, out);
      return 0;
    });
  });
}

int gwm_plan_view(gwm_state* s, int slice_k, int steps, double dt, int with_precip, double* out,
                  int* nfields) {
  return guard([&] {
    S_(s);
    require(out && nfields, "null output");
    require(steps >= 0, "steps must be >= 0");
    require(!(s->view ? s->owner : s)->pert3,
            "gwm_plan_view: the perturbation form on terrain is not supported");
    *nfields = dispatch(s->ctx, [&](auto r) {
      return plan_view<decltype(r)>(s, slice_k, steps, dt, with_precip, out);
    });
  });
}

int gwm_budget_sums_noalloc(gwm_state* s, double* out) {
  return guard([&] {
    S_(s);
    require(out, "null output");
    require(s->ctx->noalloc, "the context has no budget scratch");
    // host staging on the stack: no allocation on either side
    double host[gwm::kNoAllocScratchDoubles];
    ReduceBufs rb;
    rb.dev = s->ctx->noalloc;
    rb.host = host;
    dispatch(s->ctx, [&](auto r) {
      budget<decltype(r)>(s, out, rb);
      return 0;
    });
  });
}

int gwm_init_constant(gwm_state* s, double rho, double th, double u, double v, double w) {
  return guard([&] {
    S_(s);
    init_dispatch(s, [&](auto r) {
      using R = decltype(r);
      HostState<R> H(s);
      fill_constant<R>(s, H, static_cast<R>(rho), static_cast<R>(th), static_cast<R>(u),
                       static_cast<R>(v), static_cast<R>(w));
      if (s->pert) {  // tracers are left as they are
        upload_pert_fields(s, H.rho.data(), H.rt.data(), H.u.data(), H.v.data(), H.w.data(),
                           std::vector<std::vector<R>>{});
        return 0;
      }
      H.q.clear();
      std::vector<R>* f[5] = {&H.rho, &H.rt, &H.u, &H.v, &H.w};
      for (int a = 0; a < 5; ++a)
        scopy(s->ctx->stream, s->cur_set()[a], f[a]->data(), f[a]->size() * sizeof(R),
              hipMemcpyHostToDevice);
      s->faces_synced = false;
      return 0;
    });
  });
}

int gwm_init_hydrostatic_rest(gwm_state* s, double rho_s, double th) {
  return guard([&] {
    S_(s);
    init_dispatch(s, [&](auto r) {
      using R = decltype(r);
      HostState<R> H(s);
      init_hydro<R>(s, rho_s, th, H);
      H.upload(s);
      return 0;
    });
  });
}

int gwm_init_bubble(gwm_state* s, const double* p, int sign) {
  return guard([&] {
    S_(s);
    require(p, "null parameters");
    init_dispatch(s, [&](auto r) {
      init_bubble<decltype(r)>(s, p, sign);
      return 0;
    });
  });
}

int gwm_init_mountain_wave(gwm_state* s, double rho_s, double th, double u, double v, double w) {
  return guard([&] {
    S_(s);
    init_dispatch(s, [&](auto r) {
      init_mountain<decltype(r)>(s, rho_s, th, u, v, w);
      return 0;
    });
  });
}

int gwm_init_constant_n(gwm_state* s, double th_s, double N, double p_s, double u, double v) {
  return guard([&] {
    S_(s);
    require(N > 0 && th_s > 0 && p_s > 0, "bad sounding parameters");
    init_dispatch(s, [&](auto r) {
      init_constant_n<decltype(r)>(s, th_s, N, p_s, u, v);
      return 0;
    });
  });
}

int gwm_init_convection(gwm_state* s, double p_s, unsigned long long seed, double amp, double ztop) {
  return guard([&] {
    S_(s);
    init_dispatch(s, [&](auto r) {
      init_convection<decltype(r)>(s, p_s, seed, amp, ztop);
      return 0;
    });
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Warm-rain surrogate (no reference counterpart): parameters, training-data
// sampling, FP32 training, BF16 matrix-core inference and the warm-rain step
// that uses it.
// ---------------------------------------------------------------------------
struct gwm_surrogate {
  gwm_ctx* ctx = nullptr;
  gwm::SurTrain T;
  int batch = 0;  // rows the training buffers hold
  gwm::SurNorm norm{};
  bool have_norm = false;
  long version = 0;  // bumped by set_norm (graphs bake the normalisation in)
  // cells of the last surrogate warm-rain call: [0] plain-scheme fallback
  // (outside the training box), [1] through the network
  unsigned long long* nfall = nullptr;
  double* loss_part = nullptr;
  size_t loss_n = 0;
  ~gwm_surrogate() {
    if (ctx) (void)hipStreamSynchronize(ctx->stream);
    free_batch();
    for (void* p : {static_cast<void*>(T.params), static_cast<void*>(T.grad),
                    static_cast<void*>(T.m), static_cast<void*>(T.v),
                    static_cast<void*>(nfall), static_cast<void*>(loss_part)})
      if (p) (void)hipFree(p);
    ctx_release(ctx);
  }
  void free_batch() {
    std::vector<float*> ps = {T.xn, T.wpart};
    for (auto* q : T.z) ps.push_back(q);
    for (auto* q : T.a) ps.push_back(q);
    for (auto* q : T.dz) ps.push_back(q);
    for (float* q : ps)
      if (q) (void)hipFree(q);
    T.xn = T.wpart = nullptr;
    for (auto& q : T.z) q = nullptr;
    for (auto& q : T.a) q = nullptr;
    for (auto& q : T.dz) q = nullptr;
    batch = 0;
  }
  void ensure_batch(int B) {
    if (B <= batch) return;
    free_batch();
    auto fa = [](size_t n) { return static_cast<float*>(dalloc(n * sizeof(float))); };
    T.xn = fa(static_cast<size_t>(B) * 8);
    for (int l = 0; l < 4; ++l) {
      const size_t w = l < 3 ? 64 : gwm::kSurOut;
      T.z[l] = fa(static_cast<size_t>(B) * w);
      T.dz[l] = fa(static_cast<size_t>(B) * w);
      if (l < 3) T.a[l] = fa(static_cast<size_t>(B) * 64);
    }
    T.wpart = fa(static_cast<size_t>(16) * 65 * 64);
    batch = B;
  }
  double* loss_buf(size_t n) {
    if (n > loss_n) {
      if (loss_part) HIPCHK(hipFree(loss_part));
      loss_part = static_cast<double*>(dalloc(n * sizeof(double)));
      loss_n = n;
    }
    return loss_part;
  }
};

namespace {

gwm::SurNorm to_norm(const gwm_surrogate_norm& n) {
  gwm::SurNorm s{};
  for (int f = 0; f < gwm::kSurIn; ++f) {
    require(n.x_std[f] > 0.0f, "feature std must be positive");
    require(n.box_lo[f] <= n.box_hi[f], "fallback box bounds out of order");
    s.xm[f] = n.x_mean[f];
    s.xs[f] = n.x_std[f];
    s.lo[f] = n.box_lo[f];
    s.hi[f] = n.box_hi[f];
  }
  for (int o = 0; o < gwm::kSurOut; ++o) {
    require(n.y_std[o] > 0.0f, "target std must be positive");
    s.ym[o] = n.y_mean[o];
    s.ys[o] = n.y_std[o];
  }
  return s;
}

gwm::MicroCfg<float> micro_cfg(const gwm_warm_rain_config& c) {
  return gwm::MicroCfg<float>{static_cast<float>(c.dt),
                              static_cast<float>(c.condensation_relaxation),
                              static_cast<float>(c.cloud_autoconversion_threshold),
                              static_cast<float>(c.cloud_autoconversion_rate),
                              static_cast<float>(c.rain_evaporation_rate),
                              static_cast<float>(c.accretion_rate), c.enable_latent_heating};
}

// Warm rain with the surrogate: source terms (network, or the plain scheme
// outside the box), then the fallout sweep of warm_rain (metric or nominal).
void warm_rain_surrogate(gwm_state* S, const gwm_warm_rain_config& c, gwm_surrogate* M,
                         int metric, long long* nfall_out) {
  require(S->ctx->prec == GWM_F32, "the surrogate runs on FP32 states");
  require(S->ntr >= 3, "warm rain requires the qv, qc, qr tracers");
  require(M->have_norm, "surrogate normalisation not set");
  require(M->ctx == S->ctx, "surrogate and state belong to different contexts");
  const Geo& g = S->g();
  void** ds = S->cur_set();
  float* rho = static_cast<float*>(ds[0]);
  float* rt = static_cast<float*>(ds[1]);
  float* qv = static_cast<float*>(S->field(GWM_TRACER0 + 0));
  float* qc = static_cast<float*>(S->field(GWM_TRACER0 + 1));
  float* qr = static_cast<float*>(S->field(GWM_TRACER0 + 2));
  gwm::launch_zero_counters(S->ctx->stream, M->nfall, 2);
  const float vt = static_cast<float>(c.rain_terminal_velocity);
  const float depth = static_cast<float>(c.sedimentation_layer_depth_m);
  const int fallout = metric ? (vt > 0.0f ? 2 : 0) : ((vt > 0.0f && depth > 0.0f) ? 1 : 0);
  {
    // source terms and fallout (the classify pass carries the fallout down
    // each column to its first network cell, a fallout pass finishes it)
    Timed t(S->ctx, "warm_rain_surrogate");
    gwm_state* O = S->view ? S->owner : S;
    unsigned* q = O->surrogate_queue();
    const size_t nq = static_cast<size_t>(gwm::sur_segments(g)) * 64 * g.nz;
    unsigned* segs = q + nq;
    unsigned* cols = segs + 2 * static_cast<size_t>(gwm::sur_segments(g));
    gwm::SurArgs a{g, rho, rt, qv, qc, qr, micro_cfg(c), M->T.params, M->norm, M->nfall,
                   S->base<float>(), q, segs};
    a.precip = static_cast<float*>(O->precip);
    a.idz_c = static_cast<const float*>(S->grid->idz_c);
    a.lev_idz_c = static_cast<const float*>(S->grid->lev_idz_c);
    a.fdt = static_cast<float>(c.dt);
    a.vt = vt;
    a.depth = depth;
    a.fallout = fallout;
    a.col_stop = reinterpret_cast<int*>(cols);
    a.col_inc = reinterpret_cast<float*>(cols + static_cast<size_t>(g.nx) * g.ny);
    gwm::launch_warm_rain_sur(S->ctx->stream, a, S->grid->terrain_nonzero);
  }
  if (nfall_out) {
    unsigned long long h = 0;
    HIPCHK(hipMemcpyAsync(&h, M->nfall, sizeof h, hipMemcpyDeviceToHost, S->ctx->stream));
    HIPCHK(hipStreamSynchronize(S->ctx->stream));
    *nfall_out = static_cast<long long>(h);
  }
}

}  // namespace

extern "C" {

int gwm_surrogate_create(gwm_ctx* ctx, unsigned long long seed, gwm_surrogate** out) {
  return guard([&] {
    require(ctx && out, "null argument");
    require(!ctx->closing, "context already destroyed");
    HIPCHK(hipSetDevice(ctx->device));
    auto M = std::make_unique<gwm_surrogate>();
    M->ctx = ctx;
    ctx->refs += 1;
    // Xavier-uniform weights from a seeded counter hash, zero biases.
    std::vector<float> p(gwm::kSurParams, 0.0f);
    const int wofs[4] = {gwm::kSurW1, gwm::kSurW2, gwm::kSurW3, gwm::kSurW4};
    const int fin[4] = {gwm::kSurIn, 64, 64, 64};
    const int fout[4] = {64, 64, 64, gwm::kSurOut};
    unsigned long long ctr = seed * 0x9e3779b97f4a7c15ull + 1;
    for (int l = 0; l < 4; ++l) {
      const double lim = std::sqrt(6.0 / (fin[l] + fout[l]));
      for (int q = 0; q < fin[l] * fout[l]; ++q) {
        unsigned long long z = (ctr += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        const double u = static_cast<double>(z >> 11) * (1.0 / 9007199254740992.0);
        p[wofs[l] + q] = static_cast<float>((2.0 * u - 1.0) * lim);
      }
    }
    auto fa = [](size_t n) { return static_cast<float*>(dalloc(n * sizeof(float))); };
    M->T.params = fa(gwm::kSurParams);
    M->T.grad = fa(gwm::kSurParams);
    M->T.m = fa(gwm::kSurParams);
    M->T.v = fa(gwm::kSurParams);
    scopy(M->ctx->stream, M->T.params, p.data(), p.size() * sizeof(float), hipMemcpyHostToDevice);
    HIPCHK(hipMemsetAsync(M->T.m, 0, gwm::kSurParams * sizeof(float), ctx->stream));
    HIPCHK(hipMemsetAsync(M->T.v, 0, gwm::kSurParams * sizeof(float), ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    M->nfall = static_cast<unsigned long long*>(dalloc(2 * sizeof(unsigned long long)));
    gwm::launch_zero_counters(ctx->stream, M->nfall, 2);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    *out = M.release();
  });
}

int gwm_surrogate_destroy(gwm_surrogate* m) {
  return guard([&] { delete m; });
}

int gwm_surrogate_set_norm(gwm_surrogate* m, const gwm_surrogate_norm* n) {
  return guard([&] {
    require(m && n, "null argument");
    m->norm = to_norm(*n);
    m->have_norm = true;
    m->version += 1;
  });
}

int gwm_surrogate_get_norm(gwm_surrogate* m, gwm_surrogate_norm* n) {
  return guard([&] {
    require(m && n, "null argument");
    require(m->have_norm, "surrogate normalisation not set");
    for (int f = 0; f < gwm::kSurIn; ++f) {
      n->x_mean[f] = m->norm.xm[f];
      n->x_std[f] = m->norm.xs[f];
      n->box_lo[f] = m->norm.lo[f];
      n->box_hi[f] = m->norm.hi[f];
    }
    for (int o = 0; o < gwm::kSurOut; ++o) {
      n->y_mean[o] = m->norm.ym[o];
      n->y_std[o] = m->norm.ys[o];
    }
  });
}

int gwm_surrogate_set_params(gwm_surrogate* m, const float* host) {
  return guard([&] {
    require(m && host, "null argument");
    HIPCHK(hipMemcpyAsync(m->T.params, host, gwm::kSurParams * sizeof(float),
                          hipMemcpyHostToDevice, m->ctx->stream));
    HIPCHK(hipMemsetAsync(m->T.m, 0, gwm::kSurParams * sizeof(float), m->ctx->stream));
    HIPCHK(hipMemsetAsync(m->T.v, 0, gwm::kSurParams * sizeof(float), m->ctx->stream));
    HIPCHK(hipStreamSynchronize(m->ctx->stream));
    m->T.step = 0;
  });
}

int gwm_surrogate_get_params(gwm_surrogate* m, float* host) {
  return guard([&] {
    require(m && host, "null argument");
    HIPCHK(hipMemcpyAsync(host, m->T.params, gwm::kSurParams * sizeof(float),
                          hipMemcpyDeviceToHost, m->ctx->stream));
    HIPCHK(hipStreamSynchronize(m->ctx->stream));
  });
}

int gwm_surrogate_sample(gwm_state* s, const gwm_warm_rain_config* c, unsigned long long seed,
                         long n, void* dev_x, void* dev_y, void* dev_cell) {
  return guard([&] {
    S_(s);
    require(c && dev_x && dev_y, "null argument");
    require(n > 0, "sample count must be positive");
    require(s->ctx->prec == GWM_F32, "sampling runs on FP32 states");
    require(s->ntr >= 3, "sampling requires the qv, qc, qr tracers");
    void** ds = s->cur_set();
    gwm::launch_sur_sample(s->ctx->stream, s->g(), static_cast<const float*>(ds[0]),
                           static_cast<const float*>(ds[1]),
                           static_cast<const float*>(s->field(GWM_TRACER0 + 0)),
                           static_cast<const float*>(s->field(GWM_TRACER0 + 1)),
                           static_cast<const float*>(s->field(GWM_TRACER0 + 2)), micro_cfg(*c),
                           seed, n, static_cast<float*>(dev_x), static_cast<float*>(dev_y),
                           static_cast<long long*>(dev_cell), s->base<float>());
    HIPCHK(hipGetLastError());
  });
}

int gwm_surrogate_train(gwm_surrogate* m, const void* dev_x, const void* dev_y, long n, int batch,
                        int epochs, double lr, double lr_final, double* epoch_loss) {
  return guard([&] {
    require(m && dev_x && dev_y, "null argument");
    require(m->have_norm, "surrogate normalisation not set");
    require(batch >= 64 && batch % 64 == 0, "batch must be a positive multiple of 64");
    require(n >= batch && epochs >= 1, "need at least one batch and one epoch");
    require(lr > 0.0 && lr_final > 0.0, "learning rates must be positive");
    m->ensure_batch(batch);
    const long nmb = n / batch;
    const int per = (batch * gwm::kSurOut + 255) / 256;
    double* part = m->loss_buf(static_cast<size_t>(nmb) * per);
    const long total = nmb * epochs;
    const float* X = static_cast<const float*>(dev_x);
    const float* Y = static_cast<const float*>(dev_y);
    std::vector<double> h(static_cast<size_t>(nmb) * per);
    long step = 0;
    for (int e = 0; e < epochs; ++e) {
      for (long b = 0; b < nmb; ++b, ++step) {
        // cosine decay from lr to lr_final over all steps
        const double f = 0.5 * (1.0 + std::cos(M_PI * static_cast<double>(step) / total));
        m->T.lr = static_cast<float>(lr_final + (lr - lr_final) * f);
        int nl = 0;
        gwm::launch_sur_train_step(m->ctx->stream, m->T, X + b * batch * gwm::kSurIn,
                                   Y + b * batch * gwm::kSurOut, batch, m->norm, true,
                                   part + b * per, &nl);
      }
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpyAsync(h.data(), part, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                            m->ctx->stream));
      HIPCHK(hipStreamSynchronize(m->ctx->stream));
      double sum = 0.0;
      for (double v : h) sum += v;
      if (epoch_loss) epoch_loss[e] = sum / (static_cast<double>(nmb) * batch * gwm::kSurOut);
    }
  });
}

int gwm_surrogate_loss_grad(gwm_surrogate* m, const void* dev_x, const void* dev_y, int n,
                            double* loss, float* grad_host) {
  return guard([&] {
    require(m && dev_x && dev_y && loss && grad_host, "null argument");
    require(m->have_norm, "surrogate normalisation not set");
    require(n >= 1, "need at least one row");
    m->ensure_batch(n);
    const int per = (n * gwm::kSurOut + 255) / 256;
    double* part = m->loss_buf(static_cast<size_t>(per));
    int nl = 0;
    gwm::launch_sur_train_step(m->ctx->stream, m->T, static_cast<const float*>(dev_x),
                               static_cast<const float*>(dev_y), n, m->norm, false, part, &nl);
    HIPCHK(hipGetLastError());
    std::vector<double> h(static_cast<size_t>(nl));
    HIPCHK(hipMemcpyAsync(h.data(), part, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                          m->ctx->stream));
    HIPCHK(hipMemcpyAsync(grad_host, m->T.grad, gwm::kSurParams * sizeof(float),
                          hipMemcpyDeviceToHost, m->ctx->stream));
    HIPCHK(hipStreamSynchronize(m->ctx->stream));
    double sum = 0.0;
    for (double v : h) sum += v;
    *loss = sum / (static_cast<double>(n) * gwm::kSurOut);
  });
}

int gwm_surrogate_infer(gwm_surrogate* m, const void* dev_x, long n, void* dev_y) {
  return guard([&] {
    require(m && dev_x && dev_y, "null argument");
    require(m->have_norm, "surrogate normalisation not set");
    require(n >= 1, "need at least one row");
    Timed t(m->ctx, "surrogate_infer");
    gwm::launch_sur_infer(m->ctx->stream, m->T.params, m->norm, static_cast<const float*>(dev_x),
                          n, static_cast<float*>(dev_y));
    HIPCHK(hipGetLastError());
  });
}

int gwm_warm_rain_surrogate(gwm_state* s, const gwm_warm_rain_config* c, gwm_surrogate* m,
                            int metric_fallout, long long* n_fallback) {
  return guard([&] {
    S_(s);
    require(c && m, "null argument");
    warm_rain_surrogate(s, *c, m, metric_fallout, n_fallback);
  });
}

int gwm_surrogate_last_counts(gwm_surrogate* m, long long* fallback, long long* network) {
  return guard([&] {
    require(m != nullptr, "null surrogate");
    unsigned long long h[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(h, m->nfall, sizeof h, hipMemcpyDeviceToHost, m->ctx->stream));
    HIPCHK(hipStreamSynchronize(m->ctx->stream));
    if (fallback) *fallback = static_cast<long long>(h[0]);
    if (network) *network = static_cast<long long>(h[1]);
  });
}

int gwm_moist_step_surrogate(gwm_state* s, const gwm_step_config* c,
                             const gwm_warm_rain_config* wr, gwm_surrogate* m,
                             const gwm_boundary_config* b) {
  return guard([&] {
    S_(s);
    require(c && wr && m, "null argument");
    require(s->ctx->prec == GWM_F32, "the surrogate runs on FP32 states");
    ssprk3<float>(s, *c, b);
    tracers_ssprk3<float>(
        s, *c, b ? (b->mode == GWM_BOUNDARY_STRIPS ? GWM_BOUNDARY_STRIPS : GWM_BOUNDARY_NONE)
                 : GWM_BOUNDARY_NONE);
    warm_rain_surrogate(s, *wr, m, 1, nullptr);
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Graph-replayed moist steps and ensembles (extension).
// ---------------------------------------------------------------------------
namespace {

void moist_step_f32(gwm_state* S, const gwm_step_config& c, const gwm_warm_rain_config& wr,
                    gwm_surrogate* M) {
  ssprk3<float>(S, c, nullptr);
  tracers_ssprk3<float>(S, c, GWM_BOUNDARY_NONE);
  if (M) warm_rain_surrogate(S, wr, M, 1, nullptr);
  else warm_rain<float>(S, wr, 1);
}

template <class R>
void moist_step_any(gwm_state* S, const gwm_step_config& c, const gwm_warm_rain_config& wr,
                    gwm_surrogate* M) {
  if constexpr (std::is_same<R, float>::value) {
    moist_step_f32(S, c, wr, M);
  } else {
    require(M == nullptr, "the surrogate runs on FP32 states");
    ssprk3<R>(S, c, nullptr);
    tracers_ssprk3<R>(S, c, GWM_BOUNDARY_NONE);
    warm_rain<R>(S, wr, 1);
  }
}

std::vector<double> graph_key(const gwm_step_config& c, const gwm_warm_rain_config& w,
                              const gwm_surrogate* M) {
  return {c.dt, static_cast<double>(c.fast_substeps), c.gravity,
          static_cast<double>(c.fast_update_horizontal_momentum),
          static_cast<double>(c.fast_update_density), w.dt, w.condensation_relaxation,
          w.cloud_autoconversion_threshold, w.cloud_autoconversion_rate,
          w.rain_evaporation_rate, w.rain_terminal_velocity, w.sedimentation_layer_depth_m,
          static_cast<double>(w.enable_latent_heating), w.accretion_rate,
          static_cast<double>(reinterpret_cast<uintptr_t>(M)),
          M ? static_cast<double>(M->version) : -1.0};
}

// The captured graph of two steps from the current buffer parity, capturing
// it on first use.  Before a capture, one eager step runs (it chooses the
// chunk counts, allocates every scratch buffer and enters the two-step
// buffer cycle); returns the number of eager steps taken (0 or 1).
// Device buffers a captured moist-step graph references: the state's and,
// with the surrogate, its parameters and fallback counter.
std::vector<const void*> graph_buffers(const gwm_state* S, const gwm_surrogate* M) {
  auto b = S->buffers();
  if (M) {
    b.push_back(M->T.params);
    b.push_back(M->nfall);
  }
  return b;
}

// A buffer that existed when the graph was captured must still be the same
// allocation (one allocated later, null at capture, is not referenced).
void require_graph_buffers(const std::vector<const void*>& captured,
                           const std::vector<const void*>& now) {
  bool same = captured.size() == now.size();
  for (size_t i = 0; same && i < captured.size(); ++i)
    same = captured[i] == nullptr || captured[i] == now[i];
  require(same, "internal: a device buffer referenced by a captured step graph was reallocated");
}

template <class R>
int ensure_graph(gwm_state* S, const gwm_step_config& c, const gwm_warm_rain_config& w,
                 gwm_surrogate* M, gwm_state::Graph** out) {
  const auto key = graph_key(c, w, M);
  for (auto& gr : S->graphs)
    if (gr.exec && gr.cur == S->cur && gr.tr0 == (S->tr.empty() ? nullptr : S->tr[0]) &&
        gr.key == key) {
      require_graph_buffers(gr.bufs, graph_buffers(S, M));
      *out = &gr;
      return 0;
    }
  int eager = 0;
  if (S->graphs[0].exec == nullptr || S->graphs[0].key != key) {
    moist_step_any<R>(S, c, w, M);
    eager = 1;
    for (auto& gr : S->graphs)
      if (gr.exec) {
        HIPCHK(hipGraphExecDestroy(gr.exec));
        gr = gwm_state::Graph{};
      }
  }
  gwm_state::Graph& gr = S->graphs[S->graphs[0].exec ? 1 : 0];
  if (gr.exec) {
    HIPCHK(hipGraphExecDestroy(gr.exec));
    gr = gwm_state::Graph{};
  }
  const int cur0 = S->cur;
  void* tr0 = S->tr.empty() ? nullptr : S->tr[0];
  const bool timing = S->ctx->timing;
  S->ctx->timing = false;
  hipGraph_t graph = nullptr;
  HIPCHK(hipStreamBeginCapture(S->ctx->stream, hipStreamCaptureModeThreadLocal));
  try {
    moist_step_any<R>(S, c, w, M);
    moist_step_any<R>(S, c, w, M);
  } catch (...) {
    (void)hipStreamEndCapture(S->ctx->stream, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    S->ctx->timing = timing;
    throw;
  }
  HIPCHK(hipStreamEndCapture(S->ctx->stream, &graph));
  S->ctx->timing = timing;
  require(S->cur == cur0 && (S->tr.empty() ? nullptr : S->tr[0]) == tr0,
          "graph capture needs the two-step buffer cycle");
  hipGraphExec_t exec = nullptr;
  const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  HIPCHK(e);
  gr.exec = exec;
  gr.cur = cur0;
  gr.tr0 = tr0;
  gr.key = key;
  gr.bufs = graph_buffers(S, M);
  *out = &gr;
  return eager;
}

// ---------------------------------------------------------------------------
// Batched ensemble: one launch per kernel of the step for all members.
// Each member's two moist steps are recorded (the launchers append their
// batched forms, nothing runs; two steps return the buffer cycle to where it
// started); position i of every member's sequence is the same kernel, so it
// becomes one launch over the members' work-groups, each member's range
// starting at a multiple of 8 (the XCD remap inside a member stays the one of
// its own launch).  The two-step sequence is captured into a graph on the
// first member's stream and replayed.
// ---------------------------------------------------------------------------
struct BatchPlan {
  struct Op {
    const void* fn;
    dim3 grid, block;
    size_t lds;
    void* items;
    gwm::BatchMap map;
  };
  std::vector<Op> ops;
  void* mem = nullptr;
  hipStream_t stream = nullptr;  // the group's own stream (default priority)
  hipGraphExec_t exec = nullptr;
  BatchPlan() = default;
  BatchPlan(const BatchPlan&) = delete;
  BatchPlan& operator=(const BatchPlan&) = delete;
  ~BatchPlan() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (stream) (void)hipStreamDestroy(stream);
    if (mem) (void)hipFree(mem);
  }
};

// Records two moist steps of a member; false if a launcher has no batched
// form (the recorded steps then cannot be batched).
template <class R>
bool record_two_steps(gwm_state* S, const gwm_step_config& c, const gwm_warm_rain_config& w,
                      gwm_surrogate* M, gwm::Recorder& rec) {
  const int cur0 = S->cur;
  void* tr0 = S->tr.empty() ? nullptr : S->tr[0];
  const bool timing = S->ctx->timing;
  S->ctx->timing = false;
  gwm::recorder() = &rec;
  try {
    moist_step_any<R>(S, c, w, M);
    moist_step_any<R>(S, c, w, M);
  } catch (...) {
    gwm::recorder() = nullptr;
    S->ctx->timing = timing;
    throw;
  }
  gwm::recorder() = nullptr;
  S->ctx->timing = timing;
  require(S->cur == cur0 && (S->tr.empty() ? nullptr : S->tr[0]) == tr0,
          "internal: two recorded steps did not return the buffer cycle to its start");
  return rec.unsupported == nullptr;
}

// The plan for the members' recorded sequences, or null if they differ.
std::unique_ptr<BatchPlan> build_plan(const std::vector<gwm::Recorder>& recs) {
  const size_t nops = recs[0].ops.size();
  for (const auto& r : recs)
    if (r.ops.size() != nops) return nullptr;
  const int n = static_cast<int>(recs.size());
  require(n < (1 << (32 - gwm::kBatchBlockBits)) - 1, "a batched ensemble takes at most 4094 members");
  auto align = [](size_t v) { return (v + 255) & ~static_cast<size_t>(255); };
  // Position i of the sequences becomes one launch per distinct kernel form
  // there (members tuned to different tile heights run different
  // instantiations); usually one.
  struct Sub {
    size_t op;
    std::vector<int> members;
    std::vector<int> starts;
    int total = 0;
    size_t off_items = 0, off_map = 0;
  };
  std::vector<Sub> subs;
  size_t bytes = 0;
  for (size_t i = 0; i < nops; ++i) {
    std::vector<Sub> here;
    for (int m = 0; m < n; ++m) {
      const gwm::LaunchRecord& r = recs[static_cast<size_t>(m)].ops[i];
      require(static_cast<long>(r.gx) * r.gy < (1L << gwm::kBatchBlockBits),
              "a batched member's grid exceeds 2^20 work-groups");
      Sub* sub = nullptr;
      for (auto& h : here) {
        const gwm::LaunchRecord& q = recs[static_cast<size_t>(h.members[0])].ops[i];
        if (q.fn == r.fn && q.bx == r.bx && q.by == r.by && q.bz == r.bz &&
            q.item.size() == r.item.size())
          sub = &h;
      }
      if (!sub) {
        // different kernels at one position must at least agree in kind
        // (the same launcher), or the sequences differ
        if (!here.empty() &&
            recs[static_cast<size_t>(here[0].members[0])].ops[i].item.size() != r.item.size())
          return nullptr;
        here.push_back(Sub{i, {}, {}, 0, 0, 0});
        sub = &here.back();
      }
      sub->total = (sub->total + 7) & ~7;
      sub->starts.push_back(sub->total);
      sub->members.push_back(m);
      sub->total += r.gx * r.gy;
    }
    for (auto& h : here) {
      h.off_items = bytes;
      bytes = align(bytes + recs[0].ops[i].item.size() * h.members.size());
      h.off_map = bytes;
      bytes = align(bytes + sizeof(unsigned) * static_cast<size_t>(h.total));
      subs.push_back(std::move(h));
    }
  }
  std::vector<unsigned char> host(bytes, 0);
  auto plan = std::make_unique<BatchPlan>();
  plan->mem = dalloc(std::max<size_t>(bytes, 256));
  unsigned char* dev = static_cast<unsigned char*>(plan->mem);
  for (const Sub& h : subs) {
    const gwm::LaunchRecord& r0 = recs[static_cast<size_t>(h.members[0])].ops[h.op];
    const size_t isz = r0.item.size();
    size_t lds = 0;
    std::vector<unsigned> map(static_cast<size_t>(h.total), gwm::kNoMember);
    for (size_t q = 0; q < h.members.size(); ++q) {
      const gwm::LaunchRecord& r = recs[static_cast<size_t>(h.members[q])].ops[h.op];
      std::memcpy(host.data() + h.off_items + isz * q, r.item.data(), isz);
      lds = std::max(lds, r.lds);
      const int s0 = h.starts[q];
      for (int b = 0; b < r.gx * r.gy; ++b)
        map[static_cast<size_t>(s0 + b)] =
            (static_cast<unsigned>(q) << gwm::kBatchBlockBits) | static_cast<unsigned>(b);
    }
    std::memcpy(host.data() + h.off_map, map.data(), sizeof(unsigned) * map.size());
    BatchPlan::Op op;
    op.fn = r0.fn;
    op.grid = dim3(static_cast<unsigned>(h.total));
    op.block = dim3(r0.bx, r0.by, r0.bz);
    op.lds = lds;
    op.items = dev + h.off_items;
    op.map = gwm::BatchMap{reinterpret_cast<const unsigned*>(dev + h.off_map)};
    plan->ops.push_back(op);
  }
  HIPCHK(hipMemcpy(plan->mem, host.data(), bytes, hipMemcpyHostToDevice));
  return plan;
}

void launch_plan_ops(const BatchPlan& plan, hipStream_t st) {
  for (const auto& op : plan.ops) {
    void* items = op.items;
    gwm::BatchMap map = op.map;
    void* args[] = {&items, &map};
    HIPCHK(hipLaunchKernel(op.fn, op.grid, op.block, args, op.lds, st));
  }
}

// Launch codes the column kernels of a batched ensemble are recorded with:
// 0 keeps each member's own tuned code (chosen for a launch of that member
// alone; a position whose members differ in tile height becomes one launch
// per height), the others apply to every member: level chunks 2, 4, 8 with
// 8-row tiles (4-row tiles for all members measured 13 % slower on the
// 32-member ensemble and are not tried).  The ensemble's own autotuner
// times each candidate on real step pairs (the candidates give identical
// results) and keeps the fastest per ensemble signature.
constexpr int kEnsCands = 4;
constexpr int kEnsChunkCands[kEnsCands] = {0, 2, 4, 8};
struct EnsembleTuning {
  float ms[kEnsCands] = {-1.f, -1.f, -1.f, -1.f};  // step-pair time, -1 untested
};
constexpr float kNoPlan = 1e30f;  // a candidate whose steps cannot be batched
std::mutex g_ens_mu;
std::map<std::string, EnsembleTuning> g_ens_tuning;

template <class R>
std::string ensemble_signature(gwm_state* const* states, int n, const gwm_step_config& c,
                               gwm_surrogate* const* surs) {
  std::string sig = std::to_string(sizeof(R)) + ":" + std::to_string(c.fast_substeps);
  for (int e = 0; e < n; ++e) {
    const gwm_state* S = states[e];
    const Geo& g = S->g();
    char b[96];
    std::snprintf(b, sizeof b, "|%dx%dx%d:%d%d%d:%d", g.nx, g.ny, g.nz, S->pert ? 1 : 0,
                  S->grid->terrain_nonzero ? 1 : 0, (surs && surs[e]) ? 1 : 0, S->ntr);
    sig += b;
  }
  return sig;
}

// A plan of two batched steps recorded with level-chunk override `chunks`,
// captured into a graph on a stream of its own; null if the steps cannot be
// batched.
template <class R>
std::unique_ptr<BatchPlan> make_batch_plan(gwm_state* const* states, int n,
                                           const gwm_step_config& c,
                                           const gwm_warm_rain_config& w,
                                           gwm_surrogate* const* surs, int chunks) {
  std::vector<std::pair<gwm_ctx*, int>> saved;
  for (int e = 0; e < n; ++e) {
    gwm_ctx* x = states[e]->ctx;
    bool seen = false;
    for (auto& p : saved) seen = seen || p.first == x;
    if (!seen) saved.push_back({x, x->force_chunks});
  }
  auto restore = [&] {
    for (auto& p : saved) p.first->force_chunks = p.second;
  };
  if (chunks > 0)
    for (auto& p : saved) p.first->force_chunks = chunks;
  std::vector<gwm::Recorder> recs(static_cast<size_t>(n));
  bool ok = true;
  try {
    for (int e = 0; e < n && ok; ++e)
      ok = record_two_steps<R>(states[e], c, w, surs ? surs[e] : nullptr,
                               recs[static_cast<size_t>(e)]);
  } catch (...) {
    restore();
    throw;
  }
  restore();
  if (!ok) return nullptr;
  auto plan = build_plan(recs);
  if (!plan) return nullptr;
  HIPCHK(hipStreamCreateWithFlags(&plan->stream, hipStreamNonBlocking));
  hipGraph_t graph = nullptr;
  HIPCHK(hipStreamBeginCapture(plan->stream, hipStreamCaptureModeThreadLocal));
  try {
    launch_plan_ops(*plan, plan->stream);
  } catch (...) {
    (void)hipStreamEndCapture(plan->stream, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    throw;
  }
  HIPCHK(hipStreamEndCapture(plan->stream, &graph));
  const hipError_t ge = hipGraphInstantiate(&plan->exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  HIPCHK(ge);
  return plan;
}

// Batched steps of the members (all of precision R); returns the steps every
// member has taken: nsteps, or fewer when their steps cannot be batched (a
// launcher without a batched form, or sequences that differ), the caller
// then running the rest otherwise.  When a member's last warm-up was for
// another configuration, every member first takes one eager step
// (allocation and tuning), counted in nsteps.
template <class R>
int ensemble_batched(gwm_state* const* states, int n, const gwm_step_config& c,
                     const gwm_warm_rain_config& w, gwm_surrogate* const* surs, int nsteps) {
  auto eager = [&](int e) { moist_step_any<R>(states[e], c, w, surs ? surs[e] : nullptr); };
  bool need_warm = false;
  for (int e = 0; e < n; ++e) {
    gwm_state* S = states[e];
    ensure_synced<R>(S, S->set[S->cur]);
    need_warm = need_warm || S->batch_key != graph_key(c, w, surs ? surs[e] : nullptr);
  }
  int left = nsteps;
  if (need_warm && left > 0) {
    for (int e = 0; e < n; ++e) {
      eager(e);
      states[e]->batch_key = graph_key(c, w, surs ? surs[e] : nullptr);
    }
    --left;
  }
  if (left >= 2) {
    const std::string sig = ensemble_signature<R>(states, n, c, surs);
    EnsembleTuning tune;
    {
      std::lock_guard<std::mutex> lock(g_ens_mu);
      tune = g_ens_tuning[sig];
    }
    std::unique_ptr<BatchPlan> plans[kEnsCands];
    auto plan_for = [&](int ci) -> BatchPlan* {
      if (!plans[ci]) plans[ci] = make_batch_plan<R>(states, n, c, w, surs, kEnsChunkCands[ci]);
      return plans[ci].get();
    };
    // every member's queued work first
    for (int e = 0; e < n; ++e) HIPCHK(hipStreamSynchronize(states[e]->ctx->stream));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    BatchPlan* last = nullptr;
    try {
      while (left >= 2) {
        int ci = -1;
        for (int q = 0; q < kEnsCands && ci < 0; ++q)
          if (tune.ms[q] < 0.f) ci = q;
        if (ci >= 0) {  // a candidate not timed yet: one timed step pair
          BatchPlan* p = plan_for(ci);
          if (!p) {
            // not batchable with this candidate (members tuned to different
            // tile heights give different kernels at one position)
            tune.ms[ci] = kNoPlan;
            continue;
          }
          if (last && last != p) HIPCHK(hipStreamSynchronize(last->stream));
          HIPCHK(hipEventRecord(e0, p->stream));
          HIPCHK(hipGraphLaunch(p->exec, p->stream));
          HIPCHK(hipEventRecord(e1, p->stream));
          HIPCHK(hipEventSynchronize(e1));
          float ms = 0.f;
          HIPCHK(hipEventElapsedTime(&ms, e0, e1));
          tune.ms[ci] = ms;
          last = p;
        } else {
          int best = 0;
          for (int q = 1; q < kEnsCands; ++q)
            if (tune.ms[q] < tune.ms[best]) best = q;
          BatchPlan* p = tune.ms[best] < kNoPlan ? plan_for(best) : nullptr;
          if (!p) break;  // the steps cannot be batched
          if (last && last != p) HIPCHK(hipStreamSynchronize(last->stream));
          HIPCHK(hipGraphLaunch(p->exec, p->stream));
          last = p;
        }
        left -= 2;
      }
      if (last) HIPCHK(hipStreamSynchronize(last->stream));
    } catch (...) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    {
      std::lock_guard<std::mutex> lock(g_ens_mu);
      g_ens_tuning[sig] = tune;
    }
    if (left >= 2) return nsteps - left;  // not batchable: the caller runs the rest
  }
  for (int e = 0; e < n; ++e)
    for (int q = 0; q < left; ++q) eager(e);
  return nsteps;
}

// The batched ensemble autotuner's times per signature, for reports.
std::string ensemble_tuning_report() {
  std::lock_guard<std::mutex> lock(g_ens_mu);
  std::string out;
  for (const auto& kv : g_ens_tuning) {
    char b[160];
    std::snprintf(b, sizeof b, "%zu members:", static_cast<size_t>(std::count(
                                                   kv.first.begin(), kv.first.end(), '|')));
    out += b;
    for (int q = 0; q < kEnsCands; ++q) {
      const int cc = kEnsChunkCands[q];
      char what[48];
      if (cc == 0)
        std::snprintf(what, sizeof what, "own codes");
      else
        std::snprintf(what, sizeof what, "chunks %d rows %d", cc & 15, (cc >> 4) ? 4 : 8);
      if (kv.second.ms[q] >= kNoPlan)
        std::snprintf(b, sizeof b, " %s not batchable", what);
      else if (kv.second.ms[q] < 0.f)
        std::snprintf(b, sizeof b, " %s untried", what);
      else
        std::snprintf(b, sizeof b, " %s %.3f ms", what, kv.second.ms[q]);
      out += b;
    }
    out += "\n";
  }
  return out;
}

template <class R>
void moist_steps_graph(gwm_state* S, const gwm_step_config& c, const gwm_warm_rain_config& w,
                       gwm_surrogate* M, int nsteps) {
  int done = 0;
  ensure_synced<R>(S, S->set[S->cur]);
  while (nsteps - done >= 2) {
    gwm_state::Graph* gr = nullptr;
    done += ensure_graph<R>(S, c, w, M, &gr);
    if (nsteps - done < 2) break;
    HIPCHK(hipGraphLaunch(gr->exec, S->ctx->stream));
    done += 2;
  }
  while (done < nsteps) {
    moist_step_any<R>(S, c, w, M);
    ++done;
  }
}

}  // namespace

extern "C" {

int gwm_moist_steps_graph(gwm_state* s, const gwm_step_config* c, const gwm_warm_rain_config* wr,
                          gwm_surrogate* m, int nsteps) {
  return guard([&] {
    S_(s);
    require(c && wr, "null config");
    require(nsteps >= 0, "nsteps must be >= 0");
    require(!s->view, "not on a stage view");
    if (m) require(m->ctx == s->ctx, "surrogate and state belong to different contexts");
    dispatch(s->ctx, [&](auto r) {
      moist_steps_graph<decltype(r)>(s, *c, *wr, m, nsteps);
      return 0;
    });
  });
}

namespace {

// Members one context (stream) each, step pairs replayed from each member's
// captured graph, interleaved across members.
void ensemble_streams(gwm_state* const* states, int n, const gwm_step_config* c,
                      const gwm_warm_rain_config* wr, gwm_surrogate* const* surrogates,
                      int nsteps) {
  for (int e = 0; e < n; ++e)
    for (int f = 0; f < e; ++f)
      require(states[f]->ctx != states[e]->ctx,
              "ensemble members stepped on their own streams need a context each");
  std::vector<int> done(static_cast<size_t>(n), 0);
  for (int e = 0; e < n; ++e)
    dispatch(states[e]->ctx, [&](auto r) {
      ensure_synced<decltype(r)>(states[e], states[e]->set[states[e]->cur]);
      return 0;
    });
  bool more = true;
  while (more) {
    more = false;
    for (int e = 0; e < n; ++e) {
      gwm_state* S = states[e];
      gwm_surrogate* M = surrogates ? surrogates[e] : nullptr;
      int& d = done[static_cast<size_t>(e)];
      if (nsteps - d < 2) continue;
      dispatch(S->ctx, [&](auto r) {
        using R = decltype(r);
        gwm_state::Graph* gr = nullptr;
        d += ensure_graph<R>(S, *c, *wr, M, &gr);
        if (nsteps - d >= 2) {
          HIPCHK(hipGraphLaunch(gr->exec, S->ctx->stream));
          d += 2;
        }
        return 0;
      });
      if (nsteps - d >= 2) more = true;
    }
  }
  for (int e = 0; e < n; ++e) {
    gwm_state* S = states[e];
    gwm_surrogate* M = surrogates ? surrogates[e] : nullptr;
    while (done[static_cast<size_t>(e)] < nsteps) {
      dispatch(S->ctx, [&](auto r) {
        moist_step_any<decltype(r)>(S, *c, *wr, M);
        return 0;
      });
      ++done[static_cast<size_t>(e)];
    }
  }
  for (int e = 0; e < n; ++e) HIPCHK(hipStreamSynchronize(states[e]->ctx->stream));
}

void check_members(gwm_state* const* states, int n, const gwm_step_config* c,
                   const gwm_warm_rain_config* wr, gwm_surrogate* const* surrogates,
                   int nsteps) {
  require(states && c && wr && n >= 1, "null argument");
  require(nsteps >= 0, "nsteps must be >= 0");
  for (int e = 0; e < n; ++e) {
    S_(states[e]);
    require(!states[e]->view, "not on a stage view");
    require(!states[e]->pert3, "ensemble members in the perturbation form on terrain are not supported");
    for (int f = 0; f < e; ++f) require(states[f] != states[e], "a member appears twice");
    if (surrogates && surrogates[e])
      require(surrogates[e]->ctx == states[e]->ctx,
              "surrogate and state belong to different contexts");
  }
}

}  // namespace

int gwm_ensemble_steps(gwm_state* const* states, int n, const gwm_step_config* c,
                       const gwm_warm_rain_config* wr, gwm_surrogate* const* surrogates,
                       int nsteps) {
  return guard([&] {
    check_members(states, n, c, wr, surrogates, nsteps);
    bool same = true;
    for (int e = 1; e < n; ++e)
      same = same && states[e]->ctx->prec == states[0]->ctx->prec &&
             states[e]->ctx->device == states[0]->ctx->device;
    int done = 0;
    if (same)
      done = dispatch(states[0]->ctx, [&](auto r) {
        return ensemble_batched<decltype(r)>(states, n, *c, *wr, surrogates, nsteps);
      });
    if (done < nsteps) {
      bool distinct = true;
      for (int e = 0; e < n; ++e)
        for (int f = 0; f < e; ++f) distinct = distinct && states[f]->ctx != states[e]->ctx;
      if (distinct) {
        ensemble_streams(states, n, c, wr, surrogates, nsteps - done);
      } else {  // members sharing a context: one after the other
        for (int e = 0; e < n; ++e)
          for (int q = done; q < nsteps; ++q)
            dispatch(states[e]->ctx, [&](auto r) {
              moist_step_any<decltype(r)>(states[e], *c, *wr,
                                          surrogates ? surrogates[e] : nullptr);
              return 0;
            });
        for (int e = 0; e < n; ++e) HIPCHK(hipStreamSynchronize(states[e]->ctx->stream));
      }
    }
  });
}

int gwm_ensemble_tuning_report(char* buf, size_t size) {
  return guard([&] {
    require(buf && size > 0, "null argument");
    std::snprintf(buf, size, "%s", ensemble_tuning_report().c_str());
  });
}

int gwm_ensemble_steps_streams(gwm_state* const* states, int n, const gwm_step_config* c,
                               const gwm_warm_rain_config* wr, gwm_surrogate* const* surrogates,
                               int nsteps) {
  return guard([&] {
    check_members(states, n, c, wr, surrogates, nsteps);
    ensemble_streams(states, n, c, wr, surrogates, nsteps);
  });
}

}  // extern "C"
