// gwm_slow.hip -- column-marching slow stage for gfx950.
//
// compute_slow_tendencies (src/dycore/dry_core.cu:400-539) = K2 continuity and
// upwind theta flux + K16-K19 momentum flux divergence + K21-K22 horizontal
// pressure gradient + K3 buoyancy, fused with the stage update / SSP blend of
// advance_dry_state_ssprk3 (dry_core.cu:572-646, K4/K5).  Modes: 0 writes
// tendencies; 1 out = max(floor, S + dt*T); 2 out = wa*Q0 + wb*max(floor,
// S + dt*T).  With out_theta the cell output is theta = (rho theta)_new /
// max(rho_new, 1e-6), the fast-mode entry value (dry_fast_modes.cu:86-95).
// A thread owns cell (i,j,k), x-face i, y-face j and z-face k, plus face
// nx / ny / nz on the last index.  Arithmetic is the step kernels' q* forms
// (gwm_internal.hpp).
//
// A work-group of 64 x 4 threads marches a 64 x 4 column tile upwards.  For
// each level it stages, over the tile plus a one-cell ring, the derived cell
// quantities every flux needs -- theta, p (dry_thermo.hpp:21-33) and the cell
// velocities (dry_momentum_flux.cu:36-73) -- into LDS once, instead of
// recomputing each of them at every use (a cell's p and velocities enter up
// to a dozen face fluxes).  Four LDS slots per quantity hold levels k-1 ..
// k+1 and the one being replaced (level k+1 is staged at the top of
// iteration k); one barrier per level.
// The column is split into level chunks (count chosen by the autotuner) so
// the chip holds more independent work-groups; a chunk stages the level
// below its first one.
// FP32 stages the raw face momenta of each level in LDS as well, so the
// upwind values of every flux come from LDS; FP64 reads them from global
// memory.  Global offsets of the staged positions are computed once per
// work-group (stage_pos), so the halo map leaves the level loop.
//
// The tracer stage below follows the same pattern.
#include "gwm_kernels.hpp"

namespace gwm {

namespace {

constexpr int TX = 64, TY = 4, PX = TX + 2, PY = TY + 2, PN = PX * PY;
constexpr int NHALO = PN - TX * TY;

// Work-group -> (tile, level chunk [k0, k1)).  Work-group b runs on XCD
// b % 8; the remap gives each XCD a contiguous band of tiles of one chunk.
__device__ inline int tile_of(int blk, int tiles, int nkc, int nz, int& b, int& k0, int& k1) {
  const int nb = tiles * nkc;
  b = blk;
  if ((nb & 7) == 0) b = (b & 7) * (nb >> 3) + (b >> 3);
  const int chunk = b / tiles;
  k0 = (nz * chunk) / nkc;
  k1 = (nz * (chunk + 1)) / nkc;
  return b % tiles;
}

template <class R>
__device__ inline R upwind(R a, R ql, R qr) {  // dry_momentum_flux.cu:75-77
  return a >= R(0) ? a * ql : a * qr;
}


}  // namespace

// Per-level values of a chunk (levels k0-1 .. k1+1), staged in LDS once per
// work-group.
enum { SL_TREF = 0, SL_IDZF, SL_IDZC, SL_BRHO, SL_BTH, SL_BRT, SL_BP, SL_N };

template <class R, bool TERRAIN, bool PERT, int MODE>
__device__ __forceinline__ void slow_body(const SlowArgs<R>& a, int tiles_x, int tiles, int nkc,
                                          const VBlock& vb) {
  // FP32 also stages the raw face momenta of each level (tile plus ring), so
  // the upwind values of every flux come from LDS; FP64 reads them from
  // global memory (its LDS budget would halve the resident work-groups).
  //
  // LDS holds four level slots (slot kk % 4): level k+1 is staged into the
  // slot of k-3, which every wave finished reading before the previous
  // level's barrier, so each level takes one barrier (before the reads).
  // The fourth slot costs a third more LDS, 52 KB per work-group in FP32;
  // three work-groups per CU still fit, and the registers allow no more.
  //
  // FP32 keeps the staged quantities of one position together, (quantity,
  // slot) innermost, with an odd stride between positions (no LDS bank
  // conflicts between neighbouring lanes): a thread's reads are its own entry
  // plus compile-time offsets, so the per-level address arithmetic is one add
  // per slot.  FP64 keeps one plane per quantity and slot (its register
  // budget is the tighter one, and the interleaved form costs it registers).
  constexpr bool RAW = sizeof(R) == 4;
  constexpr int NSL = 4;
  constexpr int NQ = RAW ? 8 : 5;  // theta, p, uc, vc, wc (+ raw u, v, w)
  constexpr int QS = (NQ * NSL) | 1;
  enum { Q_TH = 0, Q_P, Q_UC, Q_VC, Q_WC, Q_U, Q_V, Q_W };
  __shared__ R sA[RAW ? PN * QS : NQ * NSL * PN];
  auto lds = [&](int q, int slot, int e) -> R& {
    if constexpr (RAW) return sA[e * QS + slot * NQ + q];
    else return sA[(q * NSL + slot) * PN + e];
  };
  const Geo g = a.g;
  const Loader<R> L{g};
  const DState<R> s = a.s;

  int b, k0, k1;
  const int tile = tile_of(vb.x, tiles, nkc, g.nz, b, k0, k1);
  (void)b;
  const int ti0 = (tile % tiles_x) * TX, tj0 = (tile / tiles_x) * TY;
  const int lx = threadIdx.x, ly = threadIdx.y;
  const int t = lx + TX * ly;
  const int i = ti0 + lx, j = tj0 + ly;
  const bool active = i < g.nx && j < g.ny;

  int hx = 0, hy = 0;
  if (t < 2 * PX) {
    hx = (t % PX) - 1;
    hy = t < PX ? -1 : TY;
  } else if (t < NHALO) {
    const int q = t - 2 * PX;
    hx = (q & 1) ? TX : -1;
    hy = q >> 1;
  }

  const int nz = g.nz;
  constexpr bool pert = PERT;
  // Terrain perturbation form: per-cell bases (level value + deviation),
  // the base's pressure-gradient forcing per face, open sides allowed.
  constexpr bool TP = TERRAIN && PERT;
  const unsigned pc = static_cast<unsigned>(g.pxc) * g.pyc;  // cell / z-face level stride
  const unsigned pu = static_cast<unsigned>(g.pxu) * g.pyc;  // x-face level stride
  const unsigned pv = static_cast<unsigned>(g.pxc) * g.pyv;  // y-face level stride
  const StagePos own = stage_pos(g, i, j);
  const StagePos ring = stage_pos(g, ti0 + hx, tj0 + hy);
  const int eo = (lx + 1) + PX * (ly + 1), eh = (hx + 1) + PX * (hy + 1);

  // Per-level values of levels k0-1 .. k1+1 (indices clamped to the arrays).
  extern __shared__ __align__(16) unsigned char lev_raw[];
  R* const lev = reinterpret_cast<R*>(lev_raw);
  for (int q = t; q < k1 - k0 + 3; q += TX * TY) {
    const int kk = k0 - 1 + q;
    const int kc = kk < 0 ? 0 : (kk >= nz ? nz - 1 : kk), kf = kk < 0 ? 0 : (kk > nz ? nz : kk);
    R* e = lev + q * SL_N;
    e[SL_TREF] = a.theta_ref[kc];
    if (!TERRAIN) {
      e[SL_IDZF] = a.m.lev_idz_f[kf];
      e[SL_IDZC] = a.m.lev_idz_c[kc];
    }
    if (PERT) {
      e[SL_BRHO] = a.base.rho[kc];
      e[SL_BTH] = a.base.th[kc];
      e[SL_BRT] = a.base.rt[kc];
      e[SL_BP] = a.base.p[kc];
    }
  }
  auto LV = [&](int kk, int q) -> R { return lev[(kk - k0 + 1) * SL_N + q]; };
  __syncthreads();

  // Raw values of one level at a staged position.  Loads are unconditional
  // (threads without a ring cell, or outside the domain, read a valid
  // stand-in and discard it), so the level loop is straight-line code and
  // each wait covers exactly the loads it needs: the loads of level k+2 (and
  // of the SSP base of level k+1) stay in flight while level k is computed.
  struct Raw {
    R r, rt, u0, u1, v0, v1, w0, w1;
    R br, bt, bp;  // base deviations (terrain perturbation form)
    bool out;      // a position outside an open domain (zero stand-ins)
  };
  struct Lev {
    Raw o, h;  // own cell, ring cell
  };
  auto ldo = [&](const R* f, int off, unsigned plane, int kk) -> R {
    // The flat perturbation form runs on doubly periodic grids only, where
    // every staged position maps to a stored one: no zero stand-ins to select.
    if constexpr (PERT && !TERRAIN)
      return ld32(f, static_cast<unsigned>(off) + static_cast<unsigned>(kk) * plane);
    const R v = ld32(f, static_cast<unsigned>(off < 0 ? 0 : off) + static_cast<unsigned>(kk) * plane);
    return off < 0 ? R(0) : v;
  };
  const StagePos hpos = t < NHALO ? ring : own;
  auto ldb = [&](const R* f, int off, int kk) -> R {  // base deviations (stand-in: position 0)
    return ld32(f, static_cast<unsigned>(off < 0 ? 0 : off) + static_cast<unsigned>(kk) * pc);
  };
  auto load_raw = [&](const StagePos& o, int kk) -> Raw {
    Raw q{ldo(s.rho, o.c, pc, kk),  ldo(s.rt, o.c, pc, kk),  ldo(s.u, o.u0, pu, kk),
          ldo(s.u, o.u1, pu, kk),   ldo(s.v, o.v0, pv, kk),  ldo(s.v, o.v1, pv, kk),
          ldo(s.w, o.c, pc, kk),    ldo(s.w, o.c, pc, kk + 1), R(0), R(0), R(0), o.c < 0};
    if constexpr (TP) {
      q.br = ldb(a.base.drho, o.c, kk);
      q.bt = ldb(a.base.dth, o.c, kk);
      q.bp = ldb(a.base.dp, o.c, kk);
    }
    return q;
  };
  auto load_lev = [&](int kk) -> Lev { return Lev{load_raw(own, kk), load_raw(hpos, kk)}; };
  // Stages level kk at LDS entry e.
  // FP32 divides the cell velocities (and theta') by rho through one
  // reciprocal (within an ulp of the quotients); FP64 divides each.
  auto stage_raw = [&](const Raw& q, int e, int kk) {
    R rr;
    if constexpr (TP) {
      // theta staged relative to the level's base value theta_b(k) (the
      // cell's base deviation plus its departure), p' against the cell's
      // base; outside an open domain the stand-in is zero theta, as in the
      // full form (theta = -theta_b(k) here), and p' = 0 (never used)
      const R rb = LV(kk, SL_BRHO) + q.br, tb = LV(kk, SL_BTH) + q.bt;
      rr = rmax(rb + q.r, R(1.0e-6));
      lds(Q_TH, kk & (NSL - 1), e) = q.out ? -LV(kk, SL_BTH) : q.bt + (q.rt - tb * q.r) * qrcp(rr);
      lds(Q_P, kk & (NSL - 1), e) = q.out ? R(0) : p_pert(q.rt, rb * tb, LV(kk, SL_BP) + q.bp);
    } else if (pert) {  // theta' and p' of the departures
      const R rb = LV(kk, SL_BRHO);
      rr = rmax(rb + q.r, R(1.0e-6));
      lds(Q_TH, kk & (NSL - 1), e) =
          RAW ? (q.rt - LV(kk, SL_BTH) * q.r) * qrcp(rr) : theta_pert(q.r, q.rt, rb, LV(kk, SL_BTH));
      lds(Q_P, kk & (NSL - 1), e) = p_pert(q.rt, LV(kk, SL_BRT), LV(kk, SL_BP));
    } else {
      rr = rmax(q.r, R(1.0e-6));
      lds(Q_TH, kk & (NSL - 1), e) = q_theta(q.r, q.rt);
      lds(Q_P, kk & (NSL - 1), e) = q_pressure(q.r, q.rt);
    }
    if constexpr (RAW) {
      const R ir = qrcp(rr);
      lds(Q_UC, kk & (NSL - 1), e) = R(0.5) * (q.u0 + q.u1) * ir;
      lds(Q_VC, kk & (NSL - 1), e) = R(0.5) * (q.v0 + q.v1) * ir;
      lds(Q_WC, kk & (NSL - 1), e) = R(0.5) * (q.w0 + q.w1) * ir;
    } else {
      lds(Q_UC, kk & (NSL - 1), e) = qdiv(R(0.5) * (q.u0 + q.u1), rr);
      lds(Q_VC, kk & (NSL - 1), e) = qdiv(R(0.5) * (q.v0 + q.v1), rr);
      lds(Q_WC, kk & (NSL - 1), e) = qdiv(R(0.5) * (q.w0 + q.w1), rr);
    }
    if constexpr (RAW) {
      lds(Q_U, kk & (NSL - 1), e) = q.u0;
      lds(Q_V, kk & (NSL - 1), e) = q.v0;
      lds(Q_W, kk & (NSL - 1), e) = q.w0;
    }
  };
  // Own-column rho (levels k-1, k, k+1), rho*theta (k, k+1) and the top w face.
  R r_km1 = R(0), r_k = R(0), r_k1 = R(0), rt_k = R(0), rt_k1 = R(0), w_top = R(0);
  auto stage_lev = [&](const Lev& q, int kk) {
    stage_raw(q.o, eo, kk);
    r_k1 = q.o.r;
    rt_k1 = q.o.rt;
    if (kk + 1 == nz) w_top = q.o.w1;
    if (t < NHALO) stage_raw(q.h, eh, kk);
  };
  // SSP base of the own cell and faces (mode 2), clamped for threads outside.
  const int ic = i < g.nx ? i : g.nx - 1, jc = j < g.ny ? j : g.ny - 1;
  // Offsets of level kk are the level-0 offset plus kk times the field's
  // level stride (kk is uniform, so the product is scalar work).
  const unsigned q0c = static_cast<unsigned>(g.cell(ic, jc, 0));
  const unsigned q0x = static_cast<unsigned>(g.xf(ic, jc, 0));
  const unsigned q0y = static_cast<unsigned>(g.yf(ic, jc, 0));
  const unsigned oc0 = static_cast<unsigned>(g.cell(i, j, 0));
  const unsigned ox0 = static_cast<unsigned>(g.xf(i, j, 0));
  const unsigned oy0 = static_cast<unsigned>(g.yf(i, j, 0));
  struct Q0 {
    R rho, rt, u, v, w;
  };
  auto load_q0 = [&](int kk) -> Q0 {
    Q0 q{};
    if constexpr (MODE == 2) {
      const unsigned cc = q0c + static_cast<unsigned>(kk) * pc;
      const unsigned xx = q0x + static_cast<unsigned>(kk) * pu;
      const unsigned yy = q0y + static_cast<unsigned>(kk) * pv;
      q = Q0{ld32(a.q0.rho, cc), ld32(a.q0.rt, cc), ld32(a.q0.u, xx), ld32(a.q0.v, yy),
             ld32(a.q0.w, cc)};
    }
    return q;
  };

  const R dx = static_cast<R>(g.dx), dy = static_cast<R>(g.dy);
  const R rdx = R(1) / dx, rdy = R(1) / dy;
  // open sides (none in the flat perturbation form, which is doubly periodic)
  const bool ow = (!PERT || TERRAIN) && !g.per_x, oe = ow;
  const bool os = (!PERT || TERRAIN) && !g.per_y, on = os;
  const size_t plane_m = static_cast<size_t>(g.nx) * g.ny;
  auto mcol = [&](int ii, int jj) -> size_t {
    return static_cast<size_t>(L.mj(jj)) * g.nx + L.mi(ii);
  };

  // highest level a chunk stages (uniform; kept in a scalar register so the
  // clamped level indices and their offsets stay scalar work)
  const int kcap = __builtin_amdgcn_readfirstlane(k1 < nz - 1 ? k1 : nz - 1);
  if (k0 > 0) {
    stage_lev(load_lev(k0 - 1), k0 - 1);
    r_km1 = r_k1;
  }
  stage_lev(load_lev(k0), k0);
  r_k = r_k1;
  rt_k = rt_k1;
  // FP32 pipelines the raw loads a level ahead; FP64 (twice the registers per
  // value) loads each level where it stages it.
  constexpr bool PIPE = sizeof(R) == 4;
  Lev LA{};
  if constexpr (PIPE) LA = load_lev(k0 + 1 < kcap ? k0 + 1 : kcap);

  // Level k: Ln holds the raw level k+1 (loaded an iteration ago).  After
  // staging it, the loads of the SSP base of level k and then those of level
  // k+2 (into Lnn) are issued; both stay in flight while level k is computed,
  // and since loads complete in order, waiting for the base (needed first)
  // leaves level k+2 in flight.  Unrolled by two with the buffers swapped, so
  // no register copy waits for loads in flight.
  // Vertical fluxes through the top faces of the last level computed, which
  // are the bottom faces of the next one: the mass flux (w face) and the
  // theta flux (its base part in the perturbation form), and the upwind
  // fluxes of u, v and w through them.
  struct Carry {
    R m, th, thb, u, v, w;
  } cz{};
  auto level = [&](int k, const Lev& Ln, Lev& Lnn) {
    Q0 Qc;
    if constexpr (PIPE) {
      if (k + 1 < nz) stage_lev(Ln, k + 1);
      Qc = load_q0(k);
      Lnn = load_lev(k + 2 < kcap ? k + 2 : kcap);
    } else {
      if (k + 1 < nz) stage_lev(load_lev(k + 1), k + 1);
      Qc = load_q0(k);
    }
    __syncthreads();
    if (active) {
      // Levels below and above k, clamped to k at the column ends: a value
      // read at the clamped level equals the level-k one, and the fluxes
      // through the column ends are selected away, so the level loop has no
      // per-level branches and its LDS reads issue together.
      const int kb = k > 0 ? k - 1 : k, kt = k + 1 < nz ? k + 1 : k;
      const int s_b = kb & (NSL - 1), s_c = k & (NSL - 1), s_t = kt & (NSL - 1);
      auto SL = [&](int kk) { return kk == k ? s_c : (kk < k ? s_b : s_t); };
      // LDS views at (ii, jj, kk): kk in k-1 .. k+1, (ii, jj) in the tile or ring.
      // (relative to the own entry eo, so neighbour offsets fold into the
      // LDS instructions' immediate offsets)
      auto at = [&](int ii, int jj) { return eo + (ii - i) + PX * (jj - j); };
      auto TH = [&](int ii, int jj, int kk) { return lds(Q_TH, SL(kk), at(ii, jj)); };
      auto P = [&](int ii, int jj, int kk) { return lds(Q_P, SL(kk), at(ii, jj)); };
      auto UC = [&](int ii, int jj, int kk) { return lds(Q_UC, SL(kk), at(ii, jj)); };
      auto VC = [&](int ii, int jj, int kk) { return lds(Q_VC, SL(kk), at(ii, jj)); };
      auto WC = [&](int ii, int jj, int kk) { return lds(Q_WC, SL(kk), at(ii, jj)); };
      // Raw face momenta: staged (FP32) or global; Ug / Vg for the domain-edge
      // faces nx / ny, whose outer neighbour lies beyond the ring.
      auto Ug = [&](int ii, int jj, int kk) { return L.xf(s.u, ii, jj, kk); };
      auto Vg = [&](int ii, int jj, int kk) { return L.yf(s.v, ii, jj, kk); };
      auto U = [&](int ii, int jj, int kk) -> R {
        if constexpr (RAW) return lds(Q_U, SL(kk), at(ii, jj));
        else return L.xf(s.u, ii, jj, kk);
      };
      auto V = [&](int ii, int jj, int kk) -> R {
        if constexpr (RAW) return lds(Q_V, SL(kk), at(ii, jj));
        else return L.yf(s.v, ii, jj, kk);
      };
      auto W = [&](int ii, int jj, int kk) -> R {  // z-face kk < nz
        if constexpr (RAW) return lds(Q_W, SL(kk), at(ii, jj));
        else return L.zf(s.w, ii, jj, kk);
      };
      const R w_kt = W(i, j, kt);
      const R w_up = k + 1 < nz ? w_kt : w_top;  // z-face k+1 of the own cell
      auto IDZC = [&](int ii, int jj, int kk) -> R {
        if (TERRAIN) return a.m.idz_c[mcol(ii, jj) + kk * plane_m];
        return LV(kk, SL_IDZC);
      };
      auto IDZF = [&](int ii, int jj, int kk) -> R {
        if (TERRAIN) return a.m.idz_f[mcol(ii, jj) + kk * plane_m];
        return LV(kk, SL_IDZF);
      };
      // dp/dz at column (ii, jj) (dry_pressure_gradient.cu:37-73).
      auto vgrad = [&](int ii, int jj) -> R {
        if (nz == 1) return R(0);
        const R* zc = a.m.zc;
        const size_t m = mcol(ii, jj);
        if (k == 0) {
          const R dz = zc[m + plane_m] - zc[m];
          return qdiv(P(ii, jj, 1) - P(ii, jj, 0), rmax(dz, R(1.0e-6)));
        }
        if (k + 1 == nz) {
          const R dz = zc[m + k * plane_m] - zc[m + (k - 1) * plane_m];
          return qdiv(P(ii, jj, k) - P(ii, jj, k - 1), rmax(dz, R(1.0e-6)));
        }
        const R dz = zc[m + (k + 1) * plane_m] - zc[m + (k - 1) * plane_m];
        return qdiv(P(ii, jj, k + 1) - P(ii, jj, k - 1), rmax(dz, R(1.0e-6)));
      };

      // The own column's dp/dz and height enter its x face and its y face
      // (and the domain-edge faces): evaluated once, the same arithmetic as
      // at each face.
      R g_own = R(0), z_own = R(0);
      if constexpr (TERRAIN) {
        g_own = vgrad(i, j);
        z_own = a.m.zc[mcol(i, j) + k * plane_m];
      }

      // ---- K2 (dry_core.cu:62-144)
      R t_rho, t_rt;
      {
        const R mw = U(i, j, k), me = U(i + 1, j, k);
        const R ms = V(i, j, k), mn = V(i, j + 1, k);
        // below the chunk's first level the bottom face's values are the
        // top face's of the level before (carried, bit for bit the same)
        const bool carried = k > k0;
        const R mb = carried ? cz.m : W(i, j, k), mt = w_up;
        const R idz = IDZC(i, j, k);
        const R div = qdivc(me - mw, dx, rdx) + qdivc(mn - ms, dy, rdy) + (mt - mb) * idz;
        t_rho = -div;
        const R thc = TH(i, j, k);
        const R thw = TH(i - 1, j, k), the = TH(i + 1, j, k);
        const R ths = TH(i, j - 1, k), thn = TH(i, j + 1, k);
        const R tht = TH(i, j, kt);
        const R fw = mw * (mw >= R(0) ? thw : thc);
        const R fe = me * (me >= R(0) ? thc : the);
        const R fs = ms * (ms >= R(0) ? ths : thc);
        const R fn = mn * (mn >= R(0) ? thc : thn);
        R fb;
        if (carried) {
          fb = cz.th;
        } else {
          const R thb = TH(i, j, kb);
          fb = mb * (mb >= R(0) ? thb : thc);
        }
        const R ft = mt * (mt >= R(0) ? thc : tht);
        cz.m = mt;
        cz.th = ft;
        if (pert) {
          // theta = theta_b(level) + theta': the base parts of the horizontal
          // fluxes share theta_b(k); the vertical ones take the upwind level's.
          const R tb = LV(k, SL_BTH);
          const R mbb = carried ? cz.thb : mb * (mb >= R(0) ? LV(kb, SL_BTH) : tb);
          const R tbt = mt < R(0) ? LV(kt, SL_BTH) : tb;
          const R mtt = mt * tbt;
          cz.thb = mtt;
          t_rt = -(qdivc(tb * (me - mw) + (fe - fw), dx, rdx) +
                   qdivc(tb * (mn - ms) + (fn - fs), dy, rdy) +
                   ((mtt - mbb) + (ft - fb)) * idz);
        } else {
          t_rt = -(qdivc(fe - fw, dx, rdx) + qdivc(fn - fs, dy, rdy) + (ft - fb) * idz);
        }
      }

      // ---- x-face (K17 + K21)
      // (fo: the face's offset, for the base forcing of the terrain
      // perturbation form)
      auto tend_u = [&](int fi, const auto& Uf, R* carry, unsigned fo) -> R {
        // closed outer faces and open-side fluxes are selected away, not
        // branched around: every read stays inside the staged tile and ring
        const bool closed = (fi == 0 && ow) || (fi == g.nx && oe);
        const R qw = Uf(fi - 1, j, k), qc = Uf(fi, j, k), qe = Uf(fi + 1, j, k);
        const R qs = Uf(fi, j - 1, k), qn = Uf(fi, j + 1, k);
        const R qt = Uf(fi, j, kt);
        const R fw = upwind(UC(fi - 1, j, k), qw, qc);
        const R fe = upwind(UC(fi, j, k), qc, qe);
        const R fs0 = upwind(R(0.5) * (VC(fi - 1, j - 1, k) + VC(fi, j - 1, k)), qs, qc);
        const R fn0 = upwind(R(0.5) * (VC(fi - 1, j, k) + VC(fi, j, k)), qc, qn);
        const R fs = (j == 0 && os) ? R(0) : fs0, fn = (j + 1 == g.ny && on) ? R(0) : fn0;
        R fb0;
        if (carry && k > k0) {
          fb0 = *carry;
        } else {
          const R qb = Uf(fi, j, kb);
          fb0 = upwind(R(0.5) * (WC(fi - 1, j, kb) + WC(fi, j, kb)), qb, qc);
        }
        const R ft0 = upwind(R(0.5) * (WC(fi - 1, j, k) + WC(fi, j, k)), qc, qt);
        if (carry) *carry = ft0;
        const R fb = k > 0 ? fb0 : R(0), ft = k + 1 < nz ? ft0 : R(0);
        const R idz = TERRAIN ? R(0.5) * (IDZC(fi - 1, j, k) + IDZC(fi, j, k)) : IDZC(fi, j, k);
        const R A = -(qdivc(fe - fw, dx, rdx) + qdivc(fn - fs, dy, rdy) + (ft - fb) * idz);
        const R pl = P(fi - 1, j, k), pr = P(fi, j, k);
        R term;
        if (TERRAIN) {
          // face fi lies between columns fi-1 and fi, one of them the own
          // (fi == i: the right one; the edge face nx: the left one)
          const bool own_r = fi == i;
          const int nb = own_r ? fi - 1 : fi;
          const R zn = a.m.zc[mcol(nb, j) + k * plane_m], gn = vgrad(nb, j);
          const R zl = own_r ? zn : z_own, zr = own_r ? z_own : zn;
          const R gl = own_r ? gn : g_own, gr = own_r ? g_own : gn;
          const R dzdx = qdivc(zr - zl, dx, rdx);
          term = qdivc(pr - pl, dx, rdx) - dzdx * (R(0.5) * (gl + gr));
          if constexpr (TP) term = term + ld32(a.base.fu, fo);
        } else {
          term = qdivc(pr - pl, dx, rdx);
        }
        return closed ? R(0) : A + (-term);
      };
      // ---- y-face (K18 + K22)
      auto tend_v = [&](int fj, const auto& Vf, R* carry, unsigned fo) -> R {
        const bool closed = (fj == 0 && os) || (fj == g.ny && on);
        const R qw = Vf(i - 1, fj, k), qc = Vf(i, fj, k), qe = Vf(i + 1, fj, k);
        const R qs = Vf(i, fj - 1, k), qn = Vf(i, fj + 1, k);
        const R qt = Vf(i, fj, kt);
        const R fw0 = upwind(R(0.5) * (UC(i - 1, fj - 1, k) + UC(i - 1, fj, k)), qw, qc);
        const R fe0 = upwind(R(0.5) * (UC(i, fj - 1, k) + UC(i, fj, k)), qc, qe);
        const R fw = (i == 0 && ow) ? R(0) : fw0, fe = (i + 1 == g.nx && oe) ? R(0) : fe0;
        const R fs = upwind(VC(i, fj - 1, k), qs, qc);
        const R fn = upwind(VC(i, fj, k), qc, qn);
        R fb0;
        if (carry && k > k0) {
          fb0 = *carry;
        } else {
          const R qb = Vf(i, fj, kb);
          fb0 = upwind(R(0.5) * (WC(i, fj - 1, kb) + WC(i, fj, kb)), qb, qc);
        }
        const R ft0 = upwind(R(0.5) * (WC(i, fj - 1, k) + WC(i, fj, k)), qc, qt);
        if (carry) *carry = ft0;
        const R fb = k > 0 ? fb0 : R(0), ft = k + 1 < nz ? ft0 : R(0);
        const R idz = TERRAIN ? R(0.5) * (IDZC(i, fj - 1, k) + IDZC(i, fj, k)) : IDZC(i, fj, k);
        const R A = -(qdivc(fe - fw, dx, rdx) + qdivc(fn - fs, dy, rdy) + (ft - fb) * idz);
        const R ps = P(i, fj - 1, k), pn = P(i, fj, k);
        R term;
        if (TERRAIN) {
          const bool own_n = fj == j;
          const int nb = own_n ? fj - 1 : fj;
          const R zb = a.m.zc[mcol(i, nb) + k * plane_m], gb = vgrad(i, nb);
          const R zs = own_n ? zb : z_own, zn = own_n ? z_own : zb;
          const R gs = own_n ? gb : g_own, gnn = own_n ? g_own : gb;
          const R dzdy = qdivc(zn - zs, dy, rdy);
          term = qdivc(pn - ps, dy, rdy) - dzdy * (R(0.5) * (gs + gnn));
          if constexpr (TP) term = term + ld32(a.base.fv, fo);
        } else {
          term = qdivc(pn - ps, dy, rdy);
        }
        return closed ? R(0) : A + (-term);
      };
      // ---- z-face at interior face k (K19 + K3)
      auto tend_w = [&]() -> R {
        if (k == 0) return R(0);
        const int kl = k - 1, ku = k;
        const R qw = W(i - 1, j, k), qc = W(i, j, k), qe = W(i + 1, j, k);
        const R qs = W(i, j - 1, k), qn = W(i, j + 1, k);
        const R qt = w_up;
        const R fw0 = upwind(R(0.5) * (UC(i - 1, j, kl) + UC(i - 1, j, ku)), qw, qc);
        const R fe0 = upwind(R(0.5) * (UC(i, j, kl) + UC(i, j, ku)), qc, qe);
        const R fs0 = upwind(R(0.5) * (VC(i, j - 1, kl) + VC(i, j - 1, ku)), qs, qc);
        const R fn0 = upwind(R(0.5) * (VC(i, j, kl) + VC(i, j, ku)), qc, qn);
        const R fw = (i == 0 && ow) ? R(0) : fw0, fe = (i + 1 == g.nx && oe) ? R(0) : fe0;
        const R fs = (j == 0 && os) ? R(0) : fs0, fn = (j + 1 == g.ny && on) ? R(0) : fn0;
        // the level before computed this face's bottom flux as its top one
        // unless it was the chunk's first level below face 1
        R fb;
        if (k > k0 && k > 1) {
          fb = cz.w;
        } else {
          const R qb = W(i, j, k - 1);
          fb = upwind(WC(i, j, kl), qb, qc);
        }
        const R ft = upwind(WC(i, j, ku), qc, qt);
        cz.w = ft;
        const R idz = IDZF(i, j, k);
        const R A = -(qdivc(fe - fw, dx, rdx) + qdivc(fn - fs, dy, rdy) + (ft - fb) * idz);
        const R tf = R(0.5) * (TH(i, j, kl) + TH(i, j, ku));
        if (pert) {
          // theta_ref holds the level means of theta'; the bases cancel in
          // theta_f - theta_ref,f
          R rl = r_km1 + LV(kl, SL_BRHO);
          R ru = r_k + LV(ku, SL_BRHO);
          if constexpr (TP) {  // the cell's base: level value + deviation
            rl = r_km1 + (LV(kl, SL_BRHO) + ld32(a.base.drho, oc0 + static_cast<unsigned>(kl) * pc));
            ru = r_k + (LV(ku, SL_BRHO) + ld32(a.base.drho, oc0 + static_cast<unsigned>(ku) * pc));
          }
          const R trp = R(0.5) * (LV(kl, SL_TREF) + LV(ku, SL_TREF));
          const R trf = R(0.5) * (LV(kl, SL_BTH) + LV(ku, SL_BTH)) + trp;
          const R rf = R(0.5) * (rl + ru);
          const R bb = qdiv(a.gravity * (tf - trp), rmax(trf, R(1.0e-6)));
          return A + rf * bb;
        }
        const R trf = R(0.5) * (LV(kl, SL_TREF) + LV(ku, SL_TREF));
        const R rf = R(0.5) * (r_km1 + r_k);
        const R bb = qdiv(a.gravity * (tf - trf), rmax(trf, R(1.0e-6)));
        return A + rf * bb;
      };

      const R dt = a.dt;
      auto upd = [&](R base, R tend, R floor) { return rmax(floor, base + dt * tend); };
      auto fin = [&](R q0v, R v) { return MODE == 2 ? a.wa * q0v + a.wb * v : v; };

      const unsigned ci = oc0 + static_cast<unsigned>(k) * pc;
      const unsigned xf = ox0 + static_cast<unsigned>(k) * pu;
      const unsigned yf = oy0 + static_cast<unsigned>(k) * pv;
      // the extra faces: x face nx (i == nx-1), y face ny (j == ny-1), z face nz
      const unsigned xfe = xf + 1u, yfe = yf + static_cast<unsigned>(g.pxc), zft = ci + pc;
      const R tu0 = tend_u(i, U, &cz.u, xf), tv0 = tend_v(j, V, &cz.v, yf), tw0 = tend_w();
      if constexpr (MODE == 0) {
        st32(a.out.rho, ci, t_rho);
        st32(a.out.rt, ci, t_rt);
        st32(a.out.u, xf, tu0);
        st32(a.out.v, yf, tv0);
        st32(a.out.w, ci, tw0);
        if (i == g.nx - 1) st32(a.out.u, xfe, tend_u(g.nx, Ug, nullptr, xfe));
        if (j == g.ny - 1) st32(a.out.v, yfe, tend_v(g.ny, Vg, nullptr, yfe));
        if (k == nz - 1) st32(a.out.w, zft, R(0));
      } else {
        const R mf = Phys<R>::mom_floor;
        // floors on the full values (departures: 1e-6 - base)
        R rbk = LV(k, SL_BRHO), tbk = LV(k, SL_BTH), rtbk = LV(k, SL_BRT);
        if constexpr (TP) {  // the own cell's base
          rbk = rbk + ld32(a.base.drho, ci);
          tbk = tbk + ld32(a.base.dth, ci);
          rtbk = rbk * tbk;
        }
        const R flr = pert ? R(1.0e-6) - rbk : R(1.0e-6);
        const R flt = pert ? R(1.0e-6) - rtbk : R(1.0e-6);
        const R rho_new = fin(Qc.rho, upd(r_k, t_rho, flr));
        const R rt_new = fin(Qc.rt, upd(rt_k, t_rt, flt));
        st32(a.out.rho, ci, rho_new);
        if (pert)
          st32(a.out.rt, ci, a.out_theta ? theta_pert(rho_new, rt_new, rbk, tbk) : rt_new);
        else
          st32(a.out.rt, ci, a.out_theta ? qdiv(rt_new, rmax(rho_new, R(1.0e-6))) : rt_new);
        st32(a.out.u, xf, fin(Qc.u, upd(U(i, j, k), tu0, mf)));
        st32(a.out.v, yf, fin(Qc.v, upd(V(i, j, k), tv0, mf)));
        st32(a.out.w, ci, fin(Qc.w, upd(W(i, j, k), tw0, mf)));
        if (i == g.nx - 1)
          st32(a.out.u, xfe,
               fin(MODE == 2 ? ld32(a.q0.u, xfe) : R(0), upd(ld32(s.u, xfe), tend_u(g.nx, Ug, nullptr, xfe), mf)));
        if (j == g.ny - 1)
          st32(a.out.v, yfe,
               fin(MODE == 2 ? ld32(a.q0.v, yfe) : R(0), upd(ld32(s.v, yfe), tend_v(g.ny, Vg, nullptr, yfe), mf)));
        if (k == nz - 1)
          st32(a.out.w, zft, fin(MODE == 2 ? ld32(a.q0.w, zft) : R(0), upd(w_top, R(0), mf)));
      }
    }
    r_km1 = r_k;
    r_k = r_k1;
    rt_k = rt_k1;
  };
  Lev LB{};
  for (int k = k0; k < k1; k += 2) {
    level(k, LA, LB);
    if (k + 1 < k1) level(k + 1, LB, LA);
  }
}

template <class R, bool TERRAIN, bool PERT, int MODE>
__global__ __launch_bounds__(TX * TY) void k_slow_column(SlowArgs<R> a, int tiles_x, int tiles,
                                                         int nkc) {
  slow_body<R, TERRAIN, PERT, MODE>(a, tiles_x, tiles, nkc, hw_block());
}
template <class R, bool TERRAIN, bool PERT, int MODE>
__global__ __launch_bounds__(TX * TY) void k_slow_column_batch(
    const SlowItem<R>* __restrict__ items, BatchMap bm) {
  VBlock vb;
  const SlowItem<R>* it = batch_member(items, bm, vb);
  if (!it) return;
  // the member's arguments, read once before the kernel stores anything
  const SlowArgs<R> a = it->a;
  slow_body<R, TERRAIN, PERT, MODE>(a, it->tiles_x, it->tiles, it->nkc, vb);
}

template <class R, bool TERRAIN, bool PERT>
void launch_slow_mode(hipStream_t st, const SlowArgs<R>& a, dim3 grid, dim3 blk, size_t lds,
                      int tx, int tiles, int nkc) {
  if (recorder()) {
    const void* fn =
        a.mode == 0 ? reinterpret_cast<const void*>(&k_slow_column_batch<R, TERRAIN, PERT, 0>)
        : a.mode == 1 ? reinterpret_cast<const void*>(&k_slow_column_batch<R, TERRAIN, PERT, 1>)
                      : reinterpret_cast<const void*>(&k_slow_column_batch<R, TERRAIN, PERT, 2>);
    record_launch(fn, blk, lds, SlowItem<R>{a, tx, tiles, nkc, 0, 0}, static_cast<int>(grid.x), 1);
    return;
  }
  switch (a.mode) {
    case 0:
      hipLaunchKernelGGL((k_slow_column<R, TERRAIN, PERT, 0>), grid, blk, lds, st, a, tx, tiles,
                         nkc);
      break;
    case 1:
      hipLaunchKernelGGL((k_slow_column<R, TERRAIN, PERT, 1>), grid, blk, lds, st, a, tx, tiles,
                         nkc);
      break;
    default:
      hipLaunchKernelGGL((k_slow_column<R, TERRAIN, PERT, 2>), grid, blk, lds, st, a, tx, tiles,
                         nkc);
      break;
  }
}

template <class R>
void launch_slow_column(hipStream_t st, const SlowArgs<R>& a, bool terrain, int nkc) {
  const int tx = (a.g.nx + TX - 1) / TX, ty = (a.g.ny + TY - 1) / TY;
  const int tiles = tx * ty;
  nkc &= 15;  // launch code: the chunk count (a tile-height field is the fast kernel's)
  nkc = nkc < 1 ? 1 : (nkc > a.g.nz ? a.g.nz : nkc);
  const dim3 blk(TX, TY, 1), grid(tiles * nkc);
  // per-level LDS: the longest chunk plus the levels below and above it
  const size_t lds = static_cast<size_t>((a.g.nz + nkc - 1) / nkc + 3) * SL_N * sizeof(R);
  if (terrain) {
    if constexpr (sizeof(R) == 4) {
      if (a.base.drho) {  // perturbation form on terrain (per-cell bases)
        launch_slow_mode<R, true, true>(st, a, grid, blk, lds, tx, tiles, nkc);
        return;
      }
    }
    launch_slow_mode<R, true, false>(st, a, grid, blk, lds, tx, tiles, nkc);
  } else if (a.base.rho)  // perturbation form (flat, doubly periodic)
    launch_slow_mode<R, false, true>(st, a, grid, blk, lds, tx, tiles, nkc);
  else
    launch_slow_mode<R, false, false>(st, a, grid, blk, lds, tx, tiles, nkc);
}

template void launch_slow_column<float>(hipStream_t, const SlowArgs<float>&, bool, int);
template void launch_slow_column<double>(hipStream_t, const SlowArgs<double>&, bool, int);

// ===========================================================================
// Tracer stage, column march: one SSPRK3 stage of
// advance_passive_tracers_ssprk3 (src/dycore/dry_transport.cu:440-518, K23
// upwind flux form with the positivity clip).  The mixing ratios
// q = (rho q) / max(rho, 1e-6) of every tracer are staged per level over
// the tile plus ring into LDS once, instead of being divided out again at
// each of the seven uses.  mode 1: out = max(0, in - dt div); mode 2:
// out = wa*q0 + wb*(...) (dry_transport.cu:40-47).
// ===========================================================================
template <class R, bool TERRAIN, int NT, int MODE>
__device__ __forceinline__ void tracer_body(const TracerArgs<R>& a, int tiles_x, int tiles, int nkc,
                                            const VBlock& vb) {
  __shared__ R sQ[NT][4][PN];
  extern __shared__ __align__(16) unsigned char lev_raw[];
  R* const lev = reinterpret_cast<R*>(lev_raw);
  const Geo g = a.g;
  int b, k0, k1;
  const int tile = tile_of(vb.x, tiles, nkc, g.nz, b, k0, k1);
  (void)b;
  const int ti0 = (tile % tiles_x) * TX, tj0 = (tile / tiles_x) * TY;
  const int lx = threadIdx.x, ly = threadIdx.y;
  const int t = lx + TX * ly;
  const int i = ti0 + lx, j = tj0 + ly;
  const bool active = i < g.nx && j < g.ny;
  int hx = 0, hy = 0;
  if (t < 2 * PX) {
    hx = (t % PX) - 1;
    hy = t < PX ? -1 : TY;
  } else if (t < NHALO) {
    const int q = t - 2 * PX;
    hx = (q & 1) ? TX : -1;
    hy = q >> 1;
  }
  const int nz = g.nz;
  const bool pert = a.base_rho != nullptr;

  // Per-level values of levels k0-1 .. k1+1: 1/dz (flat) and the density base.
  for (int q = t; q < k1 - k0 + 3; q += TX * TY) {
    const int kk = k0 - 1 + q;
    const int kc = kk < 0 ? 0 : (kk >= nz ? nz - 1 : kk);
    lev[2 * q] = TERRAIN ? R(0) : a.lev_idz_c[kc];
    lev[2 * q + 1] = pert ? a.base_rho[kc] : R(0);
  }
  auto LV = [&](int kk, int q) -> R { return lev[2 * (kk - k0 + 1) + q]; };
  __syncthreads();

  // Loads are unconditional (stand-ins for threads without a ring cell or
  // outside the domain), so waits cover exactly the loads they need: the
  // raw level k+2 and the faces of level k+1 stay in flight while level k is
  // computed.
  const unsigned pc = static_cast<unsigned>(g.pxc) * g.pyc;
  const int oc = stage_pos(g, i, j).c;
  const int hc = t < NHALO ? stage_pos(g, ti0 + hx, tj0 + hy).c : oc;
  const int eo = (lx + 1) + PX * (ly + 1), eh = (hx + 1) + PX * (hy + 1);
  auto ldo = [&](const R* f, int off, int kk) -> R {
    const R v = ld32(f, static_cast<unsigned>(off < 0 ? 0 : off) + static_cast<unsigned>(kk) * pc);
    return off < 0 ? R(0) : v;
  };
  struct TRaw {
    R r, q[NT];
  };
  struct TLev {
    TRaw o, h;  // own cell, ring cell
  };
  auto load_traw = [&](int off, int kk) -> TRaw {
    TRaw x;
    x.r = ldo(a.rho, off, kk);
#pragma unroll
    for (int n = 0; n < NT; ++n) x.q[n] = ldo(a.in[n], off, kk);
    return x;
  };
  auto load_tlev = [&](int kk) -> TLev { return TLev{load_traw(oc, kk), load_traw(hc, kk)}; };
  auto stage_traw = [&](const TRaw& x, int e, int kk) {
    const R r = rmax(pert ? x.r + LV(kk, 1) : x.r, R(1.0e-6));
#pragma unroll
    for (int n = 0; n < NT; ++n) sQ[n][kk & 3][e] = qdiv(x.q[n], r);
  };
  auto stage_tlev = [&](const TLev& x, int kk) {
    stage_traw(x.o, eo, kk);
    if (t < NHALO) stage_traw(x.h, eh, kk);
  };
  // Face momenta (and SSP bases) of one level for the own cell; the w face
  // above a cell is the next level's lower face.  Clamped for threads outside.
  const int ic = i < g.nx ? i : g.nx - 1, jc = j < g.ny ? j : g.ny - 1;
  struct TF {
    R mw, me, ms, mn, mt, q0[NT];
  };
  // Offsets of level k: level-0 offset plus k times the level stride.
  const unsigned tpu = static_cast<unsigned>(g.pxu) * g.pyc;  // x faces
  const unsigned tpv = static_cast<unsigned>(g.pxc) * g.pyv;  // y faces
  const unsigned txf0 = static_cast<unsigned>(g.xf(ic, jc, 0));
  const unsigned tyf0 = static_cast<unsigned>(g.yf(ic, jc, 0));
  const unsigned tcc0 = static_cast<unsigned>(g.cell(ic, jc, 0));
  const unsigned tci0 = static_cast<unsigned>(g.cell(i, j, 0));
  auto load_tf = [&](int k) -> TF {
    TF f{};
    const unsigned xf = txf0 + static_cast<unsigned>(k) * tpu;
    const unsigned yf = tyf0 + static_cast<unsigned>(k) * tpv;
    const unsigned cc = tcc0 + static_cast<unsigned>(k) * pc;
    f.mw = ld32(a.u, xf);
    f.me = ld32(a.u, xf + 1);
    f.ms = ld32(a.v, yf);
    f.mn = ld32(a.v, yf + static_cast<unsigned>(g.pxc));
    f.mt = ld32(a.w, cc + pc);
    if constexpr (MODE == 2) {
#pragma unroll
      for (int n = 0; n < NT; ++n) f.q0[n] = ld32(a.q0[n], cc);
    }
    return f;
  };
  const R dx = static_cast<R>(g.dx), dy = static_cast<R>(g.dy);
  const R rdx = R(1) / dx, rdy = R(1) / dy;
  const int c0 = (lx + 1) + PX * (ly + 1);

  // highest level a chunk stages (uniform; kept in a scalar register so the
  // clamped level indices and their offsets stay scalar work)
  const int kcap = __builtin_amdgcn_readfirstlane(k1 < nz - 1 ? k1 : nz - 1);
  // own raw rho*q of level k (the update's base)
  R own[NT];
  if (k0 > 0) stage_tlev(load_tlev(k0 - 1), k0 - 1);
  {
    const TLev x = load_tlev(k0);
    stage_tlev(x, k0);
#pragma unroll
    for (int n = 0; n < NT; ++n) own[n] = x.o.q[n];
  }
  TLev LA = load_tlev(k0 + 1 < kcap ? k0 + 1 : kcap);
  TF FA = load_tf(k0);
  R mb = a.w[g.zf(ic, jc, k0)];

  // Level k: Ln holds the raw level k+1 and Fc the faces of level k (loaded
  // an iteration ago); raw level k+2 and the faces of k+1 go into Lnn / Fn.
  // Unrolled by two with the buffers swapped, so no register copy waits for
  // loads in flight.
  auto level = [&](int k, const TLev& Ln, const TF& Fc, TLev& Lnn, TF& Fn) {
    if (k + 1 < nz) stage_tlev(Ln, k + 1);
    Lnn = load_tlev(k + 2 < kcap ? k + 2 : kcap);
    Fn = load_tf(k + 1 < k1 ? k + 1 : k1 - 1);
    __syncthreads();
    if (active) {
      const int kb = k > 0 ? k - 1 : k, kt = k + 1 < nz ? k + 1 : k;
      const R idz = TERRAIN ? a.idz_c[g.metric(i, j, k)] : LV(k, 0);
      const unsigned ci = tci0 + static_cast<unsigned>(k) * pc;
#pragma unroll
      for (int n = 0; n < NT; ++n) {
        const R* q = sQ[n][k & 3];
        const R qc = q[c0];
        const R qw = q[c0 - 1], qe = q[c0 + 1], qs = q[c0 - PX], qn = q[c0 + PX];
        const R qb = sQ[n][kb & 3][c0], qt = sQ[n][kt & 3][c0];  // qc at the column ends
        const R fw = Fc.mw * (Fc.mw >= R(0) ? qw : qc);
        const R fe = Fc.me * (Fc.me >= R(0) ? qc : qe);
        const R fs = Fc.ms * (Fc.ms >= R(0) ? qs : qc);
        const R fn = Fc.mn * (Fc.mn >= R(0) ? qc : qn);
        const R fb = mb * (mb >= R(0) ? qb : qc);
        const R ft = Fc.mt * (Fc.mt >= R(0) ? qc : qt);
        const R div = qdivc(fe - fw, dx, rdx) + qdivc(fn - fs, dy, rdy) + (ft - fb) * idz;
        const R v = rmax(R(0), own[n] - a.dt * div);
        st32(a.out[n], ci, MODE == 2 ? a.wa * Fc.q0[n] + a.wb * v : v);
      }
    }
    mb = Fc.mt;
#pragma unroll
    for (int n = 0; n < NT; ++n) own[n] = Ln.o.q[n];
  };
  TLev LB{};
  TF FB{};
  for (int k = k0; k < k1; k += 2) {
    level(k, LA, FA, LB, FB);
    if (k + 1 < k1) level(k + 1, LB, FB, LA, FA);
  }
}

template <class R, bool TERRAIN, int NT, int MODE>
__global__ __launch_bounds__(TX * TY) void k_tracer_column(TracerArgs<R> a, int tiles_x, int tiles,
                                                           int nkc) {
  tracer_body<R, TERRAIN, NT, MODE>(a, tiles_x, tiles, nkc, hw_block());
}
template <class R, bool TERRAIN, int NT, int MODE>
__global__ __launch_bounds__(TX * TY) void k_tracer_column_batch(
    const TracerItem<R>* __restrict__ items, BatchMap bm) {
  VBlock vb;
  const TracerItem<R>* it = batch_member(items, bm, vb);
  if (!it) return;
  // the member's arguments, read once before the kernel stores anything
  const TracerArgs<R> a = it->a;
  tracer_body<R, TERRAIN, NT, MODE>(a, it->tiles_x, it->tiles, it->nkc, vb);
}

template <class R, int NT>
void launch_tracer_column_n(hipStream_t st, const TracerArgs<R>& a, bool terrain, int tx,
                            int tiles, int nkc) {
  const dim3 blk(TX, TY, 1), grid(tiles * nkc);
  // per-level LDS: the longest chunk plus the levels below and above it
  const size_t lds = static_cast<size_t>((a.g.nz + nkc - 1) / nkc + 3) * 2 * sizeof(R);
  const bool m2 = a.mode == 2;
  if (recorder()) {
    const void* fn =
        terrain ? (m2 ? reinterpret_cast<const void*>(&k_tracer_column_batch<R, true, NT, 2>)
                      : reinterpret_cast<const void*>(&k_tracer_column_batch<R, true, NT, 1>))
                : (m2 ? reinterpret_cast<const void*>(&k_tracer_column_batch<R, false, NT, 2>)
                      : reinterpret_cast<const void*>(&k_tracer_column_batch<R, false, NT, 1>));
    record_launch(fn, blk, lds, TracerItem<R>{a, tx, tiles, nkc, 0, 0}, static_cast<int>(grid.x),
                  1);
    return;
  }
  if (terrain) {
    if (m2)
      hipLaunchKernelGGL((k_tracer_column<R, true, NT, 2>), grid, blk, lds, st, a, tx, tiles, nkc);
    else
      hipLaunchKernelGGL((k_tracer_column<R, true, NT, 1>), grid, blk, lds, st, a, tx, tiles, nkc);
  } else {
    if (m2)
      hipLaunchKernelGGL((k_tracer_column<R, false, NT, 2>), grid, blk, lds, st, a, tx, tiles, nkc);
    else
      hipLaunchKernelGGL((k_tracer_column<R, false, NT, 1>), grid, blk, lds, st, a, tx, tiles, nkc);
  }
}

// Tracers are advanced in groups of at most three per launch.
template <class R>
void launch_tracer_column(hipStream_t st, const TracerArgs<R>& a, bool terrain, int nkc) {
  const int tx = (a.g.nx + TX - 1) / TX, ty = (a.g.ny + TY - 1) / TY;
  const int tiles = tx * ty;
  nkc &= 15;  // launch code: the chunk count
  nkc = nkc < 1 ? 1 : (nkc > a.g.nz ? a.g.nz : nkc);
  for (int t0 = 0; t0 < a.ntr; t0 += 3) {
    TracerArgs<R> s = a;
    const int n = a.ntr - t0 < 3 ? a.ntr - t0 : 3;
    for (int q = 0; q < n; ++q) {
      s.in[q] = a.in[t0 + q];
      s.q0[q] = a.q0[t0 + q];
      s.out[q] = a.out[t0 + q];
    }
    s.ntr = n;
    if (n == 3) launch_tracer_column_n<R, 3>(st, s, terrain, tx, tiles, nkc);
    else if (n == 2) launch_tracer_column_n<R, 2>(st, s, terrain, tx, tiles, nkc);
    else launch_tracer_column_n<R, 1>(st, s, terrain, tx, tiles, nkc);
  }
}

template void launch_tracer_column<float>(hipStream_t, const TracerArgs<float>&, bool, int);
template void launch_tracer_column<double>(hipStream_t, const TracerArgs<double>&, bool, int);

}  // namespace gwm
