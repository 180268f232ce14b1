// gwm_fast.hip -- column-marching fast substep for gfx950.
//
// One substep of apply_local_split_explicit_fast_modes
// (src/dycore/dry_fast_modes.cu:495-652): K10 + K20 pressure from rho and the
// fixed theta, K11 column hydrostatic reference (terrain), K12/K13
// horizontal pressure gradient, K14 vertical perturbation branch, K15
// density from the divergence of (m - m_base), and K10 rho*theta rebuild on
// the last substep.  Arithmetic is the step kernels' q* forms
// (gwm_internal.hpp).
//
// Mapping.  A work-group of 64 x 4 threads owns a 64 x 4 column tile and
// marches it upwards over a chunk of levels [k0, k1), one level per
// iteration.  Pressure of the tile plus a one-cell ring is computed once per
// cell and level into an LDS ring of four planes (levels k-1 .. k+2); the
// plane two levels ahead and the face momenta one level ahead are loaded into
// registers an iteration early, so their global loads overlap the current
// level's arithmetic.  The w face above the cell and the column reference are
// carried in registers, so each face is computed once.  One barrier per
// level.  Splitting the column into chunks (chosen by the autotuner) gives
// the chip more independent work-groups; a chunk recomputes its bottom w face
// from the level below.  The column hydrostatic reference is a recurrence
// from the ground, so terrain runs keep whole columns.  Work-group b is
// remapped so that each XCD (which receives every eighth work-group) owns a
// contiguous band of tiles and the ring rows shared by neighbouring tiles hit
// the same L2.
//
// Faces as departures.  Inside a subcycle the kernel carries the face
// momenta as departures m - m_base from the subcycle's entry faces (K15
// needs only the divergence of m - m_base, and m - m_base starts at zero), so
// a middle substep reads rho, theta and three departure fields and writes
// rho and three departures -- 9 fields instead of the 12 of the full faces
// and their bases.  The first substep reads no faces; the last reads the
// base faces and writes full faces (m_base + departure) and rho*theta.
//
// Flat terrain: the level means of rho and p for the next substep are
// emitted as per-wave FP64 partial sums of the new rho (fixed order: a DPP
// reduction, lane 63 writes), replacing a separate pass over the state.
#include "gwm_kernels.hpp"

namespace gwm {

namespace {

// Tile width: one wave per tile row.  The tile height TY (8 or 4) is a
// template parameter chosen per launch shape by the autotuner.
constexpr int TX = 64, PX = TX + 2;

}  // namespace

// Per-level values of a chunk (levels k0-1 .. k1+1), staged in LDS once per
// work-group so the level loop issues no global loads besides the pipelined
// planes and faces.
enum { LV_PREF = 0, LV_RREF, LV_IDZF, LV_IDZC, LV_BRHO, LV_BTH, LV_BRT, LV_BP, NLV };

// KIND: bit 0 first substep of the subcycle, bit 1 last substep.  DOWN:
// march the columns from the top down (whole columns of flat grids only),
// so that a substep reads first the levels the previous substep, marching
// up, wrote last; every cell's arithmetic is the upward march's, in the
// same order, so the results are identical.
template <class R, bool TERRAIN, bool COLREF, bool PERT, int KIND, int TYV, bool DOWN>
__device__ __forceinline__ void fast_body(const FastArgs<R>& a, int tiles_x, int tiles, int nkc,
                                          double* part_out, const VBlock& vb) {
  static_assert(!DOWN || (!TERRAIN && !COLREF), "the downward march is for flat grids");
  // Terrain perturbation form: per-cell bases, base forcing per face, and
  // K14 without a reference state (the column reference's terms cancel
  // between its recurrence and K14 up to the recurrence's residual, which
  // the z-face base forcing carries, computed on the base column; K14 acts
  // on the departures and adds that forcing).
  constexpr bool TP = TERRAIN && PERT;
  static_assert(!TP || !COLREF, "the terrain perturbation form has no column reference");
  constexpr int TY = TYV, PY = TY + 2, PN = PX * PY;
  constexpr int NHALO = PN - TX * TY;  // ring cells (148 at TY 8, 140 at TY 4)
  static_assert(NHALO <= TX * TY, "one thread per ring cell");
  constexpr bool FIRST = (KIND & 1) != 0, LAST = (KIND & 2) != 0;
  __shared__ R sp[4][PN];
  extern __shared__ __align__(16) unsigned char lev_raw[];
  R* const lev = reinterpret_cast<R*>(lev_raw);
  const Geo g = a.g;
  const Loader<R> L{g};
  const int nz = g.nz;

  // XCD-aware order (work-group b runs on XCD b % 8), then (chunk, tile).
  const int nb = tiles * nkc;
  int b = vb.x;
  if ((nb & 7) == 0) b = (b & 7) * (nb >> 3) + (b >> 3);
  const int chunk = b / tiles, tile = b % tiles;
  const int k0 = (nz * chunk) / nkc, k1 = (nz * (chunk + 1)) / nkc;
  const int ti0 = (tile % tiles_x) * TX, tj0 = (tile / tiles_x) * TY;
  const int lx = threadIdx.x, ly = threadIdx.y;
  const int t = lx + TX * ly;
  const int i = ti0 + lx, j = tj0 + ly;
  const bool active = i < g.nx && j < g.ny;
  const int ci = (lx + 1) + PX * (ly + 1);

  // Ring cell of this thread (t < NHALO): rows -1 and TY, then columns -1 and TX.
  int hx = 0, hy = 0;
  if (t < 2 * PX) {
    hx = (t % PX) - 1;
    hy = t < PX ? -1 : TY;
  } else if (t < NHALO) {
    const int q = t - 2 * PX;
    hx = (q & 1) ? TX : -1;
    hy = q >> 1;
  }
  const int hci = (hx + 1) + PX * (hy + 1);

  // Per-level values of levels k0-1 .. k1+1 (indices clamped to the arrays).
  for (int q = t; q < k1 - k0 + 3; q += TX * TY) {
    const int kk = k0 - 1 + q;
    const int kc = kk < 0 ? 0 : (kk >= nz ? nz - 1 : kk), kf = kk < 0 ? 0 : (kk > nz ? nz : kk);
    R* e = lev + q * NLV;
    if (!COLREF && !TP) {
      e[LV_PREF] = a.pref_lev[kc];
      e[LV_RREF] = a.rref_lev[kc];
    }
    if (!TERRAIN) {
      e[LV_IDZF] = a.m.lev_idz_f[kf];
      e[LV_IDZC] = a.m.lev_idz_c[kc];
    }
    if (PERT) {
      e[LV_BRHO] = a.base.rho[kc];
      e[LV_BTH] = a.base.th[kc];
      e[LV_BRT] = a.base.rt[kc];
      e[LV_BP] = a.base.p[kc];
    }
  }
  auto LV = [&](int kk, int q) -> R { return lev[(kk - k0 + 1) * NLV + q]; };
  __syncthreads();

  const R dx = static_cast<R>(g.dx), dy = static_cast<R>(g.dy);
  const R rdx = R(1) / dx, rdy = R(1) / dy;
  const R dtf = a.dtf, grav = a.gravity;

  // Loads are unconditional (threads without a ring cell, or outside the
  // domain, read a valid stand-in and discard it), so the level loop is
  // straight-line code and each wait covers exactly the loads it needs:
  // the loads of level k+3 planes and k+1 faces stay in flight while level
  // k is computed.
  const unsigned pc = static_cast<unsigned>(g.pxc) * g.pyc;
  const int oc = stage_pos(g, i, j).c;
  const int hc = t < NHALO ? stage_pos(g, ti0 + hx, tj0 + hy).c : oc;
  auto ldc = [&](const R* f, int off, int k) -> R {
    const R v = ld32(f, static_cast<unsigned>(off < 0 ? 0 : off) + static_cast<unsigned>(k) * pc);
    return off < 0 ? R(0) : v;
  };
  // (terrain perturbation form: the base deviations of the own and ring
  // cells too; positions outside an open domain read a stored stand-in)
  auto ldb = [&](const R* f, int off, int k) -> R {
    return ld32(f, static_cast<unsigned>(off < 0 ? 0 : off) + static_cast<unsigned>(k) * pc);
  };
  struct Plane {
    R rc, tc, rh, th;
    R brc, btc, bpc, brh, bth, bph;
  };
  auto load_plane = [&](int k) -> Plane {
    Plane P{ldc(a.rho, oc, k), ldc(a.theta, oc, k), ldc(a.rho, hc, k), ldc(a.theta, hc, k),
            R(0), R(0), R(0), R(0), R(0), R(0)};
    if constexpr (TP) {
      P.brc = ldb(a.base.drho, oc, k);
      P.btc = ldb(a.base.dth, oc, k);
      P.bpc = ldb(a.base.dp, oc, k);
      P.brh = ldb(a.base.drho, hc, k);
      P.bth = ldb(a.base.dth, hc, k);
      P.bph = ldb(a.base.dp, hc, k);
    }
    return P;
  };
  // Perturbation form: rho' and theta' in, p' out (K10 rho*theta = rho*theta_fixed
  // written for the departures: rho_b theta' + rho' (theta_b + theta')); the
  // terrain form adds the cell's base deviations (dr, dt, dp) to the level's.
  auto press = [&](R r, R th, int k, R dr, R dt, R dp) -> R {
    if (!PERT) return q_pressure(r, r * th);
    if constexpr (TP) {
      const R rb = LV(k, LV_BRHO) + dr, tb = LV(k, LV_BTH) + dt;
      return p_pert(rb * th + r * (tb + th), rb * tb, LV(k, LV_BP) + dp);
    }
    const R rb = LV(k, LV_BRHO), tb = LV(k, LV_BTH);
    return p_pert(rb * th + r * (tb + th), LV(k, LV_BRT), LV(k, LV_BP));
  };
  auto stage_plane = [&](const Plane& P, int k) {
    sp[k & 3][ci] = press(P.rc, P.tc, k, P.brc, P.btc, P.bpc);
    if (t < NHALO) sp[k & 3][hci] = press(P.rh, P.th, k, P.brh, P.bth, P.bph);
  };

  // Face departures m - m_base of one level of the own cell: none on the
  // first substep (the faces are the base faces), the previous substep's
  // output otherwise; the base faces on the last substep only, which writes
  // full faces.  Indices clamped into the domain for threads outside it.
  const int ic = i < g.nx ? i : g.nx - 1, jc = j < g.ny ? j : g.ny - 1;
  // Offsets of level k are the level-0 offset plus k times the field's level
  // stride (k is uniform, so the product is scalar work).
  const unsigned sxk = static_cast<unsigned>(g.pyc) * g.pxu;  // x faces
  const unsigned syk = static_cast<unsigned>(g.pyv) * g.pxc;  // y faces
  const unsigned xf_c = static_cast<unsigned>(g.xf(ic, jc, 0));
  const unsigned yf_c = static_cast<unsigned>(g.yf(ic, jc, 0));
  const unsigned zf_c = static_cast<unsigned>(g.zf(ic, jc, 0));
  const unsigned c_o = static_cast<unsigned>(g.cell(i, j, 0));
  const unsigned xf_o = static_cast<unsigned>(g.xf(i, j, 0));
  const unsigned yf_o = static_cast<unsigned>(g.yf(i, j, 0));
  // w1 / wb1: the z face the level computes, k+1 marching up, k marching
  // down (its departure and its base).
  struct Faces {
    R u0, u1, v0, v1, w1, ub0, ub1, vb0, vb1, wb1;
    R fu0, fu1, fv0, fv1, fw1;  // base forcing (terrain perturbation form)
  };
  auto load_faces = [&](int k) -> Faces {
    Faces F{};
    const unsigned xf0 = xf_c + static_cast<unsigned>(k) * sxk;
    const unsigned yf0 = yf_c + static_cast<unsigned>(k) * syk;
    const unsigned yf1 = yf0 + static_cast<unsigned>(g.pxc);
    const unsigned zf1 = zf_c + static_cast<unsigned>(DOWN ? k : k + 1) * pc;
    if constexpr (!FIRST) {
      F.u0 = ld32(a.u, xf0);
      F.u1 = ld32(a.u, xf0 + 1);
      F.v0 = ld32(a.v, yf0);
      F.v1 = ld32(a.v, yf1);
      F.w1 = ld32(a.w, zf1);
    }
    if constexpr (LAST) {
      F.ub0 = ld32(a.ub, xf0);
      F.ub1 = ld32(a.ub, xf0 + 1);
      F.vb0 = ld32(a.vb, yf0);
      F.vb1 = ld32(a.vb, yf1);
      F.wb1 = ld32(a.wb, zf1);
    }
    if constexpr (TP) {
      F.fu0 = ld32(a.base.fu, xf0);
      F.fu1 = ld32(a.base.fu, xf0 + 1);
      F.fv0 = ld32(a.base.fv, yf0);
      F.fv1 = ld32(a.base.fv, yf1);
      F.fw1 = ld32(a.base.fw, zf1);
    }
    return F;
  };

  // Metric column index (periodic wrap of global i, j; dry_fast_modes.cu:24-30).
  auto mcol = [&](int ii, int jj) -> size_t {
    return static_cast<size_t>(L.mj(jj)) * g.nx + L.mi(ii);
  };
  const size_t plane_m = static_cast<size_t>(g.nx) * g.ny;
  const size_t own_m = active ? mcol(i, j) : 0;

  // ---- prologue: planes k0-1 .. k0+1 staged, plane k0+2 and faces of k0
  // in flight, bottom w face.
  const int kmax = nz - 1 < k1 ? nz - 1 : k1;  // highest plane a chunk stages
  // rho_k1 / th_k1 hold the next level of the march (k+1 up, k-1 down);
  // w0 / wb0 the z face carried from the level before (the bottom face k
  // marching up, the top face k+1 marching down).
  R rho_km1 = R(0);
  R rho_k, th_k, rho_k1 = R(0), th_k1 = R(0);
  // own-cell base deviations of levels k and k+1 (terrain perturbation form)
  R br_k = R(0), bt_k = R(0), br_k1 = R(0), bt_k1 = R(0);
  Plane P2;
  Faces F;
  R w0, wb0;
  if constexpr (DOWN) {
    // whole columns: planes nz-1 and nz-2 staged, plane nz-3 and the faces
    // of level nz-1 in flight, the top w face (left as it is).
    {
      const Plane P = load_plane(nz - 1);
      stage_plane(P, nz - 1);
      rho_k = P.rc;
      th_k = P.tc;
    }
    if (nz > 1) {
      const Plane P = load_plane(nz - 2);
      stage_plane(P, nz - 2);
      rho_k1 = P.rc;
      th_k1 = P.tc;
    }
    P2 = load_plane(nz - 3 > 0 ? nz - 3 : 0);
    F = load_faces(nz - 1);
    const size_t zft = g.zf(ic, jc, nz);
    w0 = FIRST ? R(0) : a.w[zft];
    wb0 = LAST ? a.wb[zft] : R(0);
  } else {
    if (k0 > 0) {
      const Plane P = load_plane(k0 - 1);
      stage_plane(P, k0 - 1);
      rho_km1 = P.rc;
    }
    {
      const Plane P = load_plane(k0);
      stage_plane(P, k0);
      rho_k = P.rc;
      th_k = P.tc;
      br_k = P.brc;
      bt_k = P.btc;
    }
    if (k0 + 1 < nz) {
      const Plane P = load_plane(k0 + 1);
      stage_plane(P, k0 + 1);
      rho_k1 = P.rc;
      th_k1 = P.tc;
      br_k1 = P.brc;
      bt_k1 = P.btc;
    }
    P2 = load_plane(k0 + 2 < kmax ? k0 + 2 : kmax);
    F = load_faces(k0);
    const size_t zf0 = g.zf(ic, jc, k0);
    w0 = FIRST ? R(0) : a.w[zf0];
    wb0 = LAST ? a.wb[zf0] : R(0);
  }
  __syncthreads();

  // Column reference (K11) at level k0 (whole columns only: k0 == 0).
  R pref_k = R(0), rref_k = R(0);
  if (COLREF) {
    const R rt0 = rho_k * th_k;
    pref_k = q_pressure(rho_k, rt0);
    rref_k = q_rho_from_p_theta(pref_k, q_theta(rho_k, rt0));
  }

  const bool ow = !g.per_x, os = !g.per_y;
  if (!DOWN && active && k0 > 0) {
    // K14 at face k0 from levels k0-1 and k0 (flat references; the terrain
    // perturbation form: departures plus the base forcing of the face).
    const R idz = TERRAIN ? a.m.idz_f[own_m + k0 * plane_m] : LV(k0, LV_IDZF);
    if constexpr (TP) {
      const R ppl = sp[(k0 - 1) & 3][ci], ppu = sp[k0 & 3][ci];
      const R rpf = R(0.5) * (rho_km1 + rho_k);
      const R tend = -(ppu - ppl) * idz - grav * rpf;
      w0 = w0 + dtf * (tend + a.base.fw[g.zf(ic, jc, k0)]);
    } else {
      const R ppl = sp[(k0 - 1) & 3][ci] - LV(k0 - 1, LV_PREF);
      const R ppu = sp[k0 & 3][ci] - LV(k0, LV_PREF);
      const R rpf = R(0.5) * ((rho_km1 - LV(k0 - 1, LV_RREF)) + (rho_k - LV(k0, LV_RREF)));
      const R tend = -(ppu - ppl) * idz - grav * rpf;
      w0 = w0 + dtf * tend;
    }
  }

  // Level k: P2 / F hold plane k+2 and the faces of level k (loaded an
  // iteration ago); the loads of plane k+3 and of the faces of level k+1 go
  // into P3 / Fn first.  The loop is unrolled by two with the buffers
  // swapped, so no register copy waits for loads in flight.
  auto level = [&](int k, const Plane& P2, const Faces& F, Plane& P3, Faces& Fn) {
    // Stage the plane two levels ahead in the march (k+2 up, k-2 down).
    if constexpr (DOWN) {
      P3 = load_plane(k - 3 > 0 ? k - 3 : 0);
      Fn = load_faces(k - 1 > 0 ? k - 1 : 0);
      if (k - 2 >= 0) stage_plane(P2, k - 2);
    } else {
      P3 = load_plane(k + 3 < kmax ? k + 3 : kmax);
      Fn = load_faces(k + 1 < k1 ? k + 1 : k1 - 1);
      if (k + 2 <= kmax) stage_plane(P2, k + 2);
    }
    const R rho_k2 = P2.rc, th_k2 = P2.tc;
    const R br_k2 = P2.brc, bt_k2 = P2.btc;

    const R* pk = sp[k & 3];
    const R* pdn = sp[(k + 3) & 3];  // level k-1
    const R* pup = sp[(k + 1) & 3];  // level k+1

    // Column reference at level k+1 (K11 recurrence, own column).
    R pref_k1 = R(0), rref_k1 = R(0);
    if (COLREF && k + 1 < nz && active) {
      const R thk = q_theta(rho_k1, rho_k1 * th_k1);
      const R dz = a.m.zc[own_m + (k + 1) * plane_m] - a.m.zc[own_m + k * plane_m];
      R pn = rmax(R(1), pref_k - grav * rref_k * dz);
      for (int it = 0; it < 4; ++it) {
        const R rn = q_rho_from_p_theta(pn, thk);
        pn = pref_k - grav * R(0.5) * (rref_k + rn) * dz;
        pn = rmax(pn, R(1));
      }
      pref_k1 = pn;
      rref_k1 = q_rho_from_p_theta(pn, thk);
    }

    R rho_new = R(0), rt_new = R(0);
    if (active) {
      // dp/dz at a neighbour column, from the LDS planes (dry_pressure_gradient.cu:37-73).
      auto vgrad = [&](int off, size_t m) -> R {
        if (nz == 1) return R(0);
        const R* zc = a.m.zc;
        if (k == 0) {
          const R dz = zc[m + plane_m] - zc[m];
          return qdiv(pup[off] - pk[off], rmax(dz, R(1.0e-6)));
        }
        if (k + 1 == nz) {
          const R dz = zc[m + k * plane_m] - zc[m + (k - 1) * plane_m];
          return qdiv(pk[off] - pdn[off], rmax(dz, R(1.0e-6)));
        }
        const R dz = zc[m + (k + 1) * plane_m] - zc[m + (k - 1) * plane_m];
        return qdiv(pup[off] - pdn[off], rmax(dz, R(1.0e-6)));
      };
      // The own column's dp/dz and height enter all four faces: evaluated
      // once (the same arithmetic as at each face).
      R g_own = R(0), z_own = R(0);
      if constexpr (TERRAIN) {
        g_own = vgrad(ci, own_m);
        z_own = a.m.zc[own_m + k * plane_m];
      }
      auto pgf = [&](int offl, int offr, int il, int jl, int ir, int jr, R d, R rd) -> R {
        const R pl = pk[offl], pr = pk[offr];
        if (!TERRAIN) return qdivc(pr - pl, d, rd);
        const bool own_l = offl == ci;  // the own cell is the left or the right one
        const size_t mn = own_l ? mcol(ir, jr) : mcol(il, jl);
        const R zn = a.m.zc[mn + k * plane_m], gn = vgrad(own_l ? offr : offl, mn);
        const R zl = own_l ? z_own : zn, zr = own_l ? zn : z_own;
        const R gl = own_l ? g_own : gn, gr = own_l ? gn : g_own;
        const R dzd = qdivc(zr - zl, d, rd);
        return qdivc(pr - pl, d, rd) - dzd * (R(0.5) * (gl + gr));
      };
      // K12 / K13: faces i, i+1 and j, j+1 of the cell (the terrain
      // perturbation form adds the base's forcing of the face).
      auto fb = [&](R pg, R f) -> R {
        if constexpr (TP) return pg + f;
        return pg;
      };
      R u0 = F.u0, u1 = F.u1, v0 = F.v0, v1 = F.v1;
      if (a.horizontal) {
        if (!(i == 0 && ow)) u0 = F.u0 + (-dtf * fb(pgf(ci - 1, ci, i - 1, j, i, j, dx, rdx), F.fu0));
        if (!(i + 1 == g.nx && ow))
          u1 = F.u1 + (-dtf * fb(pgf(ci, ci + 1, i, j, i + 1, j, dx, rdx), F.fu1));
        if (!(j == 0 && os)) v0 = F.v0 + (-dtf * fb(pgf(ci - PX, ci, i, j - 1, i, j, dy, rdy), F.fv0));
        if (!(j + 1 == g.ny && os))
          v1 = F.v1 + (-dtf * fb(pgf(ci, ci + PX, i, j, i, j + 1, dy, rdy), F.fv1));
      }
      // K14 (faces 0 and nz are left as they are): marching up, w face k+1
      // from levels k and k+1; marching down, w face k from levels k-1 and
      // k (the same arithmetic for a face either way).
      auto k14 = [&](R wf, R p_lo, R p_hi, R pl, R pu, R r_lo, R r_hi, R rl, R ru, R idz) -> R {
        if constexpr (TP) {  // departures and the base forcing of the face
          const R rpf = R(0.5) * (r_lo + r_hi);
          const R tend = -(p_hi - p_lo) * idz - grav * rpf;
          return wf + dtf * (tend + F.fw1);
        }
        const R ppl = p_lo - pl;
        const R ppu = p_hi - pu;
        const R rpf = R(0.5) * ((r_lo - rl) + (r_hi - ru));
        const R tend = -(ppu - ppl) * idz - grav * rpf;
        return wf + dtf * tend;
      };
      R w1 = F.w1;  // the face this level computes
      if constexpr (DOWN) {
        if (k > 0)
          w1 = k14(F.w1, pdn[ci], pk[ci], LV(k - 1, LV_PREF), LV(k, LV_PREF), rho_k1, rho_k,
                   LV(k - 1, LV_RREF), LV(k, LV_RREF), LV(k, LV_IDZF));
      } else if (k + 1 < nz) {
        R rl, pl, ru, pu;
        if (COLREF) {
          rl = rref_k; pl = pref_k; ru = rref_k1; pu = pref_k1;
        } else if (TP) {
          rl = pl = ru = pu = R(0);
        } else {
          rl = LV(k, LV_RREF); pl = LV(k, LV_PREF); ru = LV(k + 1, LV_RREF); pu = LV(k + 1, LV_PREF);
        }
        const R idz = TERRAIN ? a.m.idz_f[own_m + (k + 1) * plane_m] : LV(k + 1, LV_IDZF);
        w1 = k14(F.w1, pk[ci], pup[ci], pl, pu, rho_k, rho_k1, rl, ru, idz);
      }
      // the bottom and top w faces of the cell
      const R wlo = DOWN ? w1 : w0, whi = DOWN ? w0 : w1;
      // K15
      rho_new = rho_k;
      if (a.density) {
        const R idz = TERRAIN ? a.m.idz_c[own_m + k * plane_m] : LV(k, LV_IDZC);
        // divergence of (m - m_base): the faces are departures already
        const R div = qdivc(u1 - u0, dx, rdx) + qdivc(v1 - v0, dy, rdy) + (whi - wlo) * idz;
        const R floor = PERT ? R(1.0e-6) - (TP ? LV(k, LV_BRHO) + br_k : LV(k, LV_BRHO)) : R(1.0e-6);
        rho_new = rmax(floor, rho_new - dtf * div);
      }
      const unsigned c0 = c_o + static_cast<unsigned>(k) * pc;
      const unsigned xf0 = xf_o + static_cast<unsigned>(k) * sxk;
      const unsigned yf0 = yf_o + static_cast<unsigned>(k) * syk;
      st32(a.rho_out, c0, rho_new);
      if constexpr (TP) {
        const R rb = LV(k, LV_BRHO) + br_k, tb = LV(k, LV_BTH) + bt_k;
        rt_new = rb * th_k + rho_new * (tb + th_k);
      } else {
        rt_new = PERT ? LV(k, LV_BRHO) * th_k + rho_new * (LV(k, LV_BTH) + th_k) : rho_new * th_k;
      }
      if (a.rt_out) st32(a.rt_out, c0, rt_new);
      // departures, or full faces base + departure on the last substep
      st32(a.u_out, xf0, LAST ? F.ub0 + u0 : u0);
      st32(a.v_out, yf0, LAST ? F.vb0 + v0 : v0);
      // w face k (and face nz on the top level), with its base on the last
      // substep: the carried face marching up, the computed one marching down
      const R wb_lo = DOWN ? F.wb1 : wb0, wb_hi = DOWN ? wb0 : F.wb1;
      st32(a.w_out, c0, LAST ? wb_lo + wlo : wlo);
      if (i == g.nx - 1) st32(a.u_out, xf0 + 1, LAST ? F.ub1 + u1 : u1);
      if (j == g.ny - 1) st32(a.v_out, yf0 + static_cast<unsigned>(g.pxc), LAST ? F.vb1 + v1 : v1);
      if (k == nz - 1) st32(a.w_out, c0 + pc, LAST ? wb_hi + whi : whi);
      w0 = w1;
    }
    wb0 = F.wb1;

    if (part_out) {  // level sums of the next substep (flat refs, K6/K7)
      // A row's 64 values are summed in FP64, except in the FP32
      // perturbation form: its values are departures from the level means,
      // small and of either sign, whose FP32 sum (error below 4e-7 of the
      // sum of their magnitudes) leaves the mean's error far below the
      // departures' own rounding; the partial is stored in FP64.
      using S = typename std::conditional<PERT && sizeof(R) == 4, float, double>::type;
      S sr = S(0), spp = S(0);
      if (active) {
        if (a.part_theta) {  // theta_ref of the next stage (dry_core.cu:462-490)
          sr = static_cast<S>(PERT ? theta_pert(rho_new, rt_new, LV(k, LV_BRHO), LV(k, LV_BTH))
                                   : q_theta(rho_new, rt_new));
        } else {
          sr = static_cast<S>(rho_new);
          spp = static_cast<S>(press(rho_new, th_k, k, R(0), R(0), R(0)));
        }
      }
      sr = wave_sum_to63(sr);
      spp = wave_sum_to63(spp);
      // one partial per wave (a 64-column row segment), at (level, tile
      // column, row): the layout and so the finalize's summation order do
      // not depend on the tile height, and a work-group's waves write
      // consecutive slots
      if (lx == TX - 1 && j < g.ny) {
        const size_t nw = static_cast<size_t>(g.ny) * tiles_x;
        const size_t slot = (static_cast<size_t>(k) * nw +
                             static_cast<size_t>(tile % tiles_x) * g.ny + static_cast<size_t>(j)) * 2;
        gstore(part_out, slot, static_cast<double>(sr));
        gstore(part_out, slot + 1, static_cast<double>(spp));
      }
    }

    rho_k = rho_k1;
    th_k = th_k1;
    rho_k1 = rho_k2;
    th_k1 = th_k2;
    br_k = br_k1;
    bt_k = bt_k1;
    br_k1 = br_k2;
    bt_k1 = bt_k2;
    pref_k = pref_k1;
    rref_k = rref_k1;
    __syncthreads();
  };
  Plane P3;
  Faces Fn;
  if constexpr (DOWN) {
    for (int k = nz - 1; k >= 0; k -= 2) {
      level(k, P2, F, P3, Fn);
      if (k - 1 >= 0) level(k - 1, P3, Fn, P2, F);
    }
    return;
  }
  for (int k = k0; k < k1; k += 2) {
    level(k, P2, F, P3, Fn);
    if (k + 1 < k1) level(k + 1, P3, Fn, P2, F);
  }
}

template <class R, bool TERRAIN, bool COLREF, bool PERT, int KIND, int TYV, bool DOWN>
__global__ __launch_bounds__(TX * TYV) void k_fast_column(FastArgs<R> a, int tiles_x, int tiles,
                                                          int nkc, double* part_out) {
  fast_body<R, TERRAIN, COLREF, PERT, KIND, TYV, DOWN>(a, tiles_x, tiles, nkc, part_out,
                                                       hw_block());
}
template <class R, bool TERRAIN, bool COLREF, bool PERT, int KIND, int TYV, bool DOWN>
__global__ __launch_bounds__(TX * TYV) void k_fast_column_batch(
    const FastItem<R>* __restrict__ items, BatchMap bm) {
  VBlock vb;
  const FastItem<R>* it = batch_member(items, bm, vb);
  if (!it) return;
  // the member's arguments, read once before the kernel stores anything
  const FastArgs<R> a = it->a;
  fast_body<R, TERRAIN, COLREF, PERT, KIND, TYV, DOWN>(a, it->tiles_x, it->tiles, it->nkc,
                                                       it->part_out, vb);
}

template <class R, bool TERRAIN, bool COLREF, bool PERT, int TYV, bool DOWN = false>
void launch_fast_kind(hipStream_t st, const FastArgs<R>& a, int nkc, double* part_out) {
  const int tx = (a.g.nx + TX - 1) / TX, ty = (a.g.ny + TYV - 1) / TYV;
  const int tiles = tx * ty;
  const dim3 grid(tiles * nkc), blk(TX, TYV, 1);
  // per-level LDS: the longest chunk plus the levels below and above it
  const size_t lds = static_cast<size_t>((a.g.nz + nkc - 1) / nkc + 3) * NLV * sizeof(R);
  const int kind = (a.first ? 1 : 0) | (a.last ? 2 : 0);
  if (recorder()) {
    const void* fns[4] = {
        reinterpret_cast<const void*>(&k_fast_column_batch<R, TERRAIN, COLREF, PERT, 0, TYV, DOWN>),
        reinterpret_cast<const void*>(&k_fast_column_batch<R, TERRAIN, COLREF, PERT, 1, TYV, DOWN>),
        reinterpret_cast<const void*>(&k_fast_column_batch<R, TERRAIN, COLREF, PERT, 2, TYV, DOWN>),
        reinterpret_cast<const void*>(&k_fast_column_batch<R, TERRAIN, COLREF, PERT, 3, TYV, DOWN>)};
    record_launch(fns[kind], blk, lds, FastItem<R>{a, tx, tiles, nkc, part_out, 0, 0},
                  static_cast<int>(grid.x), 1);
    return;
  }
  switch (kind) {
    case 0:
      hipLaunchKernelGGL((k_fast_column<R, TERRAIN, COLREF, PERT, 0, TYV, DOWN>), grid, blk, lds, st, a,
                         tx, tiles, nkc, part_out);
      break;
    case 1:
      hipLaunchKernelGGL((k_fast_column<R, TERRAIN, COLREF, PERT, 1, TYV, DOWN>), grid, blk, lds, st, a,
                         tx, tiles, nkc, part_out);
      break;
    case 2:
      hipLaunchKernelGGL((k_fast_column<R, TERRAIN, COLREF, PERT, 2, TYV, DOWN>), grid, blk, lds, st, a,
                         tx, tiles, nkc, part_out);
      break;
    default:
      hipLaunchKernelGGL((k_fast_column<R, TERRAIN, COLREF, PERT, 3, TYV, DOWN>), grid, blk, lds, st, a,
                         tx, tiles, nkc, part_out);
      break;
  }
}

template <class R>
void launch_fast_column(hipStream_t st, const FastArgs<R>& a, bool terrain, int code,
                        double* part_out, int* nparts_out) {
  // code: level chunks (low 4 bits), tile height (bit 4 set: 4 rows, else
  // 8; flat launches without the column reference only), bit 6 set: every
  // substep marches upwards (a.down ignored)
  int nkc = code & 15;
  const bool down = a.down && (code & 64) == 0;
  // 4-row tiles: flat launches, and the column-reference launches, whose
  // whole columns give few work-groups (one per tile)
  const bool ty4 = ((code >> 4) & 1) != 0 && (a.col_ref || !terrain);
  if (nparts_out) *nparts_out = fast_column_parts(a.g.nx, a.g.ny);
  if (a.col_ref || nkc < 1) nkc = 1;
  if (nkc > a.g.nz) nkc = a.g.nz;
  if (a.col_ref) {
    if (terrain) {
      if (ty4) launch_fast_kind<R, true, true, false, 4>(st, a, 1, nullptr);
      else launch_fast_kind<R, true, true, false, 8>(st, a, 1, nullptr);
    } else {
      if (ty4) launch_fast_kind<R, false, true, false, 4>(st, a, 1, nullptr);
      else launch_fast_kind<R, false, true, false, 8>(st, a, 1, nullptr);
    }
  } else if (terrain) {
    if constexpr (sizeof(R) == 4) {
      if (a.base.drho) {  // perturbation form on terrain (per-cell bases)
        launch_fast_kind<R, true, false, true, 8>(st, a, nkc, part_out);
        return;
      }
    }
    launch_fast_kind<R, true, false, false, 8>(st, a, nkc, part_out);
  } else if (down && nkc == 1) {
    // whole columns marched from the top down (flat grids)
    if (a.base.rho) {
      if (ty4) launch_fast_kind<R, false, false, true, 4, true>(st, a, 1, part_out);
      else launch_fast_kind<R, false, false, true, 8, true>(st, a, 1, part_out);
    } else {
      if (ty4) launch_fast_kind<R, false, false, false, 4, true>(st, a, 1, part_out);
      else launch_fast_kind<R, false, false, false, 8, true>(st, a, 1, part_out);
    }
  } else if (a.base.rho) {  // perturbation form (flat, doubly periodic)
    if (ty4) launch_fast_kind<R, false, false, true, 4>(st, a, nkc, part_out);
    else launch_fast_kind<R, false, false, true, 8>(st, a, nkc, part_out);
  } else {
    if (ty4) launch_fast_kind<R, false, false, false, 4>(st, a, nkc, part_out);
    else launch_fast_kind<R, false, false, false, 8>(st, a, nkc, part_out);
  }
}

// Level partials per level: one per 64-column row segment.
int fast_column_parts(int nx, int ny) { return ((nx + TX - 1) / TX) * ny; }

template void launch_fast_column<float>(hipStream_t, const FastArgs<float>&, bool, int, double*,
                                        int*);
template void launch_fast_column<double>(hipStream_t, const FastArgs<double>&, bool, int,
                                         double*, int*);

}  // namespace gwm
