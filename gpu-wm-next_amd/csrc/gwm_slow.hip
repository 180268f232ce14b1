// gwm_slow.hip -- column-marching slow stage for gfx950.
//
// compute_slow_tendencies (src/dycore/dry_core.cu:400-539) = K2 continuity and
// upwind theta flux + K16-K19 momentum flux divergence + K21-K22 horizontal
// pressure gradient + K3 buoyancy, fused with the stage update / SSP blend of
// advance_dry_state_ssprk3 (dry_core.cu:572-646).  Modes as in the point
// kernel k_slow_stage (gwm_kernels.hip): 0 tendencies, 1 update, 2 update +
// blend; out_theta writes theta instead of rho*theta.
//
// A work-group of 64 x 4 threads marches a 64 x 4 column tile upwards.  For
// each level it stages, over the tile plus a one-cell ring, the derived cell
// quantities every flux needs -- theta, p (dry_thermo.hpp:21-33) and the cell
// velocities (dry_momentum_flux.cu:36-73) -- into LDS once, instead of
// recomputing each of them at every use (a cell's p and velocities enter up
// to a dozen face fluxes).  Four LDS slots per quantity hold levels k-1 ..
// k+1 while level k+1 is staged one iteration ahead: one barrier per level.
// The column is split into level chunks (count chosen by the autotuner) so
// the chip holds more independent work-groups; a chunk stages the level
// below its first one.
// Raw face momenta for the upwind values are read from global memory.
//
// The tracer stage below follows the same pattern.
#include "gwm_kernels.hpp"

namespace gwm {

namespace {

constexpr int TX = 64, TY = 4, PX = TX + 2, PY = TY + 2, PN = PX * PY;
constexpr int NHALO = PN - TX * TY;

// Work-group -> (tile, level chunk [k0, k1)).  Work-group b runs on XCD
// b % 8; the remap gives each XCD a contiguous band of tiles of one chunk.
__device__ inline int tile_of(int tiles, int nkc, int nz, int& b, int& k0, int& k1) {
  const int nb = tiles * nkc;
  b = blockIdx.x;
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

template <class R, bool TERRAIN>
__global__ __launch_bounds__(256) void k_slow_column(SlowArgs<R> a, int tiles_x, int tiles,
                                                     int nkc) {
  __shared__ R sTH[4][PN], sP[4][PN], sUC[4][PN], sVC[4][PN], sWC[4][PN];
  const Geo g = a.g;
  const Loader<R> L{g};
  const DState<R> s = a.s;

  int b, k0, k1;
  const int tile = tile_of(tiles, nkc, g.nz, b, k0, k1);
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
  const bool pert = a.base.rho != nullptr;
  auto stage_one = [&](int ex, int ey, int kk) {
    const int ii = ti0 + ex, jj = tj0 + ey;
    const int e = (ex + 1) + PX * (ey + 1);
    const R r = L.cell(s.rho, ii, jj, kk);
    const R rt = L.cell(s.rt, ii, jj, kk);
    R rr;
    if (pert) {  // theta' and p' of the departures
      const R rb = a.base.rho[kk];
      rr = rmax(rb + r, R(1.0e-6));
      sTH[kk & 3][e] = theta_pert(r, rt, rb, a.base.th[kk]);
      sP[kk & 3][e] = p_pert(rt, a.base.rt[kk], a.base.p[kk]);
    } else {
      rr = rmax(r, R(1.0e-6));
      sTH[kk & 3][e] = d_theta(r, rt);
      sP[kk & 3][e] = d_pressure(r, rt);
    }
    sUC[kk & 3][e] = R(0.5) * (L.xf(s.u, ii, jj, kk) + L.xf(s.u, ii + 1, jj, kk)) / rr;
    sVC[kk & 3][e] = R(0.5) * (L.yf(s.v, ii, jj, kk) + L.yf(s.v, ii, jj + 1, kk)) / rr;
    sWC[kk & 3][e] = R(0.5) * (L.zf(s.w, ii, jj, kk) + L.zf(s.w, ii, jj, kk + 1)) / rr;
  };
  auto stage = [&](int kk) {
    stage_one(lx, ly, kk);
    if (t < NHALO) stage_one(hx, hy, kk);
  };

  const R dx = static_cast<R>(g.dx), dy = static_cast<R>(g.dy);
  const bool ow = !g.per_x, oe = !g.per_x, os = !g.per_y, on = !g.per_y;
  const size_t plane_m = static_cast<size_t>(g.nx) * g.ny;
  auto mcol = [&](int ii, int jj) -> size_t {
    return static_cast<size_t>(L.mj(jj)) * g.nx + L.mi(ii);
  };

  if (k0 > 0) stage(k0 - 1);
  stage(k0);
  for (int k = k0; k < k1; ++k) {
    if (k + 1 < nz) stage(k + 1);
    __syncthreads();
    if (!active) continue;

    // LDS views of the staged quantities at (ii, jj, kk), kk in k-1 .. k+1.
    auto at = [&](int ii, int jj) { return (ii - ti0 + 1) + PX * (jj - tj0 + 1); };
    auto TH = [&](int ii, int jj, int kk) { return sTH[kk & 3][at(ii, jj)]; };
    auto P = [&](int ii, int jj, int kk) { return sP[kk & 3][at(ii, jj)]; };
    auto UC = [&](int ii, int jj, int kk) { return sUC[kk & 3][at(ii, jj)]; };
    auto VC = [&](int ii, int jj, int kk) { return sVC[kk & 3][at(ii, jj)]; };
    auto WC = [&](int ii, int jj, int kk) { return sWC[kk & 3][at(ii, jj)]; };
    auto U = [&](int ii, int jj, int kk) { return L.xf(s.u, ii, jj, kk); };
    auto V = [&](int ii, int jj, int kk) { return L.yf(s.v, ii, jj, kk); };
    auto W = [&](int ii, int jj, int kk) { return L.zf(s.w, ii, jj, kk); };
    auto IDZC = [&](int ii, int jj, int kk) -> R {
      if (TERRAIN) return a.m.idz_c[mcol(ii, jj) + kk * plane_m];
      return a.m.lev_idz_c[kk];
    };
    auto IDZF = [&](int ii, int jj, int kk) -> R {
      if (TERRAIN) return a.m.idz_f[mcol(ii, jj) + kk * plane_m];
      return a.m.lev_idz_f[kk];
    };
    // dp/dz at column (ii, jj) (dry_pressure_gradient.cu:37-73).
    auto vgrad = [&](int ii, int jj) -> R {
      if (nz == 1) return R(0);
      const R* zc = a.m.zc;
      const size_t m = mcol(ii, jj);
      if (k == 0) {
        const R dz = zc[m + plane_m] - zc[m];
        return (P(ii, jj, 1) - P(ii, jj, 0)) / rmax(dz, R(1.0e-6));
      }
      if (k + 1 == nz) {
        const R dz = zc[m + k * plane_m] - zc[m + (k - 1) * plane_m];
        return (P(ii, jj, k) - P(ii, jj, k - 1)) / rmax(dz, R(1.0e-6));
      }
      const R dz = zc[m + (k + 1) * plane_m] - zc[m + (k - 1) * plane_m];
      return (P(ii, jj, k + 1) - P(ii, jj, k - 1)) / rmax(dz, R(1.0e-6));
    };

    // ---- K2 (dry_core.cu:62-144)
    R t_rho, t_rt;
    {
      const R mw = U(i, j, k), me = U(i + 1, j, k);
      const R ms = V(i, j, k), mn = V(i, j + 1, k);
      const R mb = W(i, j, k), mt = W(i, j, k + 1);
      const R idz = IDZC(i, j, k);
      const R div = (me - mw) / dx + (mn - ms) / dy + (mt - mb) * idz;
      t_rho = -div;
      const R thc = TH(i, j, k);
      const R thw = TH(i - 1, j, k), the = TH(i + 1, j, k);
      const R ths = TH(i, j - 1, k), thn = TH(i, j + 1, k);
      const R thb = k > 0 ? TH(i, j, k - 1) : thc;
      const R tht = k + 1 < nz ? TH(i, j, k + 1) : thc;
      const R fw = mw * (mw >= R(0) ? thw : thc);
      const R fe = me * (me >= R(0) ? thc : the);
      const R fs = ms * (ms >= R(0) ? ths : thc);
      const R fn = mn * (mn >= R(0) ? thc : thn);
      const R fb = mb * (mb >= R(0) ? thb : thc);
      const R ft = mt * (mt >= R(0) ? thc : tht);
      if (pert) {
        // theta = theta_b(level) + theta': the base parts of the horizontal
        // fluxes share theta_b(k); the vertical ones take the upwind level's.
        const R tb = a.base.th[k];
        const R tbb = (mb >= R(0) && k > 0) ? a.base.th[k - 1] : tb;
        const R tbt = (mt < R(0) && k + 1 < nz) ? a.base.th[k + 1] : tb;
        t_rt = -((tb * (me - mw) + (fe - fw)) / dx + (tb * (mn - ms) + (fn - fs)) / dy +
                 ((mt * tbt - mb * tbb) + (ft - fb)) * idz);
      } else {
        t_rt = -((fe - fw) / dx + (fn - fs) / dy + (ft - fb) * idz);
      }
    }

    // ---- x-face (K17 + K21)
    auto tend_u = [&](int fi) -> R {
      if ((fi == 0 && ow) || (fi == g.nx && oe)) return R(0);
      const R qw = U(fi - 1, j, k), qc = U(fi, j, k), qe = U(fi + 1, j, k);
      const R qs = U(fi, j - 1, k), qn = U(fi, j + 1, k);
      const R qb = k > 0 ? U(fi, j, k - 1) : qc;
      const R qt = k + 1 < nz ? U(fi, j, k + 1) : qc;
      const R fw = upwind(UC(fi - 1, j, k), qw, qc);
      const R fe = upwind(UC(fi, j, k), qc, qe);
      R fs = 0, fn = 0, fb = 0, ft = 0;
      if (!(j == 0 && os)) fs = upwind(R(0.5) * (VC(fi - 1, j - 1, k) + VC(fi, j - 1, k)), qs, qc);
      if (!(j + 1 == g.ny && on)) fn = upwind(R(0.5) * (VC(fi - 1, j, k) + VC(fi, j, k)), qc, qn);
      if (k > 0) fb = upwind(R(0.5) * (WC(fi - 1, j, k - 1) + WC(fi, j, k - 1)), qb, qc);
      if (k + 1 < nz) ft = upwind(R(0.5) * (WC(fi - 1, j, k) + WC(fi, j, k)), qc, qt);
      const R idz = TERRAIN ? R(0.5) * (IDZC(fi - 1, j, k) + IDZC(fi, j, k)) : IDZC(fi, j, k);
      const R A = -((fe - fw) / dx + (fn - fs) / dy + (ft - fb) * idz);
      const R pl = P(fi - 1, j, k), pr = P(fi, j, k);
      R term;
      if (TERRAIN) {
        const R zl = a.m.zc[mcol(fi - 1, j) + k * plane_m];
        const R zr = a.m.zc[mcol(fi, j) + k * plane_m];
        const R dzdx = (zr - zl) / dx;
        term = (pr - pl) / dx - dzdx * (R(0.5) * (vgrad(fi - 1, j) + vgrad(fi, j)));
      } else {
        term = (pr - pl) / dx;
      }
      return A + (-term);
    };
    // ---- y-face (K18 + K22)
    auto tend_v = [&](int fj) -> R {
      if ((fj == 0 && os) || (fj == g.ny && on)) return R(0);
      const R qw = V(i - 1, fj, k), qc = V(i, fj, k), qe = V(i + 1, fj, k);
      const R qs = V(i, fj - 1, k), qn = V(i, fj + 1, k);
      const R qb = k > 0 ? V(i, fj, k - 1) : qc;
      const R qt = k + 1 < nz ? V(i, fj, k + 1) : qc;
      R fw = 0, fe = 0, fb = 0, ft = 0;
      if (!(i == 0 && ow)) fw = upwind(R(0.5) * (UC(i - 1, fj - 1, k) + UC(i - 1, fj, k)), qw, qc);
      if (!(i + 1 == g.nx && oe)) fe = upwind(R(0.5) * (UC(i, fj - 1, k) + UC(i, fj, k)), qc, qe);
      const R fs = upwind(VC(i, fj - 1, k), qs, qc);
      const R fn = upwind(VC(i, fj, k), qc, qn);
      if (k > 0) fb = upwind(R(0.5) * (WC(i, fj - 1, k - 1) + WC(i, fj, k - 1)), qb, qc);
      if (k + 1 < nz) ft = upwind(R(0.5) * (WC(i, fj - 1, k) + WC(i, fj, k)), qc, qt);
      const R idz = TERRAIN ? R(0.5) * (IDZC(i, fj - 1, k) + IDZC(i, fj, k)) : IDZC(i, fj, k);
      const R A = -((fe - fw) / dx + (fn - fs) / dy + (ft - fb) * idz);
      const R ps = P(i, fj - 1, k), pn = P(i, fj, k);
      R term;
      if (TERRAIN) {
        const R zs = a.m.zc[mcol(i, fj - 1) + k * plane_m];
        const R zn = a.m.zc[mcol(i, fj) + k * plane_m];
        const R dzdy = (zn - zs) / dy;
        term = (pn - ps) / dy - dzdy * (R(0.5) * (vgrad(i, fj - 1) + vgrad(i, fj)));
      } else {
        term = (pn - ps) / dy;
      }
      return A + (-term);
    };
    // ---- z-face at interior face k (K19 + K3)
    auto tend_w = [&]() -> R {
      if (k == 0) return R(0);
      const int kl = k - 1, ku = k;
      const R qw = W(i - 1, j, k), qc = W(i, j, k), qe = W(i + 1, j, k);
      const R qs = W(i, j - 1, k), qn = W(i, j + 1, k);
      const R qb = W(i, j, k - 1), qt = W(i, j, k + 1);
      R fw = 0, fe = 0, fs = 0, fn = 0;
      if (!(i == 0 && ow)) fw = upwind(R(0.5) * (UC(i - 1, j, kl) + UC(i - 1, j, ku)), qw, qc);
      if (!(i + 1 == g.nx && oe)) fe = upwind(R(0.5) * (UC(i, j, kl) + UC(i, j, ku)), qc, qe);
      if (!(j == 0 && os)) fs = upwind(R(0.5) * (VC(i, j - 1, kl) + VC(i, j - 1, ku)), qs, qc);
      if (!(j + 1 == g.ny && on)) fn = upwind(R(0.5) * (VC(i, j, kl) + VC(i, j, ku)), qc, qn);
      const R fb = upwind(WC(i, j, kl), qb, qc);
      const R ft = upwind(WC(i, j, ku), qc, qt);
      const R idz = IDZF(i, j, k);
      const R A = -((fe - fw) / dx + (fn - fs) / dy + (ft - fb) * idz);
      if (pert) {
        // theta_ref holds the level means of theta'; the bases cancel in
        // theta_f - theta_ref,f
        const R rl = L.cell(s.rho, i, j, kl) + a.base.rho[kl];
        const R ru = L.cell(s.rho, i, j, ku) + a.base.rho[ku];
        const R tf = R(0.5) * (TH(i, j, kl) + TH(i, j, ku));
        const R trp = R(0.5) * (a.theta_ref[kl] + a.theta_ref[ku]);
        const R trf = R(0.5) * (a.base.th[kl] + a.base.th[ku]) + trp;
        const R rf = R(0.5) * (rl + ru);
        const R bb = a.gravity * (tf - trp) / rmax(trf, R(1.0e-6));
        return A + rf * bb;
      }
      const R rl = L.cell(s.rho, i, j, kl), ru = L.cell(s.rho, i, j, ku);
      const R tl = L.cell(s.rt, i, j, kl) / rmax(rl, R(1.0e-6));
      const R tu = L.cell(s.rt, i, j, ku) / rmax(ru, R(1.0e-6));
      const R tf = R(0.5) * (tl + tu);
      const R trf = R(0.5) * (a.theta_ref[kl] + a.theta_ref[ku]);
      const R rf = R(0.5) * (rl + ru);
      const R bb = a.gravity * (tf - trf) / rmax(trf, R(1.0e-6));
      return A + rf * bb;
    };

    const R dt = a.dt;
    auto upd = [&](R base, R tend, R floor) { return rmax(floor, base + dt * tend); };
    auto fin = [&](R q0v, R v) { return a.mode == 2 ? a.wa * q0v + a.wb * v : v; };

    const size_t ci = g.cell(i, j, k);
    const R tu0 = tend_u(i), tv0 = tend_v(j), tw0 = tend_w();
    if (a.mode == 0) {
      a.out.rho[ci] = t_rho;
      a.out.rt[ci] = t_rt;
      a.out.u[g.xf(i, j, k)] = tu0;
      a.out.v[g.yf(i, j, k)] = tv0;
      a.out.w[g.zf(i, j, k)] = tw0;
      if (i == g.nx - 1) a.out.u[g.xf(g.nx, j, k)] = tend_u(g.nx);
      if (j == g.ny - 1) a.out.v[g.yf(i, g.ny, k)] = tend_v(g.ny);
      if (k == nz - 1) a.out.w[g.zf(i, j, nz)] = R(0);
      continue;
    }
    const R mf = Phys<R>::mom_floor;
    // floors on the full values (departures: 1e-6 - base)
    const R flr = pert ? R(1.0e-6) - a.base.rho[k] : R(1.0e-6);
    const R flt = pert ? R(1.0e-6) - a.base.rt[k] : R(1.0e-6);
    const R rho_new = fin(a.mode == 2 ? a.q0.rho[ci] : R(0), upd(s.rho[ci], t_rho, flr));
    const R rt_new = fin(a.mode == 2 ? a.q0.rt[ci] : R(0), upd(s.rt[ci], t_rt, flt));
    a.out.rho[ci] = rho_new;
    if (pert)
      a.out.rt[ci] = a.out_theta ? theta_pert(rho_new, rt_new, a.base.rho[k], a.base.th[k])
                                 : rt_new;
    else
      a.out.rt[ci] = a.out_theta ? rt_new / rmax(rho_new, R(1.0e-6)) : rt_new;
    {
      const size_t f = g.xf(i, j, k);
      a.out.u[f] = fin(a.mode == 2 ? a.q0.u[f] : R(0), upd(s.u[f], tu0, mf));
    }
    {
      const size_t f = g.yf(i, j, k);
      a.out.v[f] = fin(a.mode == 2 ? a.q0.v[f] : R(0), upd(s.v[f], tv0, mf));
    }
    {
      const size_t f = g.zf(i, j, k);
      a.out.w[f] = fin(a.mode == 2 ? a.q0.w[f] : R(0), upd(s.w[f], tw0, mf));
    }
    if (i == g.nx - 1) {
      const size_t f = g.xf(g.nx, j, k);
      a.out.u[f] = fin(a.mode == 2 ? a.q0.u[f] : R(0), upd(s.u[f], tend_u(g.nx), mf));
    }
    if (j == g.ny - 1) {
      const size_t f = g.yf(i, g.ny, k);
      a.out.v[f] = fin(a.mode == 2 ? a.q0.v[f] : R(0), upd(s.v[f], tend_v(g.ny), mf));
    }
    if (k == nz - 1) {
      const size_t f = g.zf(i, j, nz);
      a.out.w[f] = fin(a.mode == 2 ? a.q0.w[f] : R(0), upd(s.w[f], R(0), mf));
    }
  }
}

template <class R>
void launch_slow_column(hipStream_t st, const SlowArgs<R>& a, bool terrain, int nkc) {
  const int tx = (a.g.nx + TX - 1) / TX, ty = (a.g.ny + TY - 1) / TY;
  const int tiles = tx * ty;
  nkc = nkc < 1 ? 1 : (nkc > a.g.nz ? a.g.nz : nkc);
  const dim3 blk(TX, TY, 1), grid(tiles * nkc);
  if (terrain)
    hipLaunchKernelGGL((k_slow_column<R, true>), grid, blk, 0, st, a, tx, tiles, nkc);
  else
    hipLaunchKernelGGL((k_slow_column<R, false>), grid, blk, 0, st, a, tx, tiles, nkc);
}

template void launch_slow_column<float>(hipStream_t, const SlowArgs<float>&, bool, int);
template void launch_slow_column<double>(hipStream_t, const SlowArgs<double>&, bool, int);

// ===========================================================================
// Tracer stage, column march: one SSPRK3 stage of
// advance_passive_tracers_ssprk3 (src/dycore/dry_transport.cu:440-518, K23
// upwind flux form with the positivity clip).  The mixing ratios
// q = (rho q) / max(rho, 1e-6) of every tracer are staged per level over
// the tile plus ring into LDS once, instead of being divided out again at
// each of the seven uses.  Same arithmetic per value as k_tracer_stage.
// ===========================================================================
template <class R, bool TERRAIN, int NT>
__global__ __launch_bounds__(256) void k_tracer_column(TracerArgs<R> a, int tiles_x, int tiles,
                                                       int nkc) {
  __shared__ R sQ[NT][4][PN];
  const Geo g = a.g;
  const Loader<R> L{g};
  int b, k0, k1;
  const int tile = tile_of(tiles, nkc, g.nz, b, k0, k1);
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
  auto stage_one = [&](int ex, int ey, int kk) {
    const int ii = ti0 + ex, jj = tj0 + ey;
    const int e = (ex + 1) + PX * (ey + 1);
    const R r0 = L.cell(a.rho, ii, jj, kk);
    const R r = rmax(a.base_rho ? r0 + a.base_rho[kk] : r0, R(1.0e-6));
#pragma unroll
    for (int n = 0; n < NT; ++n) sQ[n][kk & 3][e] = L.cell(a.in[n], ii, jj, kk) / r;
  };
  auto stage = [&](int kk) {
    stage_one(lx, ly, kk);
    if (t < NHALO) stage_one(hx, hy, kk);
  };
  const R dx = static_cast<R>(g.dx), dy = static_cast<R>(g.dy);
  const int c0 = (lx + 1) + PX * (ly + 1);
  if (k0 > 0) stage(k0 - 1);
  stage(k0);
  for (int k = k0; k < k1; ++k) {
    if (k + 1 < nz) stage(k + 1);
    __syncthreads();
    if (!active) continue;
    const R mw = a.u[g.xf(i, j, k)], me = a.u[g.xf(i + 1, j, k)];
    const R ms = a.v[g.yf(i, j, k)], mn = a.v[g.yf(i, j + 1, k)];
    const R mb = a.w[g.zf(i, j, k)], mt = a.w[g.zf(i, j, k + 1)];
    const R idz = TERRAIN ? a.idz_c[g.metric(i, j, k)] : a.lev_idz_c[k];
    const size_t ci = g.cell(i, j, k);
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      const R* q = sQ[n][k & 3];
      const R qc = q[c0];
      const R qw = q[c0 - 1], qe = q[c0 + 1], qs = q[c0 - PX], qn = q[c0 + PX];
      const R qb = k > 0 ? sQ[n][(k - 1) & 3][c0] : qc;
      const R qt = k + 1 < nz ? sQ[n][(k + 1) & 3][c0] : qc;
      const R fw = mw * (mw >= R(0) ? qw : qc);
      const R fe = me * (me >= R(0) ? qc : qe);
      const R fs = ms * (ms >= R(0) ? qs : qc);
      const R fn = mn * (mn >= R(0) ? qc : qn);
      const R fb = mb * (mb >= R(0) ? qb : qc);
      const R ft = mt * (mt >= R(0) ? qc : qt);
      const R div = (fe - fw) / dx + (fn - fs) / dy + (ft - fb) * idz;
      const R v = rmax(R(0), a.in[n][ci] - a.dt * div);
      a.out[n][ci] = a.mode == 2 ? a.wa * a.q0[n][ci] + a.wb * v : v;
    }
  }
}

template <class R, int NT>
void launch_tracer_column_n(hipStream_t st, const TracerArgs<R>& a, bool terrain, int tx,
                            int tiles, int nkc) {
  const dim3 blk(TX, TY, 1), grid(tiles * nkc);
  if (terrain)
    hipLaunchKernelGGL((k_tracer_column<R, true, NT>), grid, blk, 0, st, a, tx, tiles, nkc);
  else
    hipLaunchKernelGGL((k_tracer_column<R, false, NT>), grid, blk, 0, st, a, tx, tiles, nkc);
}

// Tracers are advanced in groups of at most three per launch.
template <class R>
void launch_tracer_column(hipStream_t st, const TracerArgs<R>& a, bool terrain, int nkc) {
  const int tx = (a.g.nx + TX - 1) / TX, ty = (a.g.ny + TY - 1) / TY;
  const int tiles = tx * ty;
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
