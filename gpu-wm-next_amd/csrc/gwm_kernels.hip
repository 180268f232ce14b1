// gwm_kernels.hip -- gfx950 kernels of the dry + warm-rain time step.
//
// Each kernel states the reference operator(s) it covers.  Semantics are the
// reference's 1x1 layout: a neighbour outside the owned range is the value
// the reference's halo exchange places there (map_cell / map_face), so no
// kernel depends on halo storage.  Periodic wrap faces (face 0 and face n of
// a face field) are kept equal as a state invariant, which makes the
// reference's owned-interface averaging (halo_exchange.cpp:217-285) an
// identity inside the step; gwm_sync_interfaces() establishes it on entry.
#include "gwm_kernels.hpp"
#include "gwm_micro.hpp"

#include <mutex>

namespace gwm {

namespace {

// Deterministic block sum of a double (256 threads).
__device__ inline double block_sum(double v, double* sh) {
  const int t = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
  const int n = blockDim.x * blockDim.y * blockDim.z;
  sh[t] = v;
  __syncthreads();
  for (int s = n / 2; s > 0; s >>= 1) {
    if (t < s) sh[t] += sh[t + s];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

}  // namespace

// ===========================================================================
// Level sums (deterministic).  mode 0: theta = rho_theta/max(rho,1e-6) of a
// state (dry_core.cu:462-490); mode 1: rho and p(rho, rho*theta_fixed) of the
// fast subcycle (dry_fast_modes.cu:42-67, FP32 atomics replaced by FP64
// partials in fixed order).  Block = 256 threads over a chunk of rows of one
// level; partials[(k*nchunk + c)*2 + {0,1}].
// ===========================================================================
constexpr int kPartialRows = 8;  // rows of one level per work-group

template <class R>
__device__ __forceinline__ void level_partials_body(const Geo& g, const R* rho, const R* aux,
                                                    int mode, double* partials,
                                                    const Base<R>& base, const VBlock& vb) {
  __shared__ double sh[256];
  const int k = vb.y;
  const int c = vb.x;
  const int nchunk = vb.gx;
  const int j0 = c * kPartialRows;
  const int t = threadIdx.x;
  double s0 = 0.0, s1 = 0.0;
  const bool pert = base.rho != nullptr;
  const R rb = pert ? base.rho[k] : R(0), thb = pert ? base.th[k] : R(0);
  const R rtb = pert ? base.rt[k] : R(0), pb = pert ? base.p[k] : R(0);
  // Each thread loads a column segment of the chunk's rows at once (all
  // loads in flight), then accumulates in row order.
  for (int i = t; i < g.nx; i += 256) {
    R rv[kPartialRows], av[kPartialRows];
#pragma unroll
    for (int q = 0; q < kPartialRows; ++q) {
      rv[q] = R(0);
      av[q] = R(0);
      if (j0 + q < g.ny) {
        const size_t c0 = g.cell(i, j0 + q, k);
        rv[q] = rho[c0];
        av[q] = aux[c0];
      }
    }
#pragma unroll
    for (int q = 0; q < kPartialRows; ++q) {
      if (j0 + q >= g.ny) break;
      const R r = rv[q];
      if (pert && base.drho) {
        // terrain form: theta relative to the level's base value (the
        // cell's base deviation plus its departure), for theta_ref
        const size_t c0 = g.cell(i, j0 + q, k);
        const R dt = base.dth[c0];
        s0 += static_cast<double>(dt + theta_pert(r, av[q], rb + base.drho[c0], thb + dt));
      } else if (pert) {
        // departures: theta' of (rho', (rho theta)'), or rho' and p' of
        // (rho', theta' fixed)
        if (mode == 0) {
          s0 += static_cast<double>(theta_pert(r, av[q], rb, thb));
        } else {
          const R tp = av[q];
          s0 += static_cast<double>(r);
          s1 += static_cast<double>(p_pert(rb * tp + r * (thb + tp), rtb, pb));
        }
      } else if (mode == 0) {
        s0 += static_cast<double>(q_theta(r, av[q]));
      } else {
        s0 += static_cast<double>(r);
        s1 += static_cast<double>(q_pressure(r, r * av[q]));
      }
    }
  }
  s0 = block_sum(s0, sh);
  s1 = block_sum(s1, sh);
  if (t == 0) {
    partials[(static_cast<size_t>(k) * nchunk + c) * 2 + 0] = s0;
    partials[(static_cast<size_t>(k) * nchunk + c) * 2 + 1] = s1;
  }
}

template <class R>
__global__ __launch_bounds__(256) void k_level_partials(Geo g, const R* rho, const R* aux,
                                                         int mode, double* partials,
                                                         Base<R> base) {
  level_partials_body(g, rho, aux, mode, partials, base, hw_block());
}
template <class R>
__global__ __launch_bounds__(256) void k_level_partials_batch(const PartialsItem<R>* items,
                                                               BatchMap bm) {
  VBlock vb;
  const PartialsItem<R>* it = batch_member(items, bm, vb);
  if (!it) return;
  const PartialsItem<R> a = *it;
  level_partials_body(a.g, a.rho, a.aux, a.mode, a.partials, a.base, vb);
}

// One 1024-thread block per level: thread t sums partials t, t+1024, ... in
// order over four interleaved accumulators (a level's 4096 partials of the
// 512 x 512 grid in one round of four loads per thread: the kernel is bound
// by load latency, not bytes), a fixed DPP reduction adds each wave's lanes
// (total in lane 63) and thread 0 adds the sixteen waves in a fixed tree, so
// the result does not depend on scheduling.  The kernel runs between every
// two fast substeps, so its length is the round trips it waits for.
constexpr int kFinalizeThreads = 1024;
constexpr int kFinalizeAcc = 4;

template <class R>
__device__ __forceinline__ void level_finalize_body(int nchunk, const double* partials, int mode,
                                                    double count, R* out0, R* out1,
                                                    const VBlock& vb) {
  __shared__ double sh[2][kFinalizeThreads / 64];
  const int k = vb.x;
  const int t = threadIdx.x;
  const double2* p = reinterpret_cast<const double2*>(partials) + static_cast<size_t>(k) * nchunk;
  constexpr int A = kFinalizeAcc;
  double a0[A], a1[A];
#pragma unroll
  for (int q = 0; q < A; ++q) a0[q] = a1[q] = 0.0;
  int c = t;
  for (; c + (A - 1) * kFinalizeThreads < nchunk; c += A * kFinalizeThreads) {
    double2 v[A];
#pragma unroll
    for (int q = 0; q < A; ++q) v[q] = p[c + q * kFinalizeThreads];
#pragma unroll
    for (int q = 0; q < A; ++q) {
      a0[q] += v[q].x;
      a1[q] += v[q].y;
    }
  }
#pragma unroll
  for (int q = 0; q < A; ++q) {
    if (c + q * kFinalizeThreads < nchunk) {
      const double2 v = p[c + q * kFinalizeThreads];
      a0[q] += v.x;
      a1[q] += v.y;
    }
  }
  const double s0 = wave_sum_to63((a0[0] + a0[1]) + (a0[2] + a0[3]));
  const double s1 = wave_sum_to63((a1[0] + a1[1]) + (a1[2] + a1[3]));
  if ((t & 63) == 63) {
    sh[0][t >> 6] = s0;
    sh[1][t >> 6] = s1;
  }
  __syncthreads();
  if (t != 0) return;
  // sixteen waves, pairwise
  auto tree = [](const double* x) {
    const double q0 = (x[0] + x[1]) + (x[2] + x[3]), q1 = (x[4] + x[5]) + (x[6] + x[7]);
    const double q2 = (x[8] + x[9]) + (x[10] + x[11]), q3 = (x[12] + x[13]) + (x[14] + x[15]);
    return (q0 + q1) + (q2 + q3);
  };
  static_assert(kFinalizeThreads == 1024 && kFinalizeAcc == 4, "tree sizes");
  const double S0 = tree(sh[0]);
  const double S1 = tree(sh[1]);
  if (mode == 0) {
    out0[k] = static_cast<R>(S0 / count);  // dry_core.cu:485-490
  } else {
    const R inv = R(1) / static_cast<R>(count);  // dry_fast_modes.cu:481-482
    out0[k] = static_cast<R>(S0) * inv;
    out1[k] = static_cast<R>(S1) * inv;
  }
}

template <class R>
__global__ __launch_bounds__(kFinalizeThreads) void k_level_finalize(int nz, int nchunk,
                                                                     const double* partials,
                                                                     int mode, double count,
                                                                     R* out0, R* out1) {
  (void)nz;
  level_finalize_body(nchunk, partials, mode, count, out0, out1, hw_block());
}
template <class R>
__global__ __launch_bounds__(kFinalizeThreads) void k_level_finalize_batch(
    const FinalizeItem<R>* items, BatchMap bm) {
  VBlock vb;
  const FinalizeItem<R>* it = batch_member(items, bm, vb);
  if (!it) return;
  const FinalizeItem<R> a = *it;
  level_finalize_body(a.nchunk, a.partials, a.mode, a.count, a.out0, a.out1, vb);
}

// ===========================================================================
// Pressure over the full storage (halo included) of the halo-filled state:
// compute_dry_pressure_fields (dry_pressure_gradient.cu:177-202, K20).
// ===========================================================================
template <class R>
__global__ void k_pressure(Geo g, const R* rho, const R* rt, R* p) {
  const size_t n = g.cell_total();
  const size_t a = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (a >= n) return;
  const int pi = static_cast<int>(a % g.pxc);
  const int pj = static_cast<int>((a / g.pxc) % g.pyc);
  const int k = static_cast<int>(a / (static_cast<size_t>(g.pxc) * g.pyc));
  Loader<R> L{g};
  const R r = L.cell(rho, pi - g.h, pj - g.h, k);
  const R t = L.cell(rt, pi - g.h, pj - g.h, k);
  p[a] = d_pressure(r, t);
}

template <class R>
__global__ void k_theta(Geo g, const R* rho, const R* rt, R* th, Base<R> base) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= g.nx || j >= g.ny || k >= g.nz) return;
  const size_t c = g.cell(i, j, k);
  if (base.drho)  // terrain form: the cell's base
    th[c] = theta_pert(rho[c], rt[c], base.rho[k] + base.drho[c], base.th[k] + base.dth[c]);
  else
    th[c] = base.rho ? theta_pert(rho[c], rt[c], base.rho[k], base.th[k]) : q_theta(rho[c], rt[c]);
}

template <class R>
__global__ void k_rebuild_rt(Geo g, const R* rho, const R* th, R* rt, Base<R> base) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= g.nx || j >= g.ny) return;
  const size_t c = g.cell(i, j, k);
  if (base.drho) {  // terrain form: the cell's base
    const R rb = base.rho[k] + base.drho[c], tb = base.th[k] + base.dth[c];
    rt[c] = rb * th[c] + rho[c] * (tb + th[c]);
  } else {
    rt[c] = base.rho ? base.rho[k] * th[c] + rho[c] * (base.th[k] + th[c]) : rho[c] * th[c];
  }
}

// ===========================================================================
// Warm rain + fallout, one thread per column marching top-down: K24
// (warm_rain.cu:36-120) on cell k, then the fallout sweep on the same cell
// (K25 nominal, warm_rain.cu:122-154, or K26 metric, 156-191), which reads
// the post-microphysics rain exactly like the reference's qr_old copy.
// ===========================================================================
template <class R, bool TERRAIN>
__device__ __forceinline__ void warm_rain_body(const WarmRainArgs<R>& a, const VBlock& vb) {
  const Geo g = a.g;
  int bx, by;
  xcd_block2d(vb, bx, by);
  const int i = bx * blockDim.x + threadIdx.x;
  const int j = by * blockDim.y + threadIdx.y;
  if (i >= g.nx || j >= g.ny) return;
  const R dt = a.dt;
  R inc = 0;
  const R dz_nom = rmax(a.depth, R(1.0e-3));
  const R frac_nom = rmin(rmax(a.vt * dt / dz_nom, R(0)), R(1));
  // The nominal sweep leaves the column alone when nothing falls out
  // (warm_rain.cu:134-136); the metric sweep has no such exit.
  const bool nominal = a.fallout == 1 && frac_nom > R(0), metric = a.fallout == 2;
  const R rdz_nom = R(1) / dz_nom;
  const MicroCfg<R> mc{dt, a.relax, a.thr, a.krate, a.erate, a.kacc, a.heat};
  // The loads of the level three below are issued before the current
  // level's arithmetic: three levels of loads in flight per column keep
  // enough requests outstanding at one column per thread.
  const size_t plane = static_cast<size_t>(g.pxc) * g.pyc;
  size_t c = g.cell(i, j, g.nz - 1);
  struct Lv {
    R rho, rt, qv, qc, qr;
  };
  auto load = [&](size_t cc) -> Lv { return {a.rho[cc], a.rt[cc], a.qv[cc], a.qc[cc], a.qr[cc]}; };
  Lv l1 = load(c), l2{}, l3{};
  if (g.nz > 1) l2 = load(c - plane);
  if (g.nz > 2) l3 = load(c - 2 * plane);
  const bool pert = a.base.rho != nullptr;
  for (int k = g.nz - 1; k >= 0; --k) {
    const Lv cur = l1;
    l1 = l2;
    l2 = l3;
    if (k >= 3) l3 = load(c - 3 * plane);
    // Perturbation form: full rho and rho*theta for the microphysics; the
    // stored departure of rho*theta then changes by rho * dtheta only.
    const R rtp = cur.rt;
    const R rho = rmax(pert ? cur.rho + a.base.rho[k] : cur.rho, R(1.0e-6));
    const R rt0 = pert ? cur.rt + a.base.rt[k] : cur.rt;
    const R in_qv = cur.qv, in_qc = cur.qc, in_qr = cur.qr;
    MicroCell<R> m = micro_prepare(rho, rt0, in_qv, in_qc, in_qr);
    const R th0 = m.th;
    micro_apply(m, mc);
    const R th = m.th, qv = m.qv, qc = m.qc, qr = m.qr;
    a.qv[c] = rho * qv;
    a.qc[c] = rho * qc;
    R rqr = rho * qr;
    a.rt[c] = pert ? rtp + rho * (th - th0) : rho * th;
    if (nominal) {
      const R old = rmax(rqr, R(0)) * dz_nom;
      const R out = frac_nom * old;
      rqr = qdivc(old - out + inc, dz_nom, rdz_nom);
      inc = out;
    } else if (metric) {
      const R idz = rmax(TERRAIN ? a.idz_c[g.metric(i, j, k)] : a.lev_idz_c[k], R(1.0e-6));
      const R dz = qrcp(idz);
      const R frac = rmin(rmax(a.vt * dt * idz, R(0)), R(1));
      const R old = rmax(rqr, R(0)) * dz;
      const R out = frac * old;
      rqr = (old - out + inc) * idz;
      inc = out;
    }
    a.qr[c] = rqr;
    c -= plane;
  }
  if (nominal || metric) a.precip[static_cast<size_t>(j) * g.nx + i] += inc;
}

template <class R, bool TERRAIN>
__global__ void k_warm_rain(WarmRainArgs<R> a) {
  warm_rain_body<R, TERRAIN>(a, hw_block());
}
template <class R, bool TERRAIN>
__global__ void k_warm_rain_batch(const WarmRainItem<R>* items, BatchMap bm) {
  VBlock vb;
  const WarmRainItem<R>* it = batch_member(items, bm, vb);
  if (!it) return;
  const WarmRainArgs<R> a = it->a;
  warm_rain_body<R, TERRAIN>(a, vb);
}

// ===========================================================================
// Halo fill / interface averaging (1x1 layout) and the multi-rank exchange.
// ===========================================================================
template <class R>
__global__ void k_halo_fill(Geo g, int orient, R* f) {
  // orient: 0 cell, 1 x-face, 2 y-face, 3 z-face
  const int fx = g.nx + (orient == 1), fy = g.ny + (orient == 2), fz = g.nz + (orient == 3);
  const int px = fx + 2 * g.h, py = fy + 2 * g.h;
  const size_t n = static_cast<size_t>(px) * py * fz;
  const size_t a = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (a >= n) return;
  const int ii = static_cast<int>(a % px) - g.h;
  const int jj = static_cast<int>((a / px) % py) - g.h;
  const int k = static_cast<int>(a / (static_cast<size_t>(px) * py));
  if (ii >= 0 && ii < fx && jj >= 0 && jj < fy) return;
  const int si = orient == 1 ? map_face(ii, g.nx, g.per_x) : map_cell(ii, g.nx, g.per_x);
  const int sj = orient == 2 ? map_face(jj, g.ny, g.per_y) : map_cell(jj, g.ny, g.per_y);
  f[a] = (si < 0 || sj < 0) ? R(0)
                            : f[(static_cast<size_t>(k) * py + (sj + g.h)) * px + (si + g.h)];
}

template <class R>
__global__ void k_sync_interfaces(Geo g, int orient, R* f) {
  // synchronize_face_owned_interfaces_impl for a 1x1 periodic layout
  // (halo_exchange.cpp:255-283): face n and face 0 both take their average.
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (orient == 1) {
    if (a >= g.ny * g.nz) return;
    const int j = a % g.ny, k = a / g.ny;
    const size_t e = g.xf(g.nx, j, k), w = g.xf(0, j, k);
    const R avg = R(0.5) * (f[e] + f[w]);
    f[e] = avg;
    f[w] = avg;
  } else {
    if (a >= g.nx * g.nz) return;
    const int i = a % g.nx, k = a / g.nx;
    const size_t nn = g.yf(i, g.ny, k), s = g.yf(i, 0, k);
    const R avg = R(0.5) * (f[nn] + f[s]);
    f[nn] = avg;
    f[s] = avg;
  }
}

// Multi-rank exchange (halo_exchange.cpp:32-116, 148-215, 313-373, 563-620):
// phase 0 fills the x ghosts of owned rows, phase 1 the y ghost rows (which
// include the x ghosts).  d: 13 ints per rank + 4 neighbour ranks.
template <class R>
__global__ void k_exchange_layout(int phase, int orient, int nranks, const int* d, R* const* F,
                                  int maxn) {
  const int rank = blockIdx.y;
  if (rank >= nranks) return;
  const int* D = d + rank * 17;
  const int h = D[8], nxl = D[10] - D[9], nyl = D[12] - D[11];
  const int fx = nxl + (orient == 1), fy = nyl + (orient == 2), fz = D[7] + (orient == 3);
  const int px = fx + 2 * h, py = fy + 2 * h;
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  R* f = F[rank];
  auto at = [&](R* base, int ppx, int ppy, int i, int j, int k) -> R& {
    return base[(static_cast<size_t>(k) * ppy + (j + h)) * ppx + (i + h)];
  };
  if (phase == 0) {
    // x phase: for k < fz, j < fy, hh < h: west ghost (-h+hh) and east ghost (fx+hh)
    const int total = fz * fy * h;
    if (a >= total) return;
    const int hh = a % h, j = (a / h) % fy, k = a / (h * fy);
    const int west = D[13], east = D[14];
    // west ghost <- west neighbour's east send
    if (west >= 0) {
      const int* W = d + west * 17;
      const int wnxl = W[10] - W[9], wnyl = W[12] - W[11];
      const int wfx = wnxl + (orient == 1), wfy = wnyl + (orient == 2);
      const int ei = orient == 1 ? wnxl - h + hh : wfx - h + hh;
      at(f, px, py, -h + hh, j, k) = at(F[west], wfx + 2 * h, wfy + 2 * h, ei, j, k);
    } else {
      at(f, px, py, -h + hh, j, k) = R(0);
    }
    if (east >= 0) {
      const int* E = d + east * 17;
      const int enxl = E[10] - E[9], enyl = E[12] - E[11];
      const int efx = enxl + (orient == 1), efy = enyl + (orient == 2);
      const int wi = orient == 1 ? 1 + hh : hh;
      at(f, px, py, fx + hh, j, k) = at(F[east], efx + 2 * h, efy + 2 * h, wi, j, k);
    } else {
      at(f, px, py, fx + hh, j, k) = R(0);
    }
  } else {
    const int total = fz * h * (fx + 2 * h);
    if (a >= total) return;
    const int i = a % (fx + 2 * h) - h, hh = (a / (fx + 2 * h)) % h, k = a / ((fx + 2 * h) * h);
    const int south = D[15], north = D[16];
    if (south >= 0) {
      const int* S = d + south * 17;
      const int snxl = S[10] - S[9], snyl = S[12] - S[11];
      const int sfx = snxl + (orient == 1), sfy = snyl + (orient == 2);
      const int nj = orient == 2 ? snyl - h + hh : sfy - h + hh;
      at(f, px, py, i, -h + hh, k) = at(F[south], sfx + 2 * h, sfy + 2 * h, i, nj, k);
    } else {
      at(f, px, py, i, -h + hh, k) = R(0);
    }
    if (north >= 0) {
      const int* N = d + north * 17;
      const int nnxl = N[10] - N[9], nnyl = N[12] - N[11];
      const int nfx = nnxl + (orient == 1), nfy = nnyl + (orient == 2);
      const int sj = orient == 2 ? 1 + hh : hh;
      at(f, px, py, i, fy + hh, k) = at(F[north], nfx + 2 * h, nfy + 2 * h, i, sj, k);
    } else {
      at(f, px, py, i, fy + hh, k) = R(0);
    }
  }
  (void)maxn;
}

// ===========================================================================
// Boundary strips / relaxation on the outermost owned strips of one field
// (boundary_conditions.cpp:9-161; relaxation is the extension documented in
// the header).  orient as in k_halo_fill.  Threads cover the frame of
// cells within wx (wy) of the open x (y) sides when it is narrower than the
// field (wx = 0: the whole field): x strips of every row, then y strips of
// the columns between them; each cell's test and update are the same
// either way.
// ===========================================================================
template <class R>
__global__ void k_boundary(Geo g, int orient, R* f, const R* ref, int mode, R width, int wx,
                           int wy) {
  const int fx = g.nx + (orient == 1), fy = g.ny + (orient == 2), fz = g.nz + (orient == 3);
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  int i, j, k;
  if (wx > 0 || wy > 0) {
    const int wx0 = wx > 0 ? wx : 0;
    const int nxs = wx > 0 ? 2 * wx * fy : 0;
    const int nys = wy > 0 ? 2 * wy * (fx - 2 * wx0) : 0;
    const int per = nxs + nys;
    if (a >= per * fz) return;
    k = a / per;
    int r = a % per;
    if (r < nxs) {
      const int ii = r % (2 * wx);
      j = r / (2 * wx);
      i = ii < wx ? ii : fx - 2 * wx + ii;
    } else {
      r -= nxs;
      const int jj = r % (2 * wy);
      i = wx0 + r / (2 * wy);
      j = jj < wy ? jj : fy - 2 * wy + jj;
    }
  } else {
    if (a >= fx * fy * fz) return;
    i = a % fx;
    j = (a / fx) % fy;
    k = a / (fx * fy);
  }
  const bool w = !g.per_x, e = !g.per_x, s = !g.per_y, n = !g.per_y;
  const int px = fx + 2 * g.h, py = fy + 2 * g.h;
  const size_t idx = (static_cast<size_t>(k) * py + (j + g.h)) * px + (i + g.h);
  if (mode == 1) {
    if ((w && i == 0) || (e && i == fx - 1) || (s && j == 0) || (n && j == fy - 1)) f[idx] = ref[idx];
    return;
  }
  double dmin = 1e30;
  if (w) dmin = fmin(dmin, static_cast<double>(i));
  if (e) dmin = fmin(dmin, static_cast<double>(fx - 1 - i));
  if (s) dmin = fmin(dmin, static_cast<double>(j));
  if (n) dmin = fmin(dmin, static_cast<double>(fy - 1 - j));
  if (dmin >= static_cast<double>(width)) return;
  const double a0 = 1.0 - dmin / static_cast<double>(width);
  const R wgt = static_cast<R>(a0 * a0);
  f[idx] = f[idx] + wgt * (ref[idx] - f[idx]);
}

// ===========================================================================
// Budget reductions: compensated (Neumaier) FP64 partial sums, min and max
// of a field over owned elements, optionally weighted by dz = 1/inv_dz_cell
// (column paths, runtime_summary.cpp:70-80) or divided by rho (mixing
// ratio, runtime_summary.cpp:82-99).  One partial record per block:
// {sum, comp, min, max}.
// ===========================================================================
template <class R>
__global__ __launch_bounds__(256) void k_reduce(ReduceArgs<R> a) {
  __shared__ double sh[256];
  const int t = threadIdx.x;
  const size_t n = static_cast<size_t>(a.fx) * a.fy * a.fz;
  double sum = 0.0, comp = 0.0, mn = 1e300, mx = -1e300, bad = 0.0;
  for (size_t e = static_cast<size_t>(blockIdx.x) * 256 + t; e < n;
       e += static_cast<size_t>(gridDim.x) * 256) {
    const int i = static_cast<int>(e % a.fx);
    const int j = static_cast<int>((e / a.fx) % a.fy);
    const int k = static_cast<int>(e / (static_cast<size_t>(a.fx) * a.fy));
    const size_t idx = (static_cast<size_t>(k) * (a.fy + 2 * a.h) + (j + a.h)) * (a.fx + 2 * a.h) + (i + a.h);
    double v = static_cast<double>(a.f[idx]);
    // perturbation form: extremes of the full value, sums of the departure
    double fb = 0.0;
    if (a.fbase) {
      if (a.cb.drho && a.fbase == a.cb.rt)
        fb = (static_cast<double>(a.cb.rho[k]) + static_cast<double>(a.cb.drho[idx])) *
             (static_cast<double>(a.cb.th[k]) + static_cast<double>(a.cb.dth[idx]));
      else
        fb = static_cast<double>(a.fbase[k]) + (a.cb.drho ? static_cast<double>(a.cb.drho[idx]) : 0.0);
    }
    const double full = a.fbase ? v + fb : v;
    double ext = full;
    if (a.kind == 1) {  // column path
      const R idz = a.idz3 ? a.idz3[(static_cast<size_t>(k) * a.fy + j) * a.fx + i] : a.idz_lev[k];
      v = v * (1.0 / static_cast<double>(idz));
      ext = v;
    } else if (a.kind == 2) {  // theta_m / mixing ratio: f / max(rho, 1e-12)
      const double r = static_cast<double>(a.rho[idx]) +
                       (a.rbase ? static_cast<double>(a.rbase[k]) +
                                      (a.cb.drho ? static_cast<double>(a.cb.drho[idx]) : 0.0)
                                : 0.0);
      v = full / fmax(r, 1.0e-12);
      ext = v;
    } else if (a.kind == 3) {  // count of positive values; extremes of the value
      v = full > 0.0 ? 1.0 : 0.0;
    }
    // A NaN or infinity is counted, and kept out of the sums and extremes,
    // which then describe the finite values; the count tells the caller
    // whether those describe the whole field.
    if (!isfinite(full) || !isfinite(v)) {
      bad += 1.0;
      continue;
    }
    const double s2 = sum + v;
    comp += fabs(sum) >= fabs(v) ? (sum - s2) + v : (v - s2) + sum;
    sum = s2;
    mn = fmin(mn, ext);
    mx = fmax(mx, ext);
  }
  // Block combine in a fixed tree order; every node is a TwoSum whose
  // rounding error joins the compensation, so sum + comp stays exact to
  // working precision.
  double* out = a.partials + static_cast<size_t>(blockIdx.x) * kReduceParts;
  __shared__ double se[256];
  sh[t] = sum;
  se[t] = comp;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
      const double x = sh[t], y = sh[t + s];
      const double s2 = x + y;
      const double bb = s2 - x;
      const double err = (x - (s2 - bb)) + (y - bb);
      sh[t] = s2;
      se[t] = se[t] + se[t + s] + err;
    }
    __syncthreads();
  }
  const double bsum = sh[0];
  const double bcomp = se[0];
  __syncthreads();
  sh[t] = mn;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) sh[t] = fmin(sh[t], sh[t + s]);
    __syncthreads();
  }
  const double bmn = sh[0];
  __syncthreads();
  sh[t] = mx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) sh[t] = fmax(sh[t], sh[t + s]);
    __syncthreads();
  }
  const double bmx = sh[0];
  __syncthreads();
  sh[t] = bad;  // counts below 2^53 add exactly in any order
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) sh[t] += sh[t + s];
    __syncthreads();
  }
  if (t == 0) {
    out[0] = bsum;
    out[1] = bcomp;
    out[2] = bmn;
    out[3] = bmx;
    out[4] = sh[0];
  }
}

// ===========================================================================
// Metric construction (grid_metrics.hpp:184-236): one thread per column,
// FP64 arithmetic without contraction so the stored values match a host
// build of the same formulas.
// ===========================================================================
template <class R>
__global__ void k_metrics(Geo g, const double* zif, const double* zc, const double* wif,
                          const double* wc, const float* terrain, R* zif_m, R* zc_m, R* idz_c,
                          R* idz_f, R* slope_x, R* slope_y, R* terr_out) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= g.nx || j >= g.ny) return;
  auto tr = [&](int ii, int jj) -> float { return terrain ? terrain[static_cast<size_t>(jj) * g.nx + ii] : 0.0f; };
  auto wi = [&](int ii) { return g.per_x ? map_cell(ii, g.nx, 1) : min(max(ii, 0), g.nx - 1); };
  auto wj = [&](int jj) { return g.per_y ? map_cell(jj, g.ny, 1) : min(max(jj, 0), g.ny - 1); };
  const float th = tr(i, j);
  const size_t c2 = static_cast<size_t>(j) * g.nx + i;
  terr_out[c2] = static_cast<R>(th);
  const double ddx = (static_cast<double>(tr(wi(i + 1), j)) - static_cast<double>(tr(wi(i - 1), j))) /
                     (g.per_x || (i > 0 && i + 1 < g.nx) ? 2.0 * g.dx : g.dx);
  const double ddy = (static_cast<double>(tr(i, wj(j + 1))) - static_cast<double>(tr(i, wj(j - 1)))) /
                     (g.per_y || (j > 0 && j + 1 < g.ny) ? 2.0 * g.dy : g.dy);
  slope_x[c2] = static_cast<R>(ddx);
  slope_y[c2] = static_cast<R>(ddy);
  const double thd = static_cast<double>(th);
  for (int k = 0; k <= g.nz; ++k) {
    const double prod = wif[k] * thd;
    zif_m[g.metric(i, j, k)] = static_cast<R>(zif[k] + prod);
    if (k == 0 || k == g.nz) {
      idz_f[g.metric(i, j, k)] = R(0);
    } else {
      const double hi = zc[k] + wc[k] * thd;
      const double lo = zc[k - 1] + wc[k - 1] * thd;
      const double dzf = hi - lo;
      idz_f[g.metric(i, j, k)] = static_cast<R>(1.0 / fmax(dzf, 1.0e-6));
    }
  }
  for (int k = 0; k < g.nz; ++k) {
    const double prod = wc[k] * thd;
    zc_m[g.metric(i, j, k)] = static_cast<R>(zc[k] + prod);
    const double dzc = static_cast<double>(zif_m[g.metric(i, j, k + 1)]) -
                       static_cast<double>(zif_m[g.metric(i, j, k)]);
    idz_c[g.metric(i, j, k)] = static_cast<R>(1.0 / fmax(dzc, 1.0e-6));
  }
}

// The FP32 arithmetic helpers, elementwise (gwm_fp32_math).
__global__ void k_fp32_math(int fn, const float* x, float* y, size_t n) {
  const size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float v = x[t];
  float r;
  switch (fn) {
    case 0: r = qpow(v, 1.0f / (1.0f - Phys<float>::kappa)); break;
    case 1: r = qpow(v, Phys<float>::kappa); break;
    case 2: r = qexp(v); break;
    case 3: r = qlog1p(v); break;
    case 4: r = qexpm1(v); break;
    case 6: r = p_pert(v, 1.0f, 1.0f); break;  // (1 + x)^gamma - 1
    default: r = qrcp(v); break;
  }
  y[t] = r;
}

template <class R>
__global__ void k_copy(const R* src, R* dst, size_t n) {
  const size_t a = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (a < n) dst[a] = src[a];
}
template <class R>
__global__ void k_fill(R* dst, size_t n, R v) {
  const size_t a = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (a < n) dst[a] = v;
}

// ===========================================================================
// Launchers
// ===========================================================================
namespace {
inline dim3 cell_grid(const Geo& g, dim3 b) {
  return dim3((g.nx + b.x - 1) / b.x, (g.ny + b.y - 1) / b.y, g.nz);
}
inline unsigned blocks_for(size_t n, unsigned b) { return static_cast<unsigned>((n + b - 1) / b); }
}  // namespace

template <class R>
void launch_level_partials(hipStream_t st, const Geo& g, const R* rho, const R* aux, int mode,
                           double* partials, int* nchunk_out, Base<R> base) {
  const int nchunk = level_partials_chunks(g.ny);
  *nchunk_out = nchunk;
  if (recorder()) {
    record_launch(reinterpret_cast<const void*>(&k_level_partials_batch<R>), dim3(256), 0,
                  PartialsItem<R>{g, rho, aux, mode, partials, base, 0, 0}, nchunk, g.nz);
    return;
  }
  hipLaunchKernelGGL(k_level_partials<R>, dim3(nchunk, g.nz), dim3(256), 0, st, g, rho, aux, mode,
                     partials, base);
}
int level_partials_chunks(int ny) { return (ny + kPartialRows - 1) / kPartialRows; }
template <class R>
void launch_level_finalize(hipStream_t st, int nz, int nchunk, const double* partials, int mode,
                           double count, R* o0, R* o1) {
  if (recorder()) {
    record_launch(reinterpret_cast<const void*>(&k_level_finalize_batch<R>),
                  dim3(kFinalizeThreads), 0,
                  FinalizeItem<R>{nz, nchunk, partials, mode, count, o0, o1, 0, 0}, nz, 1);
    return;
  }
  hipLaunchKernelGGL(k_level_finalize<R>, dim3(nz), dim3(kFinalizeThreads), 0, st, nz, nchunk, partials,
                     mode, count, o0, o1);
}
template <class R>
void launch_pressure(hipStream_t st, const Geo& g, const R* rho, const R* rt, R* p) {
  if (refuse_record("pressure")) return;
  hipLaunchKernelGGL(k_pressure<R>, dim3(blocks_for(g.cell_total(), 256)), dim3(256), 0, st, g,
                     rho, rt, p);
}
template <class R>
void launch_theta(hipStream_t st, const Geo& g, const R* rho, const R* rt, R* th, Base<R> base) {
  if (refuse_record("theta")) return;
  const dim3 b(64, 4, 1);
  hipLaunchKernelGGL(k_theta<R>, cell_grid(g, b), b, 0, st, g, rho, rt, th, base);
}
template <class R>
void launch_rebuild_rt(hipStream_t st, const Geo& g, const R* rho, const R* th, R* rt,
                       Base<R> base) {
  if (refuse_record("rebuild_rt")) return;
  const dim3 b(64, 4, 1);
  hipLaunchKernelGGL(k_rebuild_rt<R>, cell_grid(g, b), b, 0, st, g, rho, th, rt, base);
}
template <class R>
void launch_warm_rain(hipStream_t st, const WarmRainArgs<R>& a, bool terrain) {
  const dim3 b(64, 4, 1);
  const dim3 gr((a.g.nx + 63) / 64, (a.g.ny + 3) / 4);
  if (recorder()) {
    const void* fn = terrain ? reinterpret_cast<const void*>(&k_warm_rain_batch<R, true>)
                             : reinterpret_cast<const void*>(&k_warm_rain_batch<R, false>);
    record_launch(fn, b, 0, WarmRainItem<R>{a, 0, 0}, static_cast<int>(gr.x),
                  static_cast<int>(gr.y));
    return;
  }
  if (terrain)
    hipLaunchKernelGGL((k_warm_rain<R, true>), gr, b, 0, st, a);
  else
    hipLaunchKernelGGL((k_warm_rain<R, false>), gr, b, 0, st, a);
}
template <class R>
void launch_halo_fill(hipStream_t st, const Geo& g, int orient, R* f) {
  if (refuse_record("halo_fill")) return;
  const int fx = g.nx + (orient == 1), fy = g.ny + (orient == 2), fz = g.nz + (orient == 3);
  const size_t n = static_cast<size_t>(fx + 2 * g.h) * (fy + 2 * g.h) * fz;
  hipLaunchKernelGGL(k_halo_fill<R>, dim3(blocks_for(n, 256)), dim3(256), 0, st, g, orient, f);
}
template <class R>
void launch_sync_interfaces(hipStream_t st, const Geo& g, int orient, R* f) {
  if (refuse_record("sync_interfaces")) return;
  const int n = orient == 1 ? g.ny * g.nz : g.nx * g.nz;
  hipLaunchKernelGGL(k_sync_interfaces<R>, dim3(blocks_for(n, 256)), dim3(256), 0, st, g, orient, f);
}
template <class R>
void launch_exchange_layout(hipStream_t st, int phase, int orient, int nranks, const int* d,
                            R* const* F, int max_elems) {
  hipLaunchKernelGGL(k_exchange_layout<R>, dim3(blocks_for(max_elems, 256), nranks), dim3(256), 0,
                     st, phase, orient, nranks, d, F, max_elems);
}
template <class R>
void launch_boundary(hipStream_t st, const Geo& g, int orient, R* f, const R* ref, int mode,
                     double width) {
  if (refuse_record("boundary")) return;
  const int fx = g.nx + (orient == 1), fy = g.ny + (orient == 2), fz = g.nz + (orient == 3);
  // frame widths: the cells a side can reach (the outermost one for strips,
  // distance < width for relaxation); the whole field when a frame would
  // cover it
  const int reach = mode == 1 ? 1 : (width > 0.0 ? static_cast<int>(std::ceil(width)) : 0);
  int wx = g.per_x ? 0 : reach, wy = g.per_y ? 0 : reach;
  size_t n = static_cast<size_t>(fx) * fy * fz;
  if ((wx > 0 || wy > 0) && 2 * wx < fx && 2 * wy < fy) {
    const size_t nxs = wx > 0 ? 2 * static_cast<size_t>(wx) * fy : 0;
    const size_t nys = wy > 0 ? 2 * static_cast<size_t>(wy) * (fx - 2 * wx) : 0;
    n = (nxs + nys) * fz;
  } else {
    wx = wy = 0;
  }
  if (n == 0) return;
  hipLaunchKernelGGL(k_boundary<R>, dim3(blocks_for(n, 256)), dim3(256), 0, st, g, orient, f, ref,
                     mode, static_cast<R>(width), wx, wy);
}
template <class R>
void launch_reduce(hipStream_t st, const ReduceArgs<R>& a, int nblocks) {
  if (refuse_record("reduce")) return;
  hipLaunchKernelGGL(k_reduce<R>, dim3(nblocks), dim3(256), 0, st, a);
}

Recorder*& recorder() {
  static thread_local Recorder* r = nullptr;
  return r;
}

// Block partials of the allocation-free budget entry point
// (gwm_budget_sums_noalloc): a file-scope array, one copy per GPU, whose
// device address is resolved once per device by noalloc_scratch_setup.  The
// name is a host-side stand-in; host code uses only the resolved address.
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

hipError_t noalloc_scratch_setup(int device, double** out) {
  static std::mutex mu;
  static double* addr[kNoAllocMaxDevices] = {};
  std::lock_guard<std::mutex> lock(mu);
  *out = nullptr;
  if (device < 0 || device >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
  if (addr[device] == nullptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[device] = static_cast<double*>(p);
  }
  *out = addr[device];
  return hipSuccess;
}
template <class R>
void launch_metrics(hipStream_t st, const Geo& g, const double* zif, const double* zc,
                    const double* wif, const double* wc, const float* terrain, R* zif_m, R* zc_m,
                    R* idz_c, R* idz_f, R* sx, R* sy, R* terr) {
  const dim3 b(64, 4, 1);
  hipLaunchKernelGGL(k_metrics<R>, dim3((g.nx + 63) / 64, (g.ny + 3) / 4), b, 0, st, g, zif, zc,
                     wif, wc, terrain, zif_m, zc_m, idz_c, idz_f, sx, sy, terr);
}
void launch_fp32_math(hipStream_t st, int fn, const float* x, float* y, size_t n) {
  hipLaunchKernelGGL(k_fp32_math, dim3(blocks_for(n, 256)), dim3(256), 0, st, fn, x, y, n);
}
template <class R>
void launch_copy(hipStream_t st, const R* src, R* dst, size_t n) {
  if (refuse_record("copy")) return;
  hipLaunchKernelGGL(k_copy<R>, dim3(blocks_for(n, 256)), dim3(256), 0, st, src, dst, n);
}
template <class R>
void launch_fill(hipStream_t st, R* dst, size_t n, R v) {
  if (refuse_record("fill")) return;
  hipLaunchKernelGGL(k_fill<R>, dim3(blocks_for(n, 256)), dim3(256), 0, st, dst, n, v);
}

#define GWM_INSTANTIATE(R)                                                                          \
  template void launch_level_partials<R>(hipStream_t, const Geo&, const R*, const R*, int,        \
                                         double*, int*, Base<R>);                                 \
  template void launch_level_finalize<R>(hipStream_t, int, int, const double*, int, double, R*,   \
                                         R*);                                                     \
  template void launch_pressure<R>(hipStream_t, const Geo&, const R*, const R*, R*);             \
  template void launch_theta<R>(hipStream_t, const Geo&, const R*, const R*, R*, Base<R>);       \
  template void launch_rebuild_rt<R>(hipStream_t, const Geo&, const R*, const R*, R*, Base<R>);  \
  template void launch_warm_rain<R>(hipStream_t, const WarmRainArgs<R>&, bool);                  \
  template void launch_halo_fill<R>(hipStream_t, const Geo&, int, R*);                           \
  template void launch_sync_interfaces<R>(hipStream_t, const Geo&, int, R*);                     \
  template void launch_exchange_layout<R>(hipStream_t, int, int, int, const int*, R* const*, int); \
  template void launch_boundary<R>(hipStream_t, const Geo&, int, R*, const R*, int, double);     \
  template void launch_reduce<R>(hipStream_t, const ReduceArgs<R>&, int);                        \
  template void launch_metrics<R>(hipStream_t, const Geo&, const double*, const double*,         \
                                  const double*, const double*, const float*, R*, R*, R*, R*, R*, \
                                  R*, R*);                                                        \
  template void launch_copy<R>(hipStream_t, const R*, R*, size_t);                               \
  template void launch_fill<R>(hipStream_t, R*, size_t, R);

GWM_INSTANTIATE(float)
GWM_INSTANTIATE(double)

}  // namespace gwm
