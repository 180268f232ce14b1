// gwm_surrogate.hpp -- layout and launchers of the warm-rain surrogate MLP
// (gwm_surrogate.hip).  No counterpart in the reference.
#pragma once

#include "gwm_kernels.hpp"
#include "gwm_micro.hpp"

namespace gwm {

constexpr int kSurIn = 7;   // T, p, rho_d, qv, qc, qr, dt
constexpr int kSurOut = 4;  // dqv, dqc, dqr, dtheta
// Flat FP32 parameter layout, W[in][out].
constexpr int kSurW1 = 0;
constexpr int kSurB1 = kSurW1 + kSurIn * 64;
constexpr int kSurW2 = kSurB1 + 64;
constexpr int kSurB2 = kSurW2 + 64 * 64;
constexpr int kSurW3 = kSurB2 + 64;
constexpr int kSurB3 = kSurW3 + 64 * 64;
constexpr int kSurW4 = kSurB3 + 64;
constexpr int kSurB4 = kSurW4 + 64 * kSurOut;
constexpr int kSurParams = kSurB4 + kSurOut;  // 9092

// Feature / target normalisation and the fallback box (raw feature units).
struct SurNorm {
  float xm[kSurIn], xs[kSurIn];
  float ym[kSurOut], ys[kSurOut];
  float lo[kSurIn], hi[kSurIn];
};

// Device buffers of the FP32 training path for a minibatch of B rows.
struct SurTrain {
  float* params = nullptr;  // kSurParams (master weights)
  float* grad = nullptr;
  float* m = nullptr;
  float* v = nullptr;
  float* xn = nullptr;  // [B][8]
  float* z[4] = {};     // pre-activations [B][64] (z[3]: [B][4] outputs)
  float* a[3] = {};     // activations [B][64]
  float* dz[4] = {};    // gradients w.r.t. z
  float* wpart = nullptr;  // [16][65][64] partial weight gradients
  long step = 0;
  float lr = 1.0e-3f, beta1 = 0.9f, beta2 = 0.999f, eps = 1.0e-8f;
};

void launch_sur_sample(hipStream_t st, const Geo& g, const float* rho, const float* rt,
                       const float* qv, const float* qc, const float* qr,
                       const MicroCfg<float>& mc, unsigned long long seed, long n, float* x,
                       float* y, long long* cell, Base<float> base);
// Forward + loss + backward on B rows (x raw [B][7], y raw [B][4]); with
// `update`, one Adam step.  Per-block loss partials (sum of squared
// normalised errors) go to loss_part, their count to *nloss.
void launch_sur_train_step(hipStream_t st, SurTrain& T, const float* x, const float* y, int B,
                           const SurNorm& nm, bool update, double* loss_part, int* nloss);
void launch_sur_infer(hipStream_t st, const float* w, const SurNorm& nm, const float* x, long n,
                      float* y);
// Arguments of the surrogate warm-rain call.  The call runs in two passes:
// the classify pass walks every cell (a wave per 64 columns, top down),
// applies the plain scheme to the cells outside the training box and writes
// quiet cells back, and lists each wave's network cells (cell offsets) in
// its own queue segment; the network pass then evaluates only the listed
// cells, packed 256 to a tile.  A segment is the 64 columns of one wave:
// nseg = ceil(nx / 64) * ny segments of 64 * nz entries.
struct SurArgs {
  Geo g;
  const float* rho;
  float *rt, *qv, *qc, *qr;
  MicroCfg<float> mc;
  const float* w;
  SurNorm nm;
  unsigned long long* counts;  // [0] += fallback cells, [1] += network cells
  Base<float> base;
  unsigned* queue;  // nseg * 64 * nz cell offsets
  unsigned* segs;   // [2][nseg]: network cells, fallback cells per segment
  // Rain fallout of the adjusted rain (0 none, 1 nominal depth, 2 metric;
  // time step fdt).  The classify pass carries it down each column until
  // the column's first network cell; the fallout pass then finishes the
  // columns that have one, from that level with the rain entering it.
  float* precip;
  const float* idz_c;
  const float* lev_idz_c;
  float fdt, vt, depth;
  int fallout;
  int* col_stop;   // [nx * ny] first network level of the column, -1 none
  float* col_inc;  // [nx * ny] rain entering that level from above
};
inline long sur_segments(const Geo& g) { return static_cast<long>((g.nx + 63) / 64) * g.ny; }
// Device words the queue, segment counts and column records of a grid need.
inline size_t sur_queue_words(const Geo& g) {
  return static_cast<size_t>(sur_segments(g)) * (64 * static_cast<size_t>(g.nz) + 2) +
         2 * static_cast<size_t>(g.nx) * g.ny;
}
// The call's passes: classify, network, and (fallout != 0) the fallout of
// the columns that hold network cells.
void launch_warm_rain_sur(hipStream_t st, const SurArgs& a, bool terrain);
// Clears n (<= 64) device counters with a kernel (no memset node in a
// captured graph).
void launch_zero_counters(hipStream_t st, unsigned long long* p, int n);
// Rain fallout of whole columns (col_stop null), or from each column's
// col_stop level with col_inc entering it (columns with col_stop < 0 are
// done).
void launch_fallout(hipStream_t st, const Geo& g, float* qr, float* precip, const float* idz_c,
                    const float* lev_idz_c, float dt, float vt, float depth, int fallout,
                    bool terrain, const int* col_stop = nullptr, const float* col_inc = nullptr);

// Items of the batched forms (gwm_kernels.hpp, Recorder).
struct SurItem {
  SurArgs a;
  int gx, gy;
};
struct ZeroItem {
  unsigned long long* p;
  int n;
  int gx, gy;
};
struct FalloutItem {
  Geo g;
  float* qr;
  float* precip;
  const float* idz_c;
  const float* lev_idz_c;
  float dt, vt, depth;
  int fallout;
  const int* col_stop;
  const float* col_inc;
  int gx, gy;
};

}  // namespace gwm
