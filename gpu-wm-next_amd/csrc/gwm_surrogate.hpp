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
void launch_warm_rain_sur(hipStream_t st, const Geo& g, const float* rho, float* rt, float* qv,
                          float* qc, float* qr, const MicroCfg<float>& mc, const float* w,
                          const SurNorm& nm, unsigned long long* counts, Base<float> base);
// Clears n (<= 64) device counters with a kernel (no memset node in a
// captured graph).
void launch_zero_counters(hipStream_t st, unsigned long long* p, int n);
void launch_fallout(hipStream_t st, const Geo& g, float* qr, float* precip, const float* idz_c,
                    const float* lev_idz_c, float dt, float vt, float depth, int fallout,
                    bool terrain);

// Items of the batched forms (gwm_kernels.hpp, Recorder).
struct SurItem {
  Geo g;
  const float* rho;
  float *rt, *qv, *qc, *qr;
  MicroCfg<float> mc;
  const float* w;
  SurNorm nm;
  unsigned long long* counts;
  Base<float> base;
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
  int gx, gy;
};

}  // namespace gwm
