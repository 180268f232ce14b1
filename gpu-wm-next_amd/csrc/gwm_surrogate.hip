// gwm_surrogate.hip -- the warm-rain surrogate: a per-cell MLP that replaces
// the source terms of K24 (src/physics/warm_rain.cu:36-120).  No counterpart
// in the reference; the plain scheme (gwm_micro.hpp) generates its training
// data, is its fallback outside the training box, and judges its error.
//
// Network: 7 features (T, p, rho_d, qv, qc, qr, dt; normalised) -> 64 -> 64
// -> 64 -> 4 outputs (dqv, dqc, dqr, dtheta; normalised), GELU (tanh form)
// after each hidden layer.  Parameters, FP32 master copy, flat:
//   W1[7][64] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64][4] b4[4]
// (9092 values, W[in][out]).
//
// Training runs in FP32 (forward, backward and Adam kernels below), so its
// gradients can be checked against an FP64 host restatement.  Inference runs
// on the matrix cores: BF16 weights and activations, FP32 accumulation.
#include "gwm_kernels.hpp"
#include "gwm_micro.hpp"
#include "gwm_surrogate.hpp"

namespace gwm {

namespace {

constexpr float kGeluA = 0.7978845608028654f;  // sqrt(2 / pi)
constexpr float kGeluB = 0.044715f;

__device__ inline float gelu_f(float x) {
  const float u = kGeluA * (x + kGeluB * x * x * x);
  return 0.5f * x * (1.0f + tanhf(u));
}
__device__ inline float gelu_grad_f(float x) {
  const float u = kGeluA * (x + kGeluB * x * x * x);
  const float t = tanhf(u);
  return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * kGeluA * (1.0f + 3.0f * kGeluB * x * x);
}
// Same function as gelu_f written as x * sigmoid(2u) = x / (1 + 2^(x (c0 +
// c1 x^2))), c0 = -2 A log2(e), c1 = c0 B, for inference: seven instructions
// (hardware exp2 and reciprocal), within a few FP32 ulp of gelu_f, far below
// the BF16 rounding that follows.
__device__ inline float gelu_fast(float x) {
  constexpr float c0 = -2.0f * kGeluA * 1.44269504088896341f;
  constexpr float c1 = c0 * kGeluB;
  const float a = x * __builtin_fmaf(c1, x * x, c0);
  return x * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(a));
}

// FP32 -> BF16, round to nearest even, NaN kept quiet (v_cvt_pk_bf16_f32).
__device__ inline __bf16 bf16_of(float f) { return static_cast<__bf16>(f); }
__device__ inline unsigned short to_bf16(float f) {
  return __builtin_bit_cast(unsigned short, bf16_of(f));
}

__device__ inline unsigned long long mix64(unsigned long long z) {  // splitmix64 finaliser
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

}  // namespace

// ===========================================================================
// Training-data sampler: n cells drawn by a counter-based hash of (seed, s)
// from the owned cells; features are the raw (T, p, rho_d, qv, qc, qr, dt)
// of the cell before the adjustment, targets the plain scheme's change
// (dqv, dqc, dqr, dtheta) over one call.
// ===========================================================================
__global__ void k_sur_sample(Geo g, const float* rho, const float* rt, const float* qv,
                             const float* qc, const float* qr, MicroCfg<float> mc,
                             unsigned long long seed, long n, float* x, float* y,
                             long long* cell_out, Base<float> base) {
  const long s = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const unsigned long long cells = static_cast<unsigned long long>(g.nx) * g.ny * g.nz;
  const unsigned long long r = mix64(seed * 0x100000001b3ull + static_cast<unsigned long long>(s));
  const unsigned long long lin = r % cells;
  const int i = static_cast<int>(lin % g.nx);
  const int j = static_cast<int>((lin / g.nx) % g.ny);
  const int k = static_cast<int>(lin / (static_cast<unsigned long long>(g.nx) * g.ny));
  const size_t c = g.cell(i, j, k);
  if (cell_out) cell_out[s] = static_cast<long long>(lin);
  const bool pert = base.rho != nullptr;
  MicroCell<float> m = micro_prepare(pert ? rho[c] + base.rho[k] : rho[c],
                                     pert ? rt[c] + base.rt[k] : rt[c], qv[c], qc[c], qr[c]);
  float* xs = x + s * kSurIn;
  xs[0] = m.T;
  xs[1] = m.p;
  xs[2] = m.rho;
  xs[3] = m.qv;
  xs[4] = m.qc;
  xs[5] = m.qr;
  xs[6] = mc.dt;
  const float th0 = m.th, qv0 = m.qv, qc0 = m.qc, qr0 = m.qr;
  micro_apply(m, mc);
  float* ys = y + s * kSurOut;
  ys[0] = m.qv - qv0;
  ys[1] = m.qc - qc0;
  ys[2] = m.qr - qr0;
  ys[3] = m.th - th0;
}

// ===========================================================================
// FP32 training kernels.  Row-major activations [B][width].
// ===========================================================================

// x_n[b][0..7] = (x[b][f] - mean) / std, feature 7 = 0.
__global__ void k_sur_norm_in(const float* x, long n, SurNorm nm, float* xn) {
  const long t = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= n * 8) return;
  const long b = t >> 3;
  const int f = static_cast<int>(t & 7);
  xn[t] = f < kSurIn ? (x[b * kSurIn + f] - nm.xm[f]) / nm.xs[f] : 0.0f;
}

// z = in W + b; a = gelu(z) (act) or z.  Thread per (row, out).
__global__ void k_dense_fwd(const float* in, int ldi, int K, const float* W, const float* bias,
                            int N, int B, int act, float* z, float* a) {
  const long t = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= static_cast<long>(B) * N) return;
  const long b = t / N;
  const int n = static_cast<int>(t % N);
  float s = bias[n];
  const float* row = in + b * ldi;
  for (int k = 0; k < K; ++k) s += row[k] * W[k * N + n];
  z[t] = s;
  if (a) a[t] = act ? gelu_f(s) : s;
}

// dy = 2 (y_pred - y_target_n) / (B * 4); per-block partial loss sums.
__global__ __launch_bounds__(256) void k_mse_grad(const float* yp, const float* yt, int B,
                                                  SurNorm nm, float* dy, double* part) {
  __shared__ double sh[256];
  const long t = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  double l = 0.0;
  if (t < static_cast<long>(B) * kSurOut) {
    const int o = static_cast<int>(t % kSurOut);
    const float target = (yt[t] - nm.ym[o]) / nm.ys[o];
    const float d = yp[t] - target;
    dy[t] = 2.0f * d / static_cast<float>(B * kSurOut);
    l = static_cast<double>(d) * d;
  }
  sh[threadIdx.x] = l;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// dz_in[b][k] = (sum_n dz[b][n] W[k][n]) * gelu'(z_in[b][k]).
__global__ void k_dense_bwd_data(const float* dz, int N, const float* W, int K, const float* zin,
                                 int B, float* dzin) {
  const long t = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= static_cast<long>(B) * K) return;
  const long b = t / K;
  const int k = static_cast<int>(t % K);
  float s = 0.0f;
  const float* row = dz + b * N;
  for (int n = 0; n < N; ++n) s += row[n] * W[k * N + n];
  dzin[t] = s * gelu_grad_f(zin[t]);
}

// Partial weight gradients over a slice of the batch: out[slice][k][n]
// (k == K is the bias row).  Fixed summation order.
__global__ void k_dense_bwd_weight(const float* in, int ldi, int K, const float* dz, int N, int B,
                                   int slices, float* out) {
  const long t = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const long per = static_cast<long>(K + 1) * N;
  if (t >= per * slices) return;
  const int sl = static_cast<int>(t / per);
  const int kn = static_cast<int>(t % per);
  const int k = kn / N, n = kn % N;
  const int b0 = static_cast<int>((static_cast<long>(B) * sl) / slices);
  const int b1 = static_cast<int>((static_cast<long>(B) * (sl + 1)) / slices);
  float s = 0.0f;
  for (int b = b0; b < b1; ++b) s += (k < K ? in[static_cast<long>(b) * ldi + k] : 1.0f) * dz[static_cast<long>(b) * N + n];
  out[t] = s;
}

// g[W offset + k*N + n] / g[b offset + n] = sum over slices.
__global__ void k_dense_grad_reduce(const float* part, int K, int N, int slices, float* gW,
                                    float* gb) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int per = (K + 1) * N;
  if (t >= per) return;
  float s = 0.0f;
  for (int sl = 0; sl < slices; ++sl) s += part[static_cast<long>(sl) * per + t];
  const int k = t / N, n = t % N;
  if (k < K) gW[k * N + n] = s;
  else gb[n] = s;
}

__global__ void k_adam(float* p, const float* g, float* m, float* v, int n, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float gt = g[t];
  const float mt = b1 * m[t] + (1.0f - b1) * gt;
  const float vt = b2 * v[t] + (1.0f - b2) * gt * gt;
  m[t] = mt;
  v[t] = vt;
  p[t] -= lr * (mt / c1) / (sqrtf(vt / c2) + eps);
}

// ===========================================================================
// BF16 MFMA inference.  A wave evaluates 16 cells at a time with the
// transposed form h' = W^T h: the output tile of v_mfma_f32_16x16x* (lane l
// holds rows 4*(l/16) .. +3 of column l%16) is, after GELU and rounding to
// BF16, exactly the operand layout of the next layer's k-slice, so
// activations never leave registers.  Layer 1 (K = 7 -> 16) uses the
// 16x16x16 form; layers 2-4 use the gfx950 16x16x32 form with the k order
// permuted identically in both operands: k slot v of lane group g stands for
// neuron 32*kt + (v < 4 ? 4g + v : 16 + 4g + v - 4).  Weights are held in
// registers (converted once per wave), biases in LDS.
// ===========================================================================
namespace {

typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

struct MlpRegs {
  s16x4 a1[4];      // layer 1, per output tile
  bf16x8 a2[4][2];  // layer 2, [out tile][k tile]
  bf16x8 a3[4][2];
  bf16x8 a4[2];     // layer 4, [k tile]
};


__device__ inline int perm_k(int kt, int g, int v) {
  return 32 * kt + (v < 4 ? 4 * g + v : 16 + 4 * g + (v - 4));
}

__device__ inline void load_regs(const float* w, MlpRegs& R, int c, int g) {
  const float* W1 = w + kSurW1;
  const float* W2 = w + kSurW2;
  const float* W3 = w + kSurW3;
  const float* W4 = w + kSurW4;
  for (int mt = 0; mt < 4; ++mt) {
    s16x4 a;
    for (int v = 0; v < 4; ++v) {
      const int k = 4 * g + v;
      a[v] = static_cast<short>(to_bf16(k < kSurIn ? W1[k * 64 + mt * 16 + c] : 0.0f));
    }
    R.a1[mt] = a;
    for (int kt = 0; kt < 2; ++kt) {
      bf16x8 b2, b3;
      for (int v = 0; v < 8; ++v) {
        const int k = perm_k(kt, g, v);
        b2[v] = bf16_of(W2[k * 64 + mt * 16 + c]);
        b3[v] = bf16_of(W3[k * 64 + mt * 16 + c]);
      }
      R.a2[mt][kt] = b2;
      R.a3[mt][kt] = b3;
    }
  }
  for (int kt = 0; kt < 2; ++kt) {
    bf16x8 b4;
    for (int v = 0; v < 8; ++v) {
      const int k = perm_k(kt, g, v);
      b4[v] = bf16_of(c < kSurOut ? W4[k * kSurOut + c] : 0.0f);
    }
    R.a4[kt] = b4;
  }
}

// Biases into LDS: sb[0..191] hidden layers, sb[192..195] outputs.
__device__ inline void load_bias(const float* w, float* sb) {
  for (int t = threadIdx.x + blockDim.x * threadIdx.y; t < 196; t += blockDim.x * blockDim.y) {
    if (t < 64) sb[t] = w[kSurB1 + t];
    else if (t < 128) sb[t] = w[kSurB2 + t - 64];
    else if (t < 192) sb[t] = w[kSurB3 + t - 128];
    else sb[t] = w[kSurB4 + t - 192];
  }
}

// GELU of two accumulator quads, rounded to BF16 in pairs (one
// v_cvt_pk_bf16_f32 per two values) and packed as the next layer's k-slice:
// slots 0-3 from lo, 4-7 from hi.
__device__ inline bf16x8 gelu_pack(const f32x4& lo, const f32x4& hi) {
  auto cvt2 = [](float a, float b) {
    const f32x2 v = {gelu_fast(a), gelu_fast(b)};
    return __builtin_convertvector(v, bf16x2);
  };
  const bf16x2 p0 = cvt2(lo[0], lo[1]), p1 = cvt2(lo[2], lo[3]);
  const bf16x2 p2 = cvt2(hi[0], hi[1]), p3 = cvt2(hi[2], hi[3]);
  const bf16x4 q0 = __builtin_shufflevector(p0, p1, 0, 1, 2, 3);
  const bf16x4 q1 = __builtin_shufflevector(p2, p3, 0, 1, 2, 3);
  return __builtin_shufflevector(q0, q1, 0, 1, 2, 3, 4, 5, 6, 7);
}

// One group of 16 cells.  xin: this lane's normalised features 4g .. 4g+3
// (zero for g >= 2).  Returns, for lanes with g == 0, the normalised
// outputs 0..3 of cell l % 16.
__device__ inline f32x4 mlp_group(const MlpRegs& R, const float* sb, const float xin[4], int g) {
  s16x4 bx;
  for (int v = 0; v < 4; ++v) bx[v] = static_cast<short>(to_bf16(xin[v]));
  f32x4 h[4];
  for (int mt = 0; mt < 4; ++mt) {
    f32x4 acc;
    for (int r = 0; r < 4; ++r) acc[r] = sb[mt * 16 + 4 * g + r];
    h[mt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(R.a1[mt], bx, acc, 0, 0, 0);
  }
  for (int layer = 0; layer < 2; ++layer) {
    bf16x8 bk[2];
    for (int kt = 0; kt < 2; ++kt) bk[kt] = gelu_pack(h[2 * kt], h[2 * kt + 1]);
    const float* b = sb + 64 * (layer + 1);
    for (int mt = 0; mt < 4; ++mt) {
      f32x4 acc;
      for (int r = 0; r < 4; ++r) acc[r] = b[mt * 16 + 4 * g + r];
      const bf16x8* a = layer == 0 ? R.a2[mt] : R.a3[mt];
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], bk[0], acc, 0, 0, 0);
      h[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], bk[1], acc, 0, 0, 0);
    }
  }
  bf16x8 bk[2];
  for (int kt = 0; kt < 2; ++kt) bk[kt] = gelu_pack(h[2 * kt], h[2 * kt + 1]);
  f32x4 acc;
  for (int r = 0; r < 4; ++r) acc[r] = g == 0 ? sb[192 + r] : 0.0f;
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(R.a4[0], bk[0], acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(R.a4[1], bk[1], acc, 0, 0, 0);
  return acc;
}

}  // namespace

// Both inference kernels run in tiles of 256 cells per work-group (64 per
// wave, one per lane): each lane prepares its own cell's features and writes
// them, normalised, to LDS; the wave then evaluates its 64 cells as four
// 16-cell MFMA groups, leaving the outputs in LDS; finally each lane applies
// its own cell's outputs.  The trip count is uniform over the work-group, so
// the barriers are too.
constexpr int kTile = 256;

// One 16-cell group: rows feat[0..15] -> out[0..15].
__device__ inline void group_mlp(const MlpRegs& R, const float* sb, const float (*feat)[8],
                                 float (*out)[kSurOut], int lane) {
  const int c = lane & 15, g = lane >> 4;
  float xin[4];
  for (int v = 0; v < 4; ++v) xin[v] = g < 2 ? feat[c][4 * g + v] : 0.0f;
  const f32x4 o = mlp_group(R, sb, xin, g);
  if (g == 0)
    for (int r = 0; r < kSurOut; ++r) out[c][r] = o[r];
}

__device__ inline void tile_mlp(const MlpRegs& R, const float* sb, const float (*feat)[8],
                                float (*out)[kSurOut], int lane) {
  for (int q = 0; q < 4; ++q) group_mlp(R, sb, feat + 16 * q, out + 16 * q, lane);
}

// Raw features x[n][7] -> raw outputs y[n][4] (tests and timing).
__global__ __launch_bounds__(256) void k_sur_infer(const float* w, SurNorm nm, const float* x,
                                                   long n, float* y) {
  __shared__ float sb[196];
  __shared__ float sfeat[4][64][8];
  __shared__ float sout[4][64][kSurOut];
  load_bias(w, sb);
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  MlpRegs R;
  load_regs(w, R, lane & 15, lane >> 4);
  for (long tile = blockIdx.x; tile * kTile < n; tile += gridDim.x) {
    const long cell = tile * kTile + wv * 64 + lane;
    const bool valid = cell < n;
    for (int q = 0; q < 8; ++q)
      sfeat[wv][lane][q] = (valid && q < kSurIn) ? qdiv(x[cell * kSurIn + q] - nm.xm[q], nm.xs[q])
                                                 : 0.0f;
    __syncthreads();
    tile_mlp(R, sb, sfeat[wv], sout[wv], lane);
    __syncthreads();
    if (valid)
      for (int q = 0; q < kSurOut; ++q) y[cell * kSurOut + q] = sout[wv][lane][q] * nm.ys[q] + nm.ym[q];
  }
}

// Warm-rain source terms with the surrogate (fallout runs separately):
// cells whose features all lie in the box use the network's tendencies,
// the others the plain scheme (counted as fallback cells); cells without
// condensate at or below saturation keep their state, the plain scheme's
// exact answer there.  Two passes (SurArgs): classify, then network.

namespace {

// Rain fallout of the post-adjustment rain (K25 nominal / K26 metric,
// warm_rain.cu:122-191), one level at a time down a column: `inc` carries
// the rain entering the level from above.
struct Fallout {
  float dt, vt, dz_nom, rdz_nom, frac_nom;
  bool nominal, metric;
  __device__ Fallout(float dt_, float vt_, float depth, int mode) : dt(dt_), vt(vt_) {
    dz_nom = fmaxf(depth, 1.0e-3f);
    frac_nom = fminf(fmaxf(vt * dt / dz_nom, 0.0f), 1.0f);
    rdz_nom = 1.0f / dz_nom;
    // nothing falls out: the nominal sweep leaves the column alone
    // (warm_rain.cu:134-136)
    nominal = mode == 1 && frac_nom > 0.0f;
    metric = mode == 2;
  }
  __device__ bool on() const { return nominal || metric; }
  template <bool TERRAIN>
  __device__ float level(float rqr, float& inc, const Geo& g, const float* idz_c,
                         const float* lev_idz_c, int i, int j, int k) const {
    if (nominal) {
      const float old = fmaxf(rqr, 0.0f) * dz_nom;
      const float out = frac_nom * old;
      rqr = qdivc(old - out + inc, dz_nom, rdz_nom);
      inc = out;
    } else {
      const float idz = fmaxf(TERRAIN ? idz_c[g.metric(i, j, k)] : lev_idz_c[k], 1.0e-6f);
      const float dz = qrcp(idz);
      const float frac = fminf(fmaxf(vt * dt * idz, 0.0f), 1.0f);
      const float old = fmaxf(rqr, 0.0f) * dz;
      const float out = frac * old;
      rqr = (old - out + inc) * idz;
      inc = out;
    }
    return rqr;
  }
};

// Features of a cell and its class.  Without cloud or rain and at or below
// saturation every source term of the plain scheme is zero (no
// condensation, nothing to evaporate, convert or accrete): such a cell is
// quiet and takes the plain scheme's answer, its own state written back.
struct SurCell {
  MicroCell<float> m;
  bool inbox, quiet;
};
__device__ inline SurCell sur_cell(float rho, float rt0, float rqv, float rqc, float rqr,
                                   const MicroCfg<float>& mc, const SurNorm& nm) {
  SurCell s;
  s.m = micro_prepare(rho, rt0, rqv, rqc, rqr);
  const float f[kSurIn] = {s.m.T, s.m.p, s.m.rho, s.m.qv, s.m.qc, s.m.qr, mc.dt};
  bool in = true;
  for (int q = 0; q < kSurIn; ++q) in = in && f[q] >= nm.lo[q] && f[q] <= nm.hi[q];
  s.inbox = in;
  s.quiet = s.m.qc <= 0.0f && s.m.qr <= 0.0f && s.m.qv <= q_qsat(s.m.T, s.m.p);
  return s;
}

}  // namespace

// Pass 1: a work-group of 64 x 4 threads owns a 64 x 4 column tile, a wave
// one row of it (its queue segment), a lane one column, walked top down
// with the loads of the next two levels in flight.  Fallback cells get the
// plain scheme, quiet cells their state written back; network cells are
// appended to the wave's segment in (level, lane) order (ballot + mbcnt,
// no atomics) and left for pass 2.  Lane 0 stores the segment's network
// and fallback counts.  The rain fallout follows each column down to its
// first network cell (all the way, and into the bucket, in a column
// without one); the column's stop level and the rain entering it are kept
// for the fallout pass.
template <bool TERRAIN>
__device__ __forceinline__ void sur_classify_body(const SurArgs& a, const VBlock& vb) {
  const Geo g = a.g;
  int bx, by;
  xcd_block2d(vb, bx, by);
  const int lane = threadIdx.x;
  const int i = bx * 64 + lane;
  const int j = by * 4 + threadIdx.y;
  if (j >= g.ny) return;  // the whole wave
  const int tiles_x = (g.nx + 63) >> 6;
  const int nseg = tiles_x * g.ny;
  const int seg = j * tiles_x + bx;
  const bool col = i < g.nx;
  unsigned* const q = a.queue + static_cast<size_t>(seg) * 64u * static_cast<unsigned>(g.nz);
  const bool pert = a.base.rho != nullptr;
  const size_t plane = static_cast<size_t>(g.pxc) * g.pyc;
  struct Lv {
    float rho, rt, qv, qc, qr;
  };
  auto load = [&](size_t cc) -> Lv {
    return {a.rho[cc], a.rt[cc], a.qv[cc], a.qc[cc], a.qr[cc]};
  };
  size_t c = g.cell(col ? i : 0, j, g.nz - 1);
  Lv l1{}, l2{};
  if (col) {
    l1 = load(c);
    if (g.nz > 1) l2 = load(c - plane);
  }
  unsigned nnet = 0, nfall = 0;  // wave-uniform
  const Fallout fo(a.fdt, a.vt, a.depth, a.fallout);
  float inc = 0.0f, inc_stop = 0.0f;
  int stop = -1;
  for (int k = g.nz - 1; k >= 0; --k, c -= plane) {
    const Lv cur = l1;
    l1 = l2;
    if (col && k >= 2) l2 = load(c - 2 * plane);
    SurCell s{};
    if (col)
      s = sur_cell(pert ? cur.rho + a.base.rho[k] : cur.rho, pert ? cur.rt + a.base.rt[k] : cur.rt,
                   cur.qv, cur.qc, cur.qr, a.mc, a.nm);
    const bool net = col && !s.quiet && s.inbox;
    const unsigned long long bal = __ballot(net);
    nfall += static_cast<unsigned>(__popcll(__ballot(col && !s.inbox)));
    if (net) {
      const unsigned below = __builtin_amdgcn_mbcnt_hi(
          static_cast<unsigned>(bal >> 32), __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(bal), 0u));
      q[nnet + below] = static_cast<unsigned>(c);
      if (stop < 0) {
        stop = k;
        inc_stop = inc;
      }
    } else if (col) {
      MicroCell<float>& m = s.m;
      const float th0 = m.th;
      if (!s.quiet) micro_apply(m, a.mc);
      a.qv[c] = m.rho * m.qv;
      a.qc[c] = m.rho * m.qc;
      float rqr = m.rho * m.qr;
      if (fo.on() && stop < 0) rqr = fo.level<TERRAIN>(rqr, inc, g, a.idz_c, a.lev_idz_c, i, j, k);
      a.qr[c] = rqr;
      // perturbation form: the departure changes by rho * dtheta only
      a.rt[c] = pert ? cur.rt + m.rho * (m.th - th0) : m.rho * m.th;
    }
    nnet += static_cast<unsigned>(__popcll(bal));
  }
  if (lane == 0) {
    a.segs[seg] = nnet;
    a.segs[nseg + seg] = nfall;
  }
  if (col && fo.on()) {
    const size_t cc = static_cast<size_t>(j) * g.nx + i;
    if (stop < 0) a.precip[cc] += inc;
    a.col_stop[cc] = stop;
    a.col_inc[cc] = inc_stop;
  }
}

// Pass 2: work-group b takes segments b, b + G, ... in tiles of 256 listed
// cells (one per lane; rows past the segment's count are zero).  Each lane
// prepares its cell's features into LDS; the waves share the tile's 16-cell
// MFMA groups; each lane then applies its own cell's outputs.  Every row of
// a group except the last tile's tail holds a network cell, and each cell's
// outputs depend on its own row only (the matrix product keeps columns
// apart), so the results equal those of running every cell through the
// network alone.  A work-group first adds its segments' counts to
// a.counts (one atomic per counter) and returns at once when they list no
// network cell, before loading the weights.
__device__ __forceinline__ void sur_network_body(const SurArgs& a, const VBlock& vb) {
  __shared__ float sb[196];
  __shared__ float sfeat[kTile][8];
  __shared__ float sout[kTile][kSurOut];
  const Geo g = a.g;
  const int tiles_x = (g.nx + 63) >> 6;
  const int nseg = tiles_x * g.ny;
  unsigned net = 0, fall = 0;
  for (int s = vb.x; s < nseg; s += vb.gx) {
    net += a.segs[s];
    fall += a.segs[nseg + s];
  }
  if (threadIdx.x == 0) {
    if (fall) atomicAdd(&a.counts[0], static_cast<unsigned long long>(fall));
    if (net) atomicAdd(&a.counts[1], static_cast<unsigned long long>(net));
  }
  if (net == 0) return;  // uniform over the work-group
  load_bias(a.w, sb);
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  MlpRegs R;
  load_regs(a.w, R, lane & 15, lane >> 4);
  const SurNorm& nm = a.nm;
  const bool pert = a.base.rho != nullptr;
  const unsigned plane = static_cast<unsigned>(g.pxc) * g.pyc;
  const size_t cap = 64u * static_cast<size_t>(g.nz);
  for (int s = vb.x; s < nseg; s += vb.gx) {
    const unsigned cnt = a.segs[s];
    const unsigned* q = a.queue + static_cast<size_t>(s) * cap;
    for (unsigned t0 = 0; t0 < cnt; t0 += kTile) {
      const unsigned idx = t0 + threadIdx.x;
      const bool valid = idx < cnt;
      unsigned c = 0;
      float rtp = 0.0f;
      MicroCell<float> m{};
      if (valid) {
        c = q[idx];
        const unsigned k = c / plane;
        rtp = a.rt[c];
        m = micro_prepare(pert ? a.rho[c] + a.base.rho[k] : a.rho[c],
                          pert ? rtp + a.base.rt[k] : rtp, a.qv[c], a.qc[c], a.qr[c]);
        const float f[kSurIn] = {m.T, m.p, m.rho, m.qv, m.qc, m.qr, a.mc.dt};
        for (int v = 0; v < kSurIn; ++v) sfeat[threadIdx.x][v] = qdiv(f[v] - nm.xm[v], nm.xs[v]);
        sfeat[threadIdx.x][7] = 0.0f;
      } else {
        for (int v = 0; v < 8; ++v) sfeat[threadIdx.x][v] = 0.0f;
      }
      __syncthreads();
      const unsigned rows = cnt - t0 < kTile ? cnt - t0 : kTile;
      for (unsigned gq = wv; 16u * gq < rows; gq += kTile / 64)
        group_mlp(R, sb, sfeat + 16 * gq, sout + 16 * gq, lane);
      __syncthreads();
      if (valid) {
        const float* o = sout[threadIdx.x];
        // The source terms only exchange water between vapour, cloud and
        // rain.  Cloud and rain cannot lose more than they hold; vapour, the
        // reservoir, takes the opposite of their change (the network's own
        // vapour change is not applied), so total water is conserved to
        // rounding; in the rare cell where that would exhaust the vapour,
        // the exchange and dtheta are scaled back so that it just does.
        const float dc = fmaxf(o[1] * nm.ys[1] + nm.ym[1], -m.qc);
        const float dr = fmaxf(o[2] * nm.ys[2] + nm.ym[2], -m.qr);
        const float dth = o[3] * nm.ys[3] + nm.ym[3];
        const float dv = -(dc + dr);
        const float lam = m.qv + dv < 0.0f ? fmaxf(qdiv(m.qv, -dv), 0.0f) : 1.0f;
        const float th0 = m.th;
        m.qv = fmaxf(m.qv + lam * dv, 0.0f);
        m.qc = fmaxf(m.qc + lam * dc, 0.0f);
        m.qr = fmaxf(m.qr + lam * dr, 0.0f);
        m.th = m.th + lam * dth;
        a.qv[c] = m.rho * m.qv;
        a.qc[c] = m.rho * m.qc;
        a.qr[c] = m.rho * m.qr;
        a.rt[c] = pert ? rtp + m.rho * (m.th - th0) : m.rho * m.th;
      }
    }
  }
}

template <bool TERRAIN>
__global__ __launch_bounds__(256) void k_sur_classify(SurArgs a) {
  sur_classify_body<TERRAIN>(a, hw_block());
}
template <bool TERRAIN>
__global__ __launch_bounds__(256) void k_sur_classify_batch(const SurItem* __restrict__ items,
                                                            BatchMap bm) {
  VBlock vb;
  const SurItem* it = batch_member(items, bm, vb);
  if (!it) return;
  const SurArgs a = it->a;
  sur_classify_body<TERRAIN>(a, vb);
}
__global__ __launch_bounds__(256) void k_sur_network(SurArgs a) {
  sur_network_body(a, hw_block());
}
__global__ __launch_bounds__(256) void k_sur_network_batch(const SurItem* __restrict__ items,
                                                           BatchMap bm) {
  VBlock vb;
  const SurItem* it = batch_member(items, bm, vb);
  if (!it) return;
  const SurArgs a = it->a;
  sur_network_body(a, vb);
}

// Rain fallout, one thread per column, top down from level nz-1, or (with
// col_stop) from the column's col_stop level with col_inc entering it.
template <bool TERRAIN>
__device__ __forceinline__ void fallout_body(const FalloutItem& a, const VBlock& vb) {
  const Geo& g = a.g;
  int bx, by;
  xcd_block2d(vb, bx, by);
  const int i = bx * blockDim.x + threadIdx.x;
  const int j = by * blockDim.y + threadIdx.y;
  if (i >= g.nx || j >= g.ny) return;
  const Fallout fo(a.dt, a.vt, a.depth, a.fallout);
  if (!fo.on()) return;
  const size_t cc = static_cast<size_t>(j) * g.nx + i;
  int k0 = g.nz - 1;
  float inc = 0.0f;
  if (a.col_stop) {
    k0 = a.col_stop[cc];
    if (k0 < 0) return;  // finished by the classify pass
    inc = a.col_inc[cc];
  }
  // unrolled so that the independent loads of several levels are in flight
#pragma unroll 4
  for (int k = k0; k >= 0; --k) {
    const size_t c = g.cell(i, j, k);
    a.qr[c] = fo.level<TERRAIN>(a.qr[c], inc, g, a.idz_c, a.lev_idz_c, i, j, k);
  }
  a.precip[cc] += inc;
}

template <bool TERRAIN>
__global__ void k_fallout(FalloutItem a) {
  fallout_body<TERRAIN>(a, hw_block());
}
template <bool TERRAIN>
__global__ void k_fallout_batch(const FalloutItem* __restrict__ items, BatchMap bm) {
  VBlock vb;
  const FalloutItem* it = batch_member(items, bm, vb);
  if (!it) return;
  const FalloutItem a = *it;
  fallout_body<TERRAIN>(a, vb);
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
  if (static_cast<int>(threadIdx.x) < n) p[threadIdx.x] = 0ull;
}
__global__ void k_zero_u64_batch(const ZeroItem* __restrict__ items, BatchMap bm) {
  VBlock vb;
  const ZeroItem* it = batch_member(items, bm, vb);
  if (!it) return;
  if (static_cast<int>(threadIdx.x) < it->n) it->p[threadIdx.x] = 0ull;
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
namespace {
inline unsigned nblk(long n, int b) { return static_cast<unsigned>((n + b - 1) / b); }
inline unsigned infer_blocks(long cells) {
  const long tiles = (cells + kTile - 1) / kTile;
  return static_cast<unsigned>(tiles < 2048 ? tiles : 2048);
}
}  // namespace

void launch_sur_sample(hipStream_t st, const Geo& g, const float* rho, const float* rt,
                       const float* qv, const float* qc, const float* qr,
                       const MicroCfg<float>& mc, unsigned long long seed, long n, float* x,
                       float* y, long long* cell, Base<float> base) {
  if (refuse_record("sur_sample")) return;
  hipLaunchKernelGGL(k_sur_sample, dim3(nblk(n, 256)), dim3(256), 0, st, g, rho, rt, qv, qc, qr,
                     mc, seed, n, x, y, cell, base);
}

void launch_sur_train_step(hipStream_t st, SurTrain& T, const float* x, const float* y, int B,
                           const SurNorm& nm, bool update, double* loss_part, int* nloss) {
  if (refuse_record("sur_train_step")) return;
  float* p = T.params;
  // forward
  hipLaunchKernelGGL(k_sur_norm_in, dim3(nblk(static_cast<long>(B) * 8, 256)), dim3(256), 0, st,
                     x, static_cast<long>(B), nm, T.xn);
  const float* in = T.xn;
  int ldi = 8, K = kSurIn;
  const int wofs[4] = {kSurW1, kSurW2, kSurW3, kSurW4};
  const int bofs[4] = {kSurB1, kSurB2, kSurB3, kSurB4};
  for (int l = 0; l < 4; ++l) {
    const int N = l < 3 ? 64 : kSurOut;
    hipLaunchKernelGGL(k_dense_fwd, dim3(nblk(static_cast<long>(B) * N, 256)), dim3(256), 0, st,
                       in, ldi, K, p + wofs[l], p + bofs[l], N, B, l < 3 ? 1 : 0, T.z[l],
                       l < 3 ? T.a[l] : nullptr);
    in = l < 3 ? T.a[l] : nullptr;
    ldi = 64;
    K = 64;
  }
  // loss and output gradient
  const unsigned lb = nblk(static_cast<long>(B) * kSurOut, 256);
  hipLaunchKernelGGL(k_mse_grad, dim3(lb), dim3(256), 0, st, T.z[3], y, B, nm, T.dz[3],
                     loss_part);
  if (nloss) *nloss = static_cast<int>(lb);
  // backward
  const int slices = 16;
  for (int l = 3; l >= 0; --l) {
    const int N = l < 3 ? 64 : kSurOut;
    const int Kl = l == 0 ? kSurIn : 64;
    const float* inl = l == 0 ? T.xn : T.a[l - 1];
    const int ldl = l == 0 ? 8 : 64;
    const long parts = static_cast<long>(Kl + 1) * N * slices;
    hipLaunchKernelGGL(k_dense_bwd_weight, dim3(nblk(parts, 256)), dim3(256), 0, st, inl, ldl,
                       Kl, T.dz[l], N, B, slices, T.wpart);
    hipLaunchKernelGGL(k_dense_grad_reduce, dim3(nblk(static_cast<long>(Kl + 1) * N, 256)),
                       dim3(256), 0, st, T.wpart, Kl, N, slices, T.grad + wofs[l],
                       T.grad + bofs[l]);
    if (l > 0)
      hipLaunchKernelGGL(k_dense_bwd_data, dim3(nblk(static_cast<long>(B) * 64, 256)), dim3(256),
                         0, st, T.dz[l], N, p + wofs[l], 64, T.z[l - 1], B, T.dz[l - 1]);
  }
  if (update) {
    T.step += 1;
    const float c1 = 1.0f - powf(T.beta1, static_cast<float>(T.step));
    const float c2 = 1.0f - powf(T.beta2, static_cast<float>(T.step));
    hipLaunchKernelGGL(k_adam, dim3(nblk(kSurParams, 256)), dim3(256), 0, st, p, T.grad, T.m, T.v,
                       kSurParams, T.lr, T.beta1, T.beta2, T.eps, c1, c2);
  }
}

void launch_sur_infer(hipStream_t st, const float* w, const SurNorm& nm, const float* x, long n,
                      float* y) {
  if (refuse_record("sur_infer")) return;
  hipLaunchKernelGGL(k_sur_infer, dim3(infer_blocks(n)), dim3(256), 0, st, w, nm, x, n, y);
}

void launch_zero_counters(hipStream_t st, unsigned long long* p, int n) {
  if (recorder()) {
    record_launch(reinterpret_cast<const void*>(&k_zero_u64_batch), dim3(64), 0,
                  ZeroItem{p, n, 0, 0}, 1, 1);
    return;
  }
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, p, n);
}

void launch_fallout(hipStream_t st, const Geo& g, float* qr, float* precip, const float* idz_c,
                    const float* lev_idz_c, float dt, float vt, float depth, int fallout,
                    bool terrain, const int* col_stop, const float* col_inc) {
  const dim3 b(64, 4, 1), gr((g.nx + 63) / 64, (g.ny + 3) / 4);
  const FalloutItem it{g, qr, precip, idz_c, lev_idz_c, dt, vt, depth, fallout,
                       col_stop, col_inc, 0, 0};
  if (recorder()) {
    const void* fn = terrain ? reinterpret_cast<const void*>(&k_fallout_batch<true>)
                             : reinterpret_cast<const void*>(&k_fallout_batch<false>);
    record_launch(fn, b, 0, it, static_cast<int>(gr.x), static_cast<int>(gr.y));
    return;
  }
  if (terrain)
    hipLaunchKernelGGL(k_fallout<true>, gr, b, 0, st, it);
  else
    hipLaunchKernelGGL(k_fallout<false>, gr, b, 0, st, it);
}

void launch_warm_rain_sur(hipStream_t st, const SurArgs& a, bool terrain) {
  const dim3 b1(64, 4, 1), g1((a.g.nx + 63) / 64, (a.g.ny + 3) / 4);
  const long nseg = sur_segments(a.g);
  const int g2 = static_cast<int>(nseg < 2048 ? nseg : 2048);
  if (recorder()) {
    const void* fn = terrain ? reinterpret_cast<const void*>(&k_sur_classify_batch<true>)
                             : reinterpret_cast<const void*>(&k_sur_classify_batch<false>);
    record_launch(fn, b1, 0, SurItem{a, 0, 0}, static_cast<int>(g1.x), static_cast<int>(g1.y));
    record_launch(reinterpret_cast<const void*>(&k_sur_network_batch), dim3(256), 0,
                  SurItem{a, 0, 0}, g2, 1);
  } else {
    if (terrain)
      hipLaunchKernelGGL(k_sur_classify<true>, g1, b1, 0, st, a);
    else
      hipLaunchKernelGGL(k_sur_classify<false>, g1, b1, 0, st, a);
    hipLaunchKernelGGL(k_sur_network, dim3(g2), dim3(256), 0, st, a);
  }
  if (a.fallout)
    launch_fallout(st, a.g, a.qr, a.precip, a.idz_c, a.lev_idz_c, a.fdt, a.vt, a.depth, a.fallout,
                   terrain, a.col_stop, a.col_inc);
}

}  // namespace gwm
