// Learned warm start of the space-time sites (no counterpart in the reference: it
// replaces the fixed site initialisation N(0, 1e2 I) of basemodels.py:130-133 through
// update_nat_params, basemodels.py:97-100).
//
// A 3 x 128 SiLU MLP maps per-site features of a local data window and the hypers to
// the converged posterior marginal (m, log v) of f at that site; the initial site is
// then the VI site step taken at that marginal (the fixed-point condition of natural-
// gradient VI, inference.py:160-195 + 21-39).  Weights are BF16, every product
// accumulates in FP32.  Training is mini-batch Adam on FP32 master weights.
//
// Kernels:
//   k_ws_features    block per (step, run): kernel-smoothed data over a +-W step window
//   k_mlp_forward    64 rows per block: features -> (m, log v)
//   k_mlp_grad       32 rows per block: forward + backward of the mean squared error,
//                    per-block partial gradients (no atomics)
//   k_mlp_reduce     per parameter: sum of block partials, optional Adam step, BF16 copy
//   k_ws_sites       per site: VI site step at the predicted marginal
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "bn_internal.h"
#include "bn_sites.h"

using namespace bn;

namespace {

constexpr int kT = 256;
constexpr int kF = BN_WS_FEATURES;  // input features
constexpr int kH = 128;             // hidden width
constexpr int kO = 2;               // outputs (m, log v)
// flat parameter layout (FP32 master and BF16 copy share it)
constexpr int oW1 = 0, oB1 = oW1 + kF * kH, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH,
              oB3 = oW3 + kH * kH, oW4 = oB3 + kH, oB4 = oW4 + kH * kO, kP = oB4 + kO;
static_assert(kP == BN_WS_PARAMS, "parameter count");

int ws_fail(int code, const std::string& m) { return set_error(code, m); }

#define WS_HIP_CHECK(expr)                                                                  \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) return ws_fail(BN_EHIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

// BF16 <-> FP32: BF16 is the top half of an IEEE binary32; rounding to nearest even.
__device__ inline float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ inline uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);  // NaN stays NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ inline float sigm(float z) { return 1.0f / (1.0f + expf(-z)); }
__device__ inline float silu(float z) { return z * sigm(z); }
__device__ inline float dsilu(float z) {
  const float s = sigm(z);
  return s * (1.0f + z * (1.0f - s));
}

struct Norm {
  float xmu[kF], xsd[kF];  // feature standardisation
  float ymu[kO], ysd[kO];  // output scaling
};

// ---------------------------------------------------------------------------
// Features of site (k, t, i):
//   0     log(1 + y)
//   1..3  log of the kernel-smoothed rate (sum w y + 1) / (sum w + 1) / binsize over the
//         steps t-W..t+W, weights exp(-|r - r'|^2 / 2hs^2 - (t - t')^2 / 2ht^2)
//   4..7  log of the hypers: temporal lengthscale, spatial lengthscale, temporal variance,
//         binsize
// ---------------------------------------------------------------------------
constexpr int kW = 6;
constexpr float kHs[3] = {0.07f, 0.15f, 0.30f};
constexpr float kHt[3] = {1.0f, 2.5f, 6.0f};

__global__ __launch_bounds__(kT) void k_ws_features(int32_t K, int32_t T, int32_t R, const float* __restrict__ r,
                                                    const float* __restrict__ y, const double* __restrict__ tt,
                                                    float h0, float h1, float h2, float binsize,
                                                    float* __restrict__ X) {
  extern __shared__ float sws[];  // (2W+1) * R * 3: x, y-location, count
  const int t = blockIdx.x, k = blockIdx.y, tid = threadIdx.x;
  const int t0 = max(0, t - kW), t1 = min(T - 1, t + kW);
  const int ns = t1 - t0 + 1;
  for (int e = tid; e < ns * R; e += kT) {
    const int64_t s = ((int64_t)k * T + t0) * R + e;
    sws[3 * e] = r[2 * s];
    sws[3 * e + 1] = r[2 * s + 1];
    const float yy = y[s];
    sws[3 * e + 2] = (yy == yy) ? yy : -1.0f;  // missing: excluded below
  }
  __syncthreads();
  const double tc = tt[(int64_t)k * T + t];
  for (int i = tid; i < R; i += kT) {
    const int64_t s = ((int64_t)k * T + t) * R + i;
    const float rx = r[2 * s], ry = r[2 * s + 1], yy = y[s];
    float num[3] = {0.f, 0.f, 0.f}, den[3] = {0.f, 0.f, 0.f};
    for (int q = 0; q < ns; ++q) {
      const float dtq = (float)(tt[(int64_t)k * T + t0 + q] - tc);
      float wt[3];
#pragma unroll
      for (int b = 0; b < 3; ++b) wt[b] = -0.5f * dtq * dtq / (kHt[b] * kHt[b]);
      for (int j = 0; j < R; ++j) {
        const float* p = sws + 3 * (q * R + j);
        const float cy = p[2];
        if (cy < 0.f) continue;
        const float dx = p[0] - rx, dy = p[1] - ry;
        const float d2 = dx * dx + dy * dy;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          const float w = expf(wt[b] - 0.5f * d2 / (kHs[b] * kHs[b]));
          num[b] = fmaf(w, cy, num[b]);
          den[b] += w;
        }
      }
    }
    float* xo = X + s * kF;
    xo[0] = log1pf((yy == yy) ? yy : 0.f);
#pragma unroll
    for (int b = 0; b < 3; ++b) xo[1 + b] = logf((num[b] + 1.0f) / (den[b] + 1.0f) / binsize);
    xo[4] = logf(h0);
    xo[5] = logf(h1);
    xo[6] = logf(h2);
    xo[7] = logf(binsize);
  }
}

// ---------------------------------------------------------------------------
// MLP layers on a block of rows.  Activations are stored transposed, act[c * ld + r].
// ---------------------------------------------------------------------------
// z[c][r] = b[c] + sum_k in(k, r) W[k][c] for NR rows, cols 0..127; in(k, r) = f(src[k*ld + r])
template <int NR, bool SILU_IN>
__device__ void layer_fwd(const float* src, int kin, int ld, const uint16_t* W, const float* b, float* dst) {
  const int tid = threadIdx.x;
  constexpr int RPT = NR / 8;  // rows per thread
  const int tr = tid >> 5, tc = tid & 31;
  float acc[RPT][4];
#pragma unroll
  for (int a = 0; a < RPT; ++a)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[a][c] = b[tc * 4 + c];
  for (int k = 0; k < kin; ++k) {
    float w[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) w[c] = bf2f(W[k * kH + tc * 4 + c]);
#pragma unroll
    for (int a = 0; a < RPT; ++a) {
      float v = src[k * ld + tr * RPT + a];
      if (SILU_IN) v = silu(v);
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[a][c] = fmaf(v, w[c], acc[a][c]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int a = 0; a < RPT; ++a)
#pragma unroll
    for (int c = 0; c < 4; ++c) dst[(tc * 4 + c) * ld + tr * RPT + a] = acc[a][c];
  __syncthreads();
}

__device__ void load_weights(const uint16_t* wbf, uint16_t* sW) {
  for (int e = threadIdx.x; e < kP; e += kT) sW[e] = wbf[e];
}

constexpr int kFR = 64;  // rows per forward block
constexpr size_t kFwdSmem = sizeof(float) * (2 * kH * kFR + kF * kFR) + sizeof(uint16_t) * (kP + 8);

__global__ __launch_bounds__(kT) void k_mlp_forward(int64_t N, const float* __restrict__ X,
                                                    const uint16_t* __restrict__ wbf, const float* __restrict__ pf,
                                                    Norm nm, float* __restrict__ out) {
  extern __shared__ float sf[];
  float* A0 = sf;                    // [kH][kFR]
  float* A1 = A0 + kH * kFR;         // [kH][kFR]
  float* X0 = A1 + kH * kFR;         // [kF][kFR]
  uint16_t* sW = (uint16_t*)(X0 + kF * kFR);
  __shared__ float sb[3 * kH + kO];
  const int tid = threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.x * kFR;
  load_weights(wbf, sW);
  for (int e = tid; e < 3 * kH + kO; e += kT)
    sb[e] = pf[e < kH ? oB1 + e : e < 2 * kH ? oB2 + e - kH : e < 3 * kH ? oB3 + e - 2 * kH : oB4 + e - 3 * kH];
  for (int e = tid; e < kF * kFR; e += kT) {
    const int r = e / kF, f = e % kF;
    const int64_t row = r0 + r;
    X0[f * kFR + r] = row < N ? (X[row * kF + f] - nm.xmu[f]) / nm.xsd[f] : 0.f;
  }
  __syncthreads();
  layer_fwd<kFR, false>(X0, kF, kFR, sW + oW1, sb, A0);
  layer_fwd<kFR, true>(A0, kH, kFR, sW + oW2, sb + kH, A1);
  layer_fwd<kFR, true>(A1, kH, kFR, sW + oW3, sb + 2 * kH, A0);
  if (tid < kFR * kO) {
    const int r = tid >> 1, o = tid & 1;
    float s = sb[3 * kH + o];
    for (int k = 0; k < kH; ++k) s = fmaf(silu(A0[k * kFR + r]), bf2f(sW[oW4 + k * kO + o]), s);
    const int64_t row = r0 + r;
    if (row < N) out[row * kO + o] = s * nm.ysd[o] + nm.ymu[o];
  }
}

// ---------------------------------------------------------------------------
// Forward + backward of L = sum_rows sum_o (out - target)^2 / (2 N) on the standardised
// outputs; block partial gradients gpart[block][kP], block loss lpart[block].
// ---------------------------------------------------------------------------
constexpr int kBR = 32;  // rows per gradient block
constexpr size_t kGradSmem =
    sizeof(float) * (3 * kH * kBR + 2 * kH * kBR + kF * kBR + kO * kBR) + sizeof(uint16_t) * (kP + 8);

// dW[k][c] = sum_r in(k, r) d[c][r] for k < kin, c < 128 (in transposed [k][r], d [c][r])
template <bool SILU_IN>
__device__ void wgrad(const float* in, int kin, const float* d, float* g) {
  const int tid = threadIdx.x;
  for (int e = tid; e < kin * kH; e += kT) {
    const int k = e / kH, c = e % kH;
    float s = 0.f;
#pragma unroll 8
    for (int r = 0; r < kBR; ++r) {
      float v = in[k * kBR + r];
      if (SILU_IN) v = silu(v);
      s = fmaf(v, d[c * kBR + r], s);
    }
    g[e] = s;
  }
  for (int c = tid; c < kH; c += kT) {
    float s = 0.f;
    for (int r = 0; r < kBR; ++r) s += d[c * kBR + r];
    g[kin * kH + c] = s;  // bias follows the weight block in the flat layout
  }
}

// dz_in[k][r] = (sum_c d[c][r] W[k][c]) * silu'(z_in[k][r])
__device__ void back_act(const float* d, const uint16_t* W, const float* zin, float* dout) {
  const int tid = threadIdx.x;
  for (int e = tid; e < kH * kBR; e += kT) {
    const int k = e / kBR, r = e % kBR;
    float s = 0.f;
    for (int c = 0; c < kH; ++c) s = fmaf(d[c * kBR + r], bf2f(W[k * kH + c]), s);
    dout[k * kBR + r] = s * dsilu(zin[k * kBR + r]);
  }
}

__global__ __launch_bounds__(kT) void k_mlp_grad(int64_t N, int64_t Ntot, const float* __restrict__ X,
                                                 const float* __restrict__ Yt, const uint16_t* __restrict__ wbf,
                                                 const float* __restrict__ pf, Norm nm, float* __restrict__ gpart,
                                                 float* __restrict__ lpart) {
  extern __shared__ float sf[];
  float* Z1 = sf;                // [kH][kBR]
  float* Z2 = Z1 + kH * kBR;
  float* Z3 = Z2 + kH * kBR;
  float* D0 = Z3 + kH * kBR;     // [kH][kBR]
  float* D1 = D0 + kH * kBR;
  float* X0 = D1 + kH * kBR;     // [kF][kBR]
  float* DO = X0 + kF * kBR;     // [kO][kBR]
  uint16_t* sW = (uint16_t*)(DO + kO * kBR);
  __shared__ float sb[3 * kH + kO];
  __shared__ float red[kT / 64];
  const int tid = threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.x * kBR;
  float* g = gpart + (int64_t)blockIdx.x * kP;
  load_weights(wbf, sW);
  for (int e = tid; e < 3 * kH + kO; e += kT)
    sb[e] = pf[e < kH ? oB1 + e : e < 2 * kH ? oB2 + e - kH : e < 3 * kH ? oB3 + e - 2 * kH : oB4 + e - 3 * kH];
  for (int e = tid; e < kF * kBR; e += kT) {
    const int r = e / kF, f = e % kF;
    const int64_t row = r0 + r;
    X0[f * kBR + r] = row < N ? (X[row * kF + f] - nm.xmu[f]) / nm.xsd[f] : 0.f;
  }
  __syncthreads();
  layer_fwd<kBR, false>(X0, kF, kBR, sW + oW1, sb, Z1);
  layer_fwd<kBR, true>(Z1, kH, kBR, sW + oW2, sb + kH, Z2);
  layer_fwd<kBR, true>(Z2, kH, kBR, sW + oW3, sb + 2 * kH, Z3);
  float lsum = 0.f;
  if (tid < kBR * kO) {
    const int r = tid >> 1, o = tid & 1;
    float s = sb[3 * kH + o];
    for (int k = 0; k < kH; ++k) s = fmaf(silu(Z3[k * kBR + r]), bf2f(sW[oW4 + k * kO + o]), s);
    const int64_t row = r0 + r;
    float d = 0.f;
    if (row < N) {
      const float tgt = (Yt[row * kO + o] - nm.ymu[o]) / nm.ysd[o];
      d = (s - tgt) / (float)Ntot;
      lsum = 0.5f * (s - tgt) * (s - tgt) / (float)Ntot;
    }
    DO[o * kBR + r] = d;
  }
  for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o, 64);
  if ((tid & 63) == 0) red[tid >> 6] = lsum;
  __syncthreads();
  if (tid == 0) {
    float s = 0.f;
    for (int w = 0; w < kT / 64; ++w) s += red[w];
    lpart[blockIdx.x] = s;
  }
  // output layer: dW4[k][o] = sum_r a3[k][r] DO[o][r], db4[o]
  for (int e = tid; e < kH * kO; e += kT) {
    const int k = e / kO, o = e % kO;
    float s = 0.f;
    for (int r = 0; r < kBR; ++r) s = fmaf(silu(Z3[k * kBR + r]), DO[o * kBR + r], s);
    g[oW4 + e] = s;
  }
  if (tid < kO) {
    float s = 0.f;
    for (int r = 0; r < kBR; ++r) s += DO[tid * kBR + r];
    g[oB4 + tid] = s;
  }
  // dZ3 = (DO W4') * silu'(Z3)
  for (int e = tid; e < kH * kBR; e += kT) {
    const int k = e / kBR, r = e % kBR;
    float s = 0.f;
#pragma unroll
    for (int o = 0; o < kO; ++o) s = fmaf(DO[o * kBR + r], bf2f(sW[oW4 + k * kO + o]), s);
    D0[k * kBR + r] = s * dsilu(Z3[k * kBR + r]);
  }
  __syncthreads();
  wgrad<true>(Z2, kH, D0, g + oW3);       // dW3, db3
  back_act(D0, sW + oW3, Z2, D1);          // dZ2
  __syncthreads();
  wgrad<true>(Z1, kH, D1, g + oW2);       // dW2, db2
  back_act(D1, sW + oW2, Z1, D0);          // dZ1
  __syncthreads();
  wgrad<false>(X0, kF, D0, g + oW1);      // dW1, db1
}

// Per parameter: g = sum of block partials; optional Adam step on the FP32 master and
// its BF16 copy.  Thread 0 also sums the block losses.
__global__ __launch_bounds__(kT) void k_mlp_reduce(int32_t nblk, const float* __restrict__ gpart,
                                                   const float* __restrict__ lpart, float* __restrict__ grad,
                                                   float* __restrict__ loss, float* __restrict__ pf,
                                                   float* __restrict__ m1, float* __restrict__ m2,
                                                   uint16_t* __restrict__ wbf, int32_t step, float lr, float beta1,
                                                   float beta2, float eps) {
  const int p = blockIdx.x * kT + threadIdx.x;
  if (p == 0 && loss != nullptr) {
    float s = 0.f;
    for (int b = 0; b < nblk; ++b) s += lpart[b];
    *loss = s;
  }
  if (p >= kP) return;
  float gs = 0.f;
  for (int b = 0; b < nblk; ++b) gs += gpart[(int64_t)b * kP + p];
  if (grad != nullptr) grad[p] = gs;
  if (step > 0) {
    const float a = beta1 * m1[p] + (1.f - beta1) * gs;
    const float v = beta2 * m2[p] + (1.f - beta2) * gs * gs;
    m1[p] = a;
    m2[p] = v;
    const float ah = a / (1.f - powf(beta1, (float)step));
    const float vh = v / (1.f - powf(beta2, (float)step));
    const float w = pf[p] - lr * ah / (sqrtf(vh) + eps);
    pf[p] = w;
    wbf[p] = f2bf(w);
  }
}

__global__ void k_to_bf16(int32_t n, const float* __restrict__ a, uint16_t* __restrict__ b) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) b[p] = f2bf(a[p]);
}

// Initial site from the predicted marginal (m, log v): the VI site step at that marginal.
__global__ __launch_bounds__(kT) void k_ws_sites(int64_t N, LikParams lik, GHTable gh, const float* __restrict__ y,
                                                 const float* __restrict__ pred, double* __restrict__ nat1,
                                                 double* __restrict__ nat2) {
  const int64_t s = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (s >= N) return;
  const double m = pred[2 * s], v = exp((double)pred[2 * s + 1]);
  SiteResult sr = site_step<M_VI>(lik, gh, (double)y[s], m, v, 0.0, 1e-2, 1.0, true);
  nat1[s] = sr.nat1;
  nat2[s] = sr.nat2;
}

Norm make_norm(const float* n) {
  Norm nm;
  for (int f = 0; f < kF; ++f) { nm.xmu[f] = n[f]; nm.xsd[f] = n[kF + f]; }
  for (int o = 0; o < kO; ++o) { nm.ymu[o] = n[2 * kF + o]; nm.ysd[o] = n[2 * kF + kO + o]; }
  return nm;
}

int set_limits() {
  static int done = 0;
  if (done) return BN_OK;
  WS_HIP_CHECK(hipFuncSetAttribute((const void*)k_mlp_forward, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)kFwdSmem));
  WS_HIP_CHECK(hipFuncSetAttribute((const void*)k_mlp_grad, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)kGradSmem));
  done = 1;
  return BN_OK;
}

}  // namespace

extern "C" {

int bn_ws_features(const bn_st_dims* dims, const float* r, const float* y, const double* t, double ls_t,
                   double ls_s, double var_t, double binsize, float* X, void* stream) {
  if (!dims || !r || !y || !t || !X) return ws_fail(BN_EINVAL, "NULL argument to bn_ws_features");
  if (dims->n_runs < 1 || dims->n_steps < 1 || dims->n_loc < 1) return ws_fail(BN_EINVAL, "empty problem");
  if (!(ls_t > 0 && ls_s > 0 && var_t > 0 && binsize > 0)) return ws_fail(BN_EINVAL, "hypers must be > 0");
  const size_t smem = sizeof(float) * 3 * (2 * kW + 1) * (size_t)dims->n_loc;
  if (smem > 160 * 1024) return ws_fail(BN_EUNSUPPORTED, "too many locations per step for the feature window");
  hipStream_t st = (hipStream_t)stream;
  WS_HIP_CHECK(hipFuncSetAttribute((const void*)k_ws_features, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)smem));
  hipLaunchKernelGGL(k_ws_features, dim3(dims->n_steps, dims->n_runs), dim3(kT), smem, st, dims->n_runs,
                     dims->n_steps, dims->n_loc, r, y, t, (float)ls_t, (float)ls_s, (float)var_t, (float)binsize, X);
  WS_HIP_CHECK(hipGetLastError());
  return BN_OK;
}

int bn_mlp_forward(int64_t n_rows, const float* X, const uint16_t* w_bf16, const float* params, const float* norm,
                   float* out, void* stream) {
  if (n_rows < 1 || !X || !w_bf16 || !params || !norm || !out) return ws_fail(BN_EINVAL, "bad bn_mlp_forward args");
  if (int rc = set_limits()) return rc;
  hipStream_t st = (hipStream_t)stream;
  const int64_t nb = (n_rows + kFR - 1) / kFR;
  if (nb > 0x7fffffff) return ws_fail(BN_EINVAL, "too many rows");
  hipLaunchKernelGGL(k_mlp_forward, dim3((uint32_t)nb), dim3(kT), kFwdSmem, st, n_rows, X, w_bf16, params,
                     make_norm(norm), out);
  WS_HIP_CHECK(hipGetLastError());
  return BN_OK;
}

int64_t bn_mlp_grad_workspace_bytes(int64_t n_rows) {
  if (n_rows < 1) return -1;
  const int64_t nb = (n_rows + kBR - 1) / kBR;
  return nb * (int64_t)(kP + 1) * (int64_t)sizeof(float) + 256;
}

int bn_mlp_grad_step(int64_t n_rows, const float* X, const float* target, float* params, uint16_t* w_bf16,
                     const float* norm, float* adam_m, float* adam_v, int32_t step, float lr, float* grad,
                     float* loss, void* workspace, void* stream) {
  if (n_rows < 1 || !X || !target || !params || !w_bf16 || !norm || !workspace)
    return ws_fail(BN_EINVAL, "bad bn_mlp_grad_step args");
  if (step > 0 && (!adam_m || !adam_v)) return ws_fail(BN_EINVAL, "Adam state missing");
  if (int rc = set_limits()) return rc;
  hipStream_t st = (hipStream_t)stream;
  const int64_t nb = (n_rows + kBR - 1) / kBR;
  if (nb > 0x7fffffff) return ws_fail(BN_EINVAL, "too many rows");
  float* gpart = (float*)workspace;
  float* lpart = gpart + nb * kP;
  hipLaunchKernelGGL(k_mlp_grad, dim3((uint32_t)nb), dim3(kT), kGradSmem, st, n_rows, n_rows, X, target, w_bf16,
                     params, make_norm(norm), gpart, lpart);
  WS_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_mlp_reduce, dim3((kP + kT - 1) / kT), dim3(kT), 0, st, (int32_t)nb, gpart, lpart, grad, loss,
                     params, adam_m, adam_v, w_bf16, step, lr, 0.9f, 0.999f, 1e-8f);
  WS_HIP_CHECK(hipGetLastError());
  return BN_OK;
}

int bn_mlp_to_bf16(const float* params, uint16_t* w_bf16, void* stream) {
  if (!params || !w_bf16) return ws_fail(BN_EINVAL, "NULL argument to bn_mlp_to_bf16");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_to_bf16, dim3((kP + kT - 1) / kT), dim3(kT), 0, st, (int32_t)kP, params, w_bf16);
  WS_HIP_CHECK(hipGetLastError());
  return BN_OK;
}

int bn_ws_sites(int64_t n_sites, const bn_lik_desc* lik, const float* y, const float* pred, double* nat1,
                double* nat2, void* stream) {
  if (n_sites < 1 || !lik || !y || !pred || !nat1 || !nat2) return ws_fail(BN_EINVAL, "bad bn_ws_sites args");
  LikParams L;
  if (int rc = check_lik(lik, L)) return rc;
  GHTable gh;
  if (int rc = make_gh_table(20, gh)) return rc;
  hipStream_t st = (hipStream_t)stream;
  const int64_t nb = (n_sites + kT - 1) / kT;
  hipLaunchKernelGGL(k_ws_sites, dim3((uint32_t)nb), dim3(kT), 0, st, n_sites, L, gh, y, pred, nat1, nat2);
  WS_HIP_CHECK(hipGetLastError());
  return BN_OK;
}

}  // extern "C"
