// Space-time Markov GP (SpatioTemporalKernel, bayesnewton/kernels.py:385-586, with the
// model path of basemodels.py:667-764) on gfx950: K independent runs, each a
// sequence of T steps with R observation locations and M spatial inducing points.
//
// State layout: x = [c = 0 block: f_0..f_{M-1} | c = 1 block: f'_0..f'_{M-1} | ...],
// state index c*M + m, D = KT*M (KT = temporal state dim).  H picks the c = 0 block,
// A = A_t (x) I_M and Q = Q_t (x) I_M act on (c, c') pairs of M x M blocks.
//
// Kernels (all FP64):
//   k_st_conditional  block per (step, run): B_t = Krz Lzz^-T rows, diag C_t     (kernels.py:486-506)
//   k_st_sites        block per (step, run): posterior -> data (basemodels.py:743-764), site
//                     update (inference.py:65-90, 105-195) or energy terms, and the projected
//                     pseudo observation in whitened form: W = B' diag(nat2) B + 1e-12 I = L L',
//                     z = L^-1 B' nat1 (basemodels.py:676-687 without forming W^-1)
//   k_st_filter       block per run, loop over T: predict, measurement update with
//                     U = P-[:, :M] L, Mm = I + L' Sh L (inverted in LDS by the sweep operator),
//                     q = z - L' H m-, m = m- + U Mm^-1 q, P = P- - U Mm^-1 U', log N terms.
//                     Equal to ops.py:159-171 with noise (W + 1e-12 I)^-1.
//   k_st_smoother     block per run, loop back over T: the RTS smoother of ops.py:288-311
//                     (G' = P-^-1 A P with P-^-1 by an in-LDS sweep, P_s = P + G (P_s' - P-) G').
//                     An adjoint (Bryson-Frazier) backward pass avoids the factorisation but
//                     its covariance error grows geometrically over long runs with
//                     precise sites, so the filtered moments are stored instead.
//   k_st_sum          per-run FP64 sums of per-step partials.
//   k_ih_*            the infinite-horizon forms (ops.py:796-1068): per-run Riccati iterations
//                     for a steady-state covariance and gain, means only over the data.
//   k_mf_*            the mean-field forms of the filter and smoother (block-diagonal state
//                     covariance, ops.py:429-706): per-run sequential filter with one M x M
//                     inverse per step, the associative form as per-step M x M work plus a
//                     scalar filter per (run, block) lane, and a per-lane RTS smoother.
// Dense products are workgroup GEMMs on FP64 operands staged through LDS.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "bn_core.h"
#include "bn_internal.h"
#include "bn_sites.h"

using namespace bn;

namespace {

constexpr int kT = 256;      // threads per workgroup
constexpr int kMaxM = 64;    // inducing points per run
constexpr int kMaxKT = 2;    // temporal state dim (Matern12 / Matern32)
constexpr int kMaxD = kMaxKT * kMaxM;  // state dim
static_assert(kT == 2 * kMaxD, "the smoother's mean update splits the workgroup in two halves of kMaxD");
constexpr int kPK = 16;      // GEMM K-panel
constexpr int kPS = 68;      // LDS panel row stride (doubles)
constexpr double kPseudoJitter = 1e-12;  // basemodels.py:683

int st_fail(int code, const std::string& msg) { return set_error(code, msg); }

#define ST_HIP_CHECK(expr)                                                                  \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) return st_fail(BN_EHIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

// element (i, j) of a strided matrix view
struct MatRef {
  const double* p;
  int64_t rs, cs;
  __device__ double at(int i, int j) const { return p[i * rs + j * cs]; }
};

typedef double f64x4 __attribute__((ext_vector_type(4)));

// C = alpha A B + beta Cin (m x n, inner k; Cin may be C, neither may alias A or B) for
// m, n > 64: 128 x 128 output tiles, wave w owns the 64 x 64 quadrant (rows (w/2)*64,
// columns (w%2)*64) as 4 x 4 blocks of v_mfma_f64_16x16x4f64.  K panels of 8 go
// through LDS (row stride 136 doubles: the four k rows a fragment read touches fall on
// disjoint bank halves), the next panel is fetched into registers while the current
// one is used.  `sm` holds 2 * 8 * 136 doubles.
constexpr int kBPK = 8, kBLD = 136;
static_assert(2 * kBPK * kBLD <= 2 * kPK * kPS, "wg_gemm_128 panels fit the wg_gemm panel space");
__device__ __noinline__ void wg_gemm_128(int m, int n, int k, double alpha, MatRef A, MatRef B, double beta, const double* Cin,
                            int64_t ldci, double* C, int64_t ldc, double* sm) {
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = (wave >> 1) * 64, wc = (wave & 1) * 64;
  __syncthreads();  // the caller's writes to the operands are complete
  double* As = sm;                // [kBPK][kBLD]: As[kk][ii] = A(i0+ii, k0+kk)
  double* Bs = sm + kBPK * kBLD;  // [kBPK][kBLD]: Bs[kk][jj] = B(k0+kk, j0+jj)
  constexpr int NL = kBPK * 128 / kT;
  for (int i0 = 0; i0 < m; i0 += 128) {
    for (int j0 = 0; j0 < n; j0 += 128) {
      f64x4 acc[4][4];
#pragma unroll
      for (int bi = 0; bi < 4; ++bi)
#pragma unroll
        for (int bj = 0; bj < 4; ++bj) acc[bi][bj] = f64x4{0.0, 0.0, 0.0, 0.0};
      // per-thread panel elements: A(i0 + ai[q], k0 + ak[q]), B(k0 + bk[q], j0 + bj[q])
      int ai[NL], ak[NL], bk[NL], bjx[NL];
#pragma unroll
      for (int q = 0; q < NL; ++q) {
        const int e = tid + q * kT;
        if (A.cs == 1) { ai[q] = e >> 3; ak[q] = e & 7; } else { ak[q] = e >> 7; ai[q] = e & 127; }
        if (B.cs == 1) { bk[q] = e >> 7; bjx[q] = e & 127; } else { bjx[q] = e >> 3; bk[q] = e & 7; }
      }
      double ra[NL], rb[NL];
      auto fetch = [&](int k0) {
#pragma unroll
        for (int q = 0; q < NL; ++q) {
          const int gi = i0 + ai[q], gk = k0 + ak[q], gj = j0 + bjx[q], gk2 = k0 + bk[q];
          ra[q] = (gi < m && gk < k) ? A.p[gi * A.rs + gk * A.cs] : 0.0;
          rb[q] = (gj < n && gk2 < k) ? B.p[gk2 * B.rs + gj * B.cs] : 0.0;
        }
      };
      fetch(0);
#pragma unroll 1
      for (int k0 = 0; k0 < k; k0 += kBPK) {
#pragma unroll
        for (int q = 0; q < NL; ++q) {
          As[ak[q] * kBLD + ai[q]] = ra[q];
          Bs[bk[q] * kBLD + bjx[q]] = rb[q];
        }
        __syncthreads();
        if (k0 + kBPK < k) fetch(k0 + kBPK);
#pragma unroll
        for (int kk = 0; kk < kBPK; kk += 4) {
          double af[4], bf[4];
#pragma unroll
          for (int bi = 0; bi < 4; ++bi) af[bi] = As[(kk + (lane >> 4)) * kBLD + wr + bi * 16 + (lane & 15)];
#pragma unroll
          for (int bj = 0; bj < 4; ++bj) bf[bj] = Bs[(kk + (lane >> 4)) * kBLD + wc + bj * 16 + (lane & 15)];
#pragma unroll
          for (int bi = 0; bi < 4; ++bi)
#pragma unroll
            for (int bj = 0; bj < 4; ++bj)
              acc[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[bi], bf[bj], acc[bi][bj], 0, 0, 0);
        }
        __syncthreads();
      }
      // block (bi, bj), register r -> row wr + bi*16 + lane/16 + 4r, column wc + bj*16 + lane%16
#pragma unroll
      for (int bi = 0; bi < 4; ++bi)
#pragma unroll
        for (int bj = 0; bj < 4; ++bj)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int gi = i0 + wr + bi * 16 + (lane >> 4) + 4 * r;
            const int gj = j0 + wc + bj * 16 + (lane & 15);
            if (gi < m && gj < n) {
              double v = alpha * acc[bi][bj][r];
              if (beta != 0.0) v = fma(beta, Cin[gi * ldci + gj], v);
              C[gi * ldc + gj] = v;
            }
          }
    }
  }
  __syncthreads();
}

// C[i*ldc + j] = alpha * sum_k A(i,k) B(k,j) + beta * C[i*ldc + j], 0 <= i < m, 0 <= j < n.
// The whole workgroup (kT threads = 4 waves) calls it; every 64x64 output block is
// computed by FP64 matrix-core instructions, one 32x32 quadrant per wave
// (wg_gemm dispatches m, n > 64 to wg_gemm_128).  Panels of A
// and B go through LDS (`sm`: 2 * kPK * kPS doubles); the next panel is fetched from
// memory into registers while the current one is multiplied.
// Operands and C may live in global memory or LDS; C must not alias A or B.
__device__ void wg_gemm_small(int m, int n, int k, double alpha, MatRef A, MatRef B, double beta, double* C,
                              int64_t ldc, double* sm) {
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  __syncthreads();  // the caller's writes to the operands are complete
  double* As = sm;             // [kPK][kPS]: As[kk][ii] = A(i0+ii, k0+kk)
  double* Bs = sm + kPK * kPS;  // [kPK][kPS]: Bs[kk][jj] = B(k0+kk, j0+jj)
  constexpr int NL = kPK * 64 / kT;  // panel elements loaded per thread
  for (int i0 = 0; i0 < m; i0 += 64) {
    for (int j0 = 0; j0 < n; j0 += 64) {
      f64x4 acc[2][2];
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) acc[bi][bj] = f64x4{0.0, 0.0, 0.0, 0.0};
      // the next panel is fetched into registers while the current one is used
      double ra[NL], rb[NL];
      auto fetch = [&](int k0) {
#pragma unroll
        for (int q = 0; q < NL; ++q) {
          const int e = tid + q * kT;
          int ii, kk, jj, kk2;
          if (A.cs == 1) { ii = e >> 4; kk = e & 15; } else { kk = e >> 6; ii = e & 63; }
          if (B.cs == 1) { kk2 = e >> 6; jj = e & 63; } else { jj = e >> 4; kk2 = e & 15; }
          const int gi = i0 + ii, gk = k0 + kk, gj = j0 + jj, gk2 = k0 + kk2;
          ra[q] = (gi < m && gk < k) ? A.at(gi, gk) : 0.0;
          rb[q] = (gj < n && gk2 < k) ? B.at(gk2, gj) : 0.0;
        }
      };
      fetch(0);
      for (int k0 = 0; k0 < k; k0 += kPK) {
#pragma unroll
        for (int q = 0; q < NL; ++q) {
          const int e = tid + q * kT;
          int ii, kk, jj, kk2;
          if (A.cs == 1) { ii = e >> 4; kk = e & 15; } else { kk = e >> 6; ii = e & 63; }
          if (B.cs == 1) { kk2 = e >> 6; jj = e & 63; } else { jj = e >> 4; kk2 = e & 15; }
          As[kk * kPS + ii] = ra[q];
          Bs[kk2 * kPS + jj] = rb[q];
        }
        __syncthreads();
        if (k0 + kPK < k) fetch(k0 + kPK);
        // wave w: 32 x 32 quadrant (rows (w/2)*32, cols (w%2)*32) as 2 x 2 blocks of
        // v_mfma_f64_16x16x4f64: lane l holds A[l%16][k+l/16], B[k+l/16][l%16]
#pragma unroll
        for (int kk = 0; kk < kPK; kk += 4) {
          double af[2], bf[2];
#pragma unroll
          for (int bi = 0; bi < 2; ++bi) af[bi] = As[(kk + (lane >> 4)) * kPS + wr + bi * 16 + (lane & 15)];
#pragma unroll
          for (int bj = 0; bj < 2; ++bj) bf[bj] = Bs[(kk + (lane >> 4)) * kPS + wc + bj * 16 + (lane & 15)];
#pragma unroll
          for (int bi = 0; bi < 2; ++bi)
#pragma unroll
            for (int bj = 0; bj < 2; ++bj) acc[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[bi], bf[bj], acc[bi][bj], 0, 0, 0);
        }
        __syncthreads();
      }
      // output: block (bi, bj), register r -> row wr + bi*16 + lane/16 + 4r, column wc + bj*16 + lane%16
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int gi = i0 + wr + bi * 16 + (lane >> 4) + 4 * r;
            const int gj = j0 + wc + bj * 16 + (lane & 15);
            if (gi < m && gj < n) {
              double v = alpha * acc[bi][bj][r];
              if (beta != 0.0) v = fma(beta, C[gi * ldc + gj], v);
              C[gi * ldc + gj] = v;
            }
          }
    }
  }
  __syncthreads();
}

// Out-of-line entry for the filter and smoother: one call site per product keeps the
// callers' register budgets free of the GEMM's accumulators.
__device__ __noinline__ void wg_gemm(int m, int n, int k, double alpha, MatRef A, MatRef B, double beta, double* C,
                                     int64_t ldc, double* sm) {
  if (m > 64 && n > 64)
    wg_gemm_128(m, n, k, alpha, A, B, beta, C, ldc, C, ldc, sm);
  else
    wg_gemm_small(m, n, k, alpha, A, B, beta, C, ldc, sm);
}

// In-place inverse of an n x n symmetric positive definite matrix a (LDS, stride ld) by
// the sweep operator on every pivot: a := -a^-1.  Returns sum log pivots = log det a
// (thread 0 only, when LOGDET).  Diagonal pivots of an SPD matrix stay positive, so no pivoting;
// a non-positive pivot gives NaN, which propagates.  Thread t holds, in registers, the
// B x B elements at rows t/16 + 16 i and columns t%16 + 16 j (cyclic, n <= 16 B): pivot
// k = 16 ib + rr then sits in register row/column ib for every thread, a compile-time
// index once ib is unrolled, so the pivot row and column fix-ups are predicated selects.
// Per pivot the owners of row k publish it to LDS (`row`: 2 x 16B doubles,
// double-buffered); by symmetry it is also column k.
template <int B, bool LOGDET>
__device__ double wg_sweep_inv(double* a, int n, int ld, double* row) {
  const int tid = threadIdx.x;
  const int rg = tid >> 4, cg = tid & 15;
  double v[B][B];
  __syncthreads();  // the caller's writes to `a` are complete
#pragma unroll
  for (int i = 0; i < B; ++i)
#pragma unroll
    for (int j = 0; j < B; ++j) {
      const int gi = rg + 16 * i, gj = cg + 16 * j;
      v[i][j] = (gi < n && gj < n) ? a[gi * ld + gj] : 0.0;
    }
  double logdet = 0.0;
  __syncthreads();
#pragma unroll
  for (int ib = 0; ib < B; ++ib) {
    const int kend = min(16, n - 16 * ib);
    for (int rr = 0; rr < kend; ++rr) {
      const int k = 16 * ib + rr;
      double* rw = row + (k & 1) * (16 * B);
      if (rg == rr) {
#pragma unroll
        for (int j = 0; j < B; ++j) rw[cg + 16 * j] = v[ib][j];
      }
      __syncthreads();
      const double d = rw[k];
      const double id = 1.0 / d;
      if (LOGDET && tid == 0) logdet += log(d);
      double ri[B], rj[B];
#pragma unroll
      for (int i = 0; i < B; ++i) { ri[i] = rw[rg + 16 * i] * id; rj[i] = rw[cg + 16 * i]; }
      // generic rank-1 update, then the pivot row / column of the owners:
      // a_kj := a_kj / d, a_ik := a_ik / d, a_kk := -1 / d
#pragma unroll
      for (int i = 0; i < B; ++i)
#pragma unroll
        for (int j = 0; j < B; ++j) v[i][j] = fma(-ri[i], rj[j], v[i][j]);
      const bool orow = rg == rr, ocol = cg == rr;
#pragma unroll
      for (int j = 0; j < B; ++j) v[ib][j] = orow ? rj[j] * id : v[ib][j];
#pragma unroll
      for (int i = 0; i < B; ++i) v[i][ib] = ocol ? ((i == ib && orow) ? -id : ri[i]) : v[i][ib];
    }
  }
#pragma unroll
  for (int i = 0; i < B; ++i)
#pragma unroll
    for (int j = 0; j < B; ++j) {
      const int gi = rg + 16 * i, gj = cg + 16 * j;
      if (gi < n && gj < n) a[gi * ld + gj] = v[i][j];
    }
  __syncthreads();
  return logdet;
}

// In-place lower Cholesky of the n x n SPD matrix a (row stride ld, LDS), upper part zeroed.
// A non-positive pivot gives NaN, which propagates (the reference's convention).
__device__ void wg_chol(double* a, int n, int ld) {
  const int tid = threadIdx.x;
  __syncthreads();
  for (int j = 0; j < n; ++j) {
    const double d = sqrt(a[j * ld + j]);
    const double id = 1.0 / d;
    __syncthreads();  // every thread has read the pivot
    for (int i = j + 1 + tid; i < n; i += kT) a[i * ld + j] *= id;
    if (tid == 0) a[j * ld + j] = d;
    __syncthreads();
    const int r = n - j - 1;
    for (int e = tid; e < r * r; e += kT) {
      const int i = j + 1 + e / r, c = j + 1 + e % r;
      if (c <= i) a[i * ld + c] = fma(-a[i * ld + j], a[c * ld + j], a[i * ld + c]);
    }
    __syncthreads();
  }
  for (int e = tid; e < n * n; e += kT) {
    const int i = e / n, c = e % n;
    if (c > i) a[i * ld + c] = 0.0;
  }
  __syncthreads();
}

// Forward substitution L X = B for `nrhs` right-hand-side columns, one thread per
// column (L lower n x n in LDS with stride ld; B(i, c) = b[i*bs + c*bcs], solved in place).
__device__ void wg_lower_solve(const double* L, int n, int ld, double* b, int bs, int bcs, int nrhs) {
  const int tid = threadIdx.x;
  __syncthreads();
  for (int c = tid; c < nrhs; c += kT) {
    double* x = b + c * bcs;
    for (int i = 0; i < n; ++i) {
      double s = x[i * bs];
      for (int q = 0; q < i; ++q) s = fma(-L[i * ld + q], x[q * bs], s);
      x[i * bs] = s / L[i * ld + i];
    }
  }
  __syncthreads();
}

// Block sum of one double per thread; result valid in every thread.
__device__ double wg_sum(double v, double* red) {
  const int tid = threadIdx.x;
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((tid & 63) == 0) red[tid >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < kT / 64; ++w) s += red[w];
  __syncthreads();
  return s;
}

struct StDims {
  int32_t K, T, R, M, KT, D;
};

// Temporal hypers (one set per run, stride 0 shares one set) and spatial variance.
struct StHyp {
  const double* ls_t;
  const double* var_t;
  const double* var_s;
  int32_t stride;
};

template <int KT>
__device__ void temporal_AQ(double ls, double var, double dt, double A[KT * KT], double Q[KT * KT]) {
  Model<KT, 0> mdl;
  mdl.init(ls, var, 1.0, 1.0);
  double Am[KT][KT], Qm[KT][KT];
  mdl.discretise(dt, Am, Qm);
  for (int i = 0; i < KT; ++i)
    for (int j = 0; j < KT; ++j) { A[i * KT + j] = Am[i][j]; Q[i * KT + j] = Qm[i][j]; }
}

// Matern correlation at scaled distance r (kernels.py:130-133, 182-186, 240-245).
__device__ inline double spatial_corr(int32_t kind, double r) {
  if (kind == BN_MATERN12) return exp(-r);
  if (kind == BN_MATERN32) {
    const double a = sqrt(3.0) * r;
    return (1.0 + a) * exp(-a);
  }
  const double a = sqrt(5.0) * r;
  return (1.0 + a + 5.0 / 3.0 * r * r) * exp(-a);
}

// ---------------------------------------------------------------------------
// k_st_conditional: B_t rows (b_i = Lzz^-1 k(z, r_i)) and diag C_t = var_t (var_s - |b_i|^2)
// kernels.py:486-506 (K = Krz Kzz^-1, B = K Lzz = Krz Lzz^-T; C = k_t(0)(Krr - K Krz')).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kT) void k_st_conditional(StDims dm, StHyp hy, const double* __restrict__ Lzz,
                                                       const double* __restrict__ z, int32_t skind, double ls_s,
                                                       const float* __restrict__ r, double* __restrict__ B,
                                                       double* __restrict__ cdiag) {
  __shared__ double sL[kMaxM * kMaxM];
  __shared__ double sz[kMaxM * 2];
  const int t = blockIdx.x, k = blockIdx.y, M = dm.M;
  const int tid = threadIdx.x;
  for (int e = tid; e < M * M; e += kT) sL[e] = Lzz[e];
  for (int e = tid; e < 2 * M; e += kT) sz[e] = z[e];
  __syncthreads();
  const double vt = hy.var_t[k * hy.stride], vs = hy.var_s[k * hy.stride];
  const double il = 1.0 / ls_s;
  const int64_t base = ((int64_t)k * dm.T + t) * dm.R;
  for (int i = tid; i < dm.R; i += kT) {
    const double rx = r[(base + i) * 2], ry = r[(base + i) * 2 + 1];
    double b[kMaxM];
    double nb = 0.0;
    for (int j = 0; j < M; ++j) {
      const double dx = rx - sz[2 * j], dy = ry - sz[2 * j + 1];
      const double r1 = sqrt(fmax((dx * dx + dy * dy) * il * il, 1e-36));   // kernels.py:101-104
      double s = vs * spatial_corr(skind, r1);
      for (int q = 0; q < j; ++q) s = fma(-sL[j * M + q], b[q], s);
      b[j] = s / sL[j * M + j];
      nb = fma(b[j], b[j], nb);
    }
    double* Bi = B + (base + i) * M;
    for (int j = 0; j < M; ++j) Bi[j] = b[j];
    cdiag[base + i] = vt * (vs - nb);
  }
}

// ---------------------------------------------------------------------------
// k_st_sites: per (step, run).
//   mean_f = B m, var_f = diag(B P B') + diag C  (basemodels.py:743-764)
//   UPDATE: site step of METHOD at (mean_f, var_f), damping with lr, then the whitened
//           pseudo observation L L' = B' diag(nat2) B + 1e-12 I, z = L^-1 B' nat1,
//           ldl = sum log diag L.  part[0] = sum |dnat1|, part[1] = sum |dnat2|.
//   PROJECT: only the whitened pseudo observation of the current sites.
//   ENERGY: part[0] = sum of site terms t1 (inference.py:130-222),
//           part[1] = E_q[log N(pseudo_y | f, pseudo_var)] in whitened form
//           (utils.py:510-531): -1/2(|z - L'm|^2 + M log 2pi - 2 ldl) - 1/2 tr(L' P L).
// ---------------------------------------------------------------------------
struct StSiteIO {
  const float* y;         // [K,T,R]
  const double* B;        // [K,T,R,M]
  const double* cdiag;    // [K,T,R]
  const double* pm;       // [K,T,M]
  const double* pP;       // [K,T,M,M]
  double* nat1;           // [K,T,R]
  double* nat2;           // [K,T,R]
  double* Lw;             // [K,T,M,M]
  double* zw;             // [K,T,M]
  double* ldl;            // [K,T]
  double* part0;          // [K,T]
  double* part1;          // [K,T]
  float* o_mean;          // optional [K,T,R] returned site terms
  float* o_jac;
  float* o_hess;
  const float* batch;     // optional [K,T]: nonzero = step in the mini-batch (NULL: every step)
};

// In-place lower Cholesky of an n x n SPD matrix a (LDS, stride ld), the upper part
// zeroed; register blocks as in wg_sweep_inv, one barrier per pivot.  Returns
// sum log diag L (thread 0 only).  A non-positive pivot gives NaN.
template <int B>
__device__ double wg_chol_reg(double* a, int n, int ld, double* row) {
  const int tid = threadIdx.x;
  const int r0 = (tid >> 4) * B, c0 = (tid & 15) * B;
  double v[B][B];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < B; ++i)
#pragma unroll
    for (int j = 0; j < B; ++j) {
      const int gi = r0 + i, gj = c0 + j;
      v[i][j] = (gi < n && gj < n) ? a[gi * ld + gj] : 0.0;
    }
  double ldl = 0.0;
  for (int k = 0; k < n; ++k) {
    double* rw = row + (k & 1) * (16 * B);
    if (k >= r0 && k < r0 + B) {
#pragma unroll
      for (int i = 0; i < B; ++i)
        if (r0 + i == k) {
#pragma unroll
          for (int j = 0; j < B; ++j) rw[c0 + j] = v[i][j];
        }
    }
    __syncthreads();
    const double d = sqrt(rw[k]);
    const double id = 1.0 / d;
    if (tid == 0) ldl += log(d);
    double ri[B], rj[B];
#pragma unroll
    for (int i = 0; i < B; ++i) {
      ri[i] = (r0 + i > k) ? rw[r0 + i] * id : 0.0;
      rj[i] = (c0 + i > k) ? rw[c0 + i] * id : 0.0;
    }
#pragma unroll
    for (int i = 0; i < B; ++i)
#pragma unroll
      for (int j = 0; j < B; ++j) v[i][j] = fma(-ri[i], rj[j], v[i][j]);
    if (k >= c0 && k < c0 + B) {
#pragma unroll
      for (int j = 0; j < B; ++j)
        if (c0 + j == k) {
#pragma unroll
          for (int i = 0; i < B; ++i)
            if (r0 + i >= k) v[i][j] = (r0 + i == k) ? d : ri[i];
        }
    }
  }
#pragma unroll
  for (int i = 0; i < B; ++i)
#pragma unroll
    for (int j = 0; j < B; ++j) {
      const int gi = r0 + i, gj = c0 + j;
      if (gi < n && gj < n) a[gi * ld + gj] = (gj > gi) ? 0.0 : v[i][j];
    }
  __syncthreads();
  return ldl;
}

constexpr int kRC = 32;  // rows of B per chunk
constexpr int kLC = kMaxM + 1;  // LDS row stride of a chunk: consecutive rows start in different banks
static_assert(2 * kRC * kLC >= kMaxM * kMaxM, "the two chunk buffers hold an M x M matrix");

// KIND: the likelihood as a compile-time constant (0: the projection, which has none);
// with only its branch compiled, the site step spills a fraction of what it does with
// every likelihood's code in the kernel.
template <int METHOD, int MODE, int KIND>
__global__ __launch_bounds__(kT, 3) void k_st_sites(StDims dm, LikParams lik, GHTable gh, StSiteIO io, double lr,
                                                 int ensure_psd, double power, int mask_pseudo) {
  extern __shared__ double smem[];
  if (KIND != 0) lik.kind = KIND;
  const int M = dm.M, R = dm.R;
  // LDS is kept near 53 KB (three workgroups per CU): the posterior covariance is read
  // from global memory by the GEMM, the projection W accumulates in its output slot Lw
  // (global, L2-resident) and comes into LDS (the chunk buffers) for the factorisation.
  double* pan = smem;                   // 2*kPK*kPS GEMM panels
  double* sC = pan + 2 * kPK * kPS;     // kRC rows of B (stride LC)
  double* sX = sC + kRC * kLC;          // kRC rows of B P, then nat2-scaled rows of B (stride LC)
  double* sW = sC;                      // M*M after the chunk loop (sC and sX together hold it)
  double* sm = sX + kRC * kLC;          // M posterior mean
  double* sv = sm + kMaxM;              // M: B' nat1 -> z
  double* sn = sv + kMaxM;              // kRC: new nat1 of the chunk
  double* row = sn + kRC;               // 2*kMaxM pivot rows
  double* red = row + 2 * kMaxM;        // 4
  const int t = blockIdx.x, k = blockIdx.y, tid = threadIdx.x;
  const int64_t kt = (int64_t)k * dm.T + t;
  const int LC = M + 1;
  constexpr bool UPDATE = MODE == BN_ST_UPDATE;
  constexpr bool PROJECT = MODE != BN_ST_ENERGY;
  constexpr bool SITES = MODE != BN_ST_PROJECT;
  const double* gP = io.pP + kt * M * M;
  double* gW = io.Lw + kt * M * M;
  // mini-batch (inference.py:65-90, group_natural_params basemodels.py:195-199): a step
  // outside it keeps its sites and adds no data term; its projection is still formed
  const bool in_batch = io.batch == nullptr || io.batch[kt] != 0.0f;
  if (SITES) {
    for (int e = tid; e < M; e += kT) sm[e] = io.pm[kt * M + e];
  }
  for (int e = tid; e < M; e += kT) sv[e] = 0.0;
  double acc0 = 0.0, acc1 = 0.0;
  for (int i0 = 0; i0 < R; i0 += kRC) {
    const int nr = min(kRC, R - i0);
    __syncthreads();
    for (int e = tid; e < kRC * M; e += kT) {
      const int i = e / M, j = e - i * M;
      sC[i * LC + j] = (e < nr * M) ? io.B[(kt * R + i0) * M + e] : 0.0;
    }
    if (SITES) {
      // X = C P ; row i: mean_f = c_i . m, var_f = c_i . x_i + diag C
      wg_gemm_small(kRC, M, M, 1.0, MatRef{sC, LC, 1}, MatRef{gP, M, 1}, 0.0, sX, LC, pan);
      if (tid < nr) {
        const int64_t s = kt * R + i0 + tid;
        double mf = 0.0, vf = io.cdiag[s];
        for (int j = 0; j < M; ++j) {
          const double c = sC[tid * LC + j];
          mf = fma(c, sm[j], mf);
          vf = fma(c, sX[tid * LC + j], vf);
        }
        const double yv = (double)io.y[s];
        if (!in_batch) {
        } else if (UPDATE) {
          const double n1 = io.nat1[s], n2 = io.nat2[s];
          // EP: (pm, pP) is the cavity of the inducing states, so (mf, vf) is the cavity at the data
          SiteResult sr = METHOD == M_EP ? ep_site_from_cavity(lik, gh, yv, mf, vf, power, mask_pseudo != 0,
                                                               ensure_psd != 0)
                                         : site_step<METHOD>(lik, gh, yv, mf, vf, n1, n2, 1.0, ensure_psd != 0);
          acc0 += fabs(sr.nat1 - n1);
          acc1 += fabs(sr.nat2 - n2);
          const double a1 = (1.0 - lr) * n1 + lr * sr.nat1, a2 = (1.0 - lr) * n2 + lr * sr.nat2;
          io.nat1[s] = a1;
          io.nat2[s] = a2;
          if (io.o_mean != nullptr) {
            io.o_mean[s] = (float)sr.mean;
            io.o_jac[s] = (float)sr.jac;
            io.o_hess[s] = (float)sr.hess;
          }
        } else if (METHOD == M_EP) {
          acc0 += ep_lel_from_cavity(lik, gh, yv, mf, vf, power);
        } else {
          double t1, t2;
          energy_terms<METHOD>(lik, gh, yv, mf, vf, 1.0, 1.0, 1.0, t1, t2);
          acc0 += t1;
        }
      }
    }
    if (PROJECT) {
      __syncthreads();
      // X = diag(nat2) C (rows of the chunk), v += C' nat1, W += C' X
      for (int e = tid; e < kRC * M; e += kT) {
        const int i = e / M, j = e - i * M;
        sX[i * LC + j] = (i < nr) ? io.nat2[kt * R + i0 + i] * sC[i * LC + j] : 0.0;
      }
      if (tid < kRC) sn[tid] = (tid < nr) ? io.nat1[kt * R + i0 + tid] : 0.0;
      __syncthreads();
      if (tid < M) {
        double u = sv[tid];
        for (int i = 0; i < nr; ++i) u = fma(sC[i * LC + tid], sn[i], u);
        sv[tid] = u;
      }
      wg_gemm_small(M, M, kRC, 1.0, MatRef{sC, 1, LC}, MatRef{sX, LC, 1}, i0 == 0 ? 0.0 : 1.0, gW, M, pan);
    }
  }
  if (UPDATE) {
    const double d1 = wg_sum(acc0, red), d2 = wg_sum(acc1, red);
    if (tid == 0) { io.part0[kt] = d1; io.part1[kt] = d2; }
  }
  if (PROJECT) {
    __syncthreads();
    for (int e = tid; e < M * M; e += kT) sW[e] = gW[e] + ((e / M == e % M) ? kPseudoJitter : 0.0);
    const double ld = wg_chol_reg<kMaxM / 16>(sW, M, M, row);
    wg_lower_solve(sW, M, M, sv, 1, 0, 1);
    double* Lo = io.Lw + kt * M * M;
    for (int e = tid; e < M * M; e += kT) Lo[e] = sW[e];
    for (int e = tid; e < M; e += kT) io.zw[kt * M + e] = sv[e];
    if (tid == 0) io.ldl[kt] = ld;
  } else if (METHOD == M_EP) {
    // the pseudo term (lel_pseudo) comes from k_st_ep_cavity
    const double e1 = wg_sum(acc0, red);
    if (tid == 0) io.part0[kt] = e1;
  } else {
    const double e1 = wg_sum(acc0, red);
    // whitened expected density of the pseudo observation:
    // |z - L' m|^2 and tr(L' P L) = sum_ac L(a,c) (P L)(a,c), L in LDS (sW), P L in sX
    const double* L = io.Lw + kt * M * M;
    const double* z = io.zw + kt * M;
    double q2 = 0.0, tr = 0.0;
    if (tid < M) {
      double u = z[tid];
      for (int a = tid; a < M; ++a) u = fma(-L[a * M + tid], sm[a], u);   // (z - L' m)_tid
      q2 = u * u;
    }
    wg_gemm_small(M, M, M, 1.0, MatRef{gP, M, 1}, MatRef{L, M, 1}, 0.0, sW, M, pan);  // P L
    for (int e = tid; e < M * M; e += kT) tr = fma(L[e], sW[e], tr);
    const double Q2 = wg_sum(q2, red), TR = wg_sum(tr, red);
    if (tid == 0) {
      io.part0[kt] = e1;
      io.part1[kt] = -0.5 * (Q2 + M * kLog2Pi - 2.0 * io.ldl[kt]) - 0.5 * TR;
    }
  }
}

// ---------------------------------------------------------------------------
// k_st_ep_cavity: per (step, run), the power-EP cavity of the inducing states
// (compute_cavity, utils.py:534-541, on the projected sites of basemodels.py:667-674):
//   Q = P + 1e-8 I,  W = B' diag(nat2) B = L L' - 1e-12 I  (L, z from the projection),
//   cav_cov = (Q^-1 - a W)^-1,  cav_mean = cav_cov (Q^-1 m - a B' nat1),  B' nat1 = L z.
// ENERGY also forms the pseudo term of the EP energy (basemodels.py:247-262),
//   lel_pseudo = log N(py | cav_mean, pv / a + cav_cov) + pep_constant(pv, a),
// with pv = (W + 1e-12 I)^-1 and py = pv B' nat1, in whitened form: w = z - L' cav_mean,
// G = I / a + L' cav_cov L, log N = -1/2 (w' G^-1 w + M log 2pi - 2 log|L| + log|G|),
// pep_constant = M/2 ((1 - a) log 2pi - log a) - (1 - a) log|L|; so pv, whose
// eigenvalues reach 1e12, is never formed.  mask_pseudo with a missing output at a
// one-location step: 0 (mvn_logpdf and pep_constant of a fully masked observation).
// ---------------------------------------------------------------------------
struct StCavityIO {
  const float* y;      // [K,T,R]
  const double* pm;    // [K,T,M]
  const double* pP;    // [K,T,M,M]
  const double* Lw;    // [K,T,M,M]
  const double* zw;    // [K,T,M]
  const double* ldl;   // [K,T]
  double* cm;          // [K,T,M]
  double* cP;          // [K,T,M,M]
  double* lel_pseudo;  // [K,T] (ENERGY)
};

template <bool ENERGY>
__global__ __launch_bounds__(kT) void k_st_ep_cavity(StDims dm, StCavityIO io, double power, int mask_pseudo) {
  extern __shared__ double smem[];
  const int M = dm.M;
  double* pan = smem;                    // 2*kPK*kPS
  double* sQ = pan + 2 * kPK * kPS;      // M*M
  double* sA = sQ + kMaxM * kMaxM;       // M*M
  double* sm = sA + kMaxM * kMaxM;       // M: posterior mean, then cav_mean
  double* sr = sm + kMaxM;               // M: right-hand side, then w
  double* row = sr + kMaxM;              // 2*kMaxM pivot rows
  double* red = row + 2 * kMaxM;         // 4
  const int t = blockIdx.x, k = blockIdx.y, tid = threadIdx.x;
  const int64_t kt = (int64_t)k * dm.T + t;
  const double* P = io.pP + kt * M * M;
  const double* L = io.Lw + kt * M * M;
  const double* z = io.zw + kt * M;
  for (int e = tid; e < M * M; e += kT) sQ[e] = P[e] + ((e / M == e % M) ? 1e-8 : 0.0);
  for (int e = tid; e < M; e += kT) sm[e] = io.pm[kt * M + e];
  wg_sweep_inv<kMaxM / 16, false>(sQ, M, M, row);                          // sQ = -Q^-1
  wg_gemm_small(M, M, M, 1.0, MatRef{L, M, 1}, MatRef{L, 1, M}, 0.0, sA, M, pan);  // L L'
  for (int e = tid; e < M * M; e += kT)
    sA[e] = -sQ[e] - power * sA[e] + ((e / M == e % M) ? power * kPseudoJitter : 0.0);
  if (tid < M) {
    double u = 0.0;
    for (int j = 0; j < M; ++j) u = fma(-sQ[tid * M + j], sm[j], u);
    for (int j = 0; j <= tid; ++j) u = fma(-power * L[tid * M + j], z[j], u);
    sr[tid] = u;
  }
  wg_sweep_inv<kMaxM / 16, false>(sA, M, M, row);                          // sA = -cav_cov
  double cmv = 0.0;
  if (tid < M) {
    for (int j = 0; j < M; ++j) cmv = fma(-sA[tid * M + j], sr[j], cmv);
    io.cm[kt * M + tid] = cmv;
  }
  double* gC = io.cP + kt * M * M;
  for (int e = tid; e < M * M; e += kT) gC[e] = -sA[e];
  if (!ENERGY) return;
  __syncthreads();
  if (tid < M) sm[tid] = cmv;
  __syncthreads();
  if (tid < M) {
    double u = z[tid];
    for (int a = tid; a < M; ++a) u = fma(-L[a * M + tid], sm[a], u);
    sr[tid] = u;                                                             // w = z - L' cav_mean
  }
  wg_gemm_small(M, M, M, -1.0, MatRef{sA, M, 1}, MatRef{L, M, 1}, 0.0, sQ, M, pan);  // cav_cov L
  wg_gemm_small(M, M, M, 1.0, MatRef{L, 1, M}, MatRef{sQ, M, 1}, 0.0, sA, M, pan);   // L' cav_cov L
  for (int e = tid; e < M * M; e += kT)
    if (e / M == e % M) sA[e] += 1.0 / power;
  const double ldg = wg_sweep_inv<kMaxM / 16, true>(sA, M, M, row);       // sA = -G^-1
  double q = 0.0;
  if (tid < M) {
    double u = 0.0;
    for (int j = 0; j < M; ++j) u = fma(-sA[tid * M + j], sr[j], u);
    q = sr[tid] * u;
  }
  const double Q2 = wg_sum(q, red);
  if (tid == 0) {
    const double ldl = io.ldl[kt];
    const double lp = -0.5 * (Q2 + M * kLog2Pi - 2.0 * ldl + ldg);
    const double pep = 0.5 * M * ((1.0 - power) * kLog2Pi - log(power)) - (1.0 - power) * ldl;
    const bool masked = mask_pseudo != 0 && dm.R == 1 && io.y[kt] != io.y[kt];
    io.lel_pseudo[kt] = masked ? 0.0 : lp + pep;
  }
}

// ---------------------------------------------------------------------------
// k_st_marginals: q(f) at every data location, mean_f = B m, var_f = diag(B P B') + diag C
// (conditional_posterior_to_data, basemodels.py:743-764, diagonal only).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kT) void k_st_marginals(StDims dm, const double* __restrict__ B,
                                                     const double* __restrict__ cdiag, const double* __restrict__ pm,
                                                     const double* __restrict__ pP, double* __restrict__ mean_f,
                                                     double* __restrict__ var_f) {
  __shared__ double sP[kMaxM * kMaxM];
  __shared__ double sm[kMaxM];
  const int M = dm.M, R = dm.R;
  const int t = blockIdx.x, k = blockIdx.y, tid = threadIdx.x;
  const int64_t kt = (int64_t)k * dm.T + t;
  for (int e = tid; e < M * M; e += kT) sP[e] = pP[kt * M * M + e];
  for (int e = tid; e < M; e += kT) sm[e] = pm[kt * M + e];
  __syncthreads();
  for (int i = tid; i < R; i += kT) {
    const int64_t s = kt * R + i;
    const double* b = B + s * M;
    double mf = 0.0, vf = cdiag[s];
    for (int j = 0; j < M; ++j) {
      const double bj = b[j];
      mf = fma(bj, sm[j], mf);
      double u = 0.0;
      for (int q = 0; q < M; ++q) u = fma(sP[j * M + q], b[q], u);
      vf = fma(bj, u, vf);
    }
    mean_f[s] = mf;
    var_f[s] = vf;
  }
}

// ---------------------------------------------------------------------------
// k_st_filter: one workgroup per run, sequential over T (ops.py:154-180 in whitened form).
// Per step outputs for the smoother: U0 = P-[:, :M] (D x M), Y (M x M), w (M), H m- (M).
// ---------------------------------------------------------------------------
struct StFilterIO {
  const double* dt;   // [K,T], dt[0] = 0
  const double* Lw;   // [K,T,M,M]
  const double* zw;   // [K,T,M]
  const double* ldl;  // [K,T]
  double* P;          // per-run scratch [K,D,D]
  double* U;          // per-run scratch [K,D,M]
  double* V;          // per-run scratch [K,D,M]
  double* fmean;      // [K,T,D] filtered means
  double* fcov;       // [K,T,D(D+1)/2] filtered covariances, packed upper triangle by rows
  double* ell;        // [K,T] per-step log-likelihood terms
};

// packed upper-triangle index of (i, j), i <= j, in an n x n matrix
__device__ inline int64_t pidx(int i, int j, int n) { return (int64_t)i * n - (int64_t)i * (i - 1) / 2 + (j - i); }

template <int KT>
__global__ __launch_bounds__(kT) void k_st_filter(StDims dm, StHyp hy, StFilterIO io) {
  extern __shared__ double smem[];
  const int M = dm.M, D = dm.D, tid = threadIdx.x;
  const int k = blockIdx.x;
  double* pan = smem;                        // 2*kPK*kPS
  double* sL = pan + 2 * kPK * kPS;          // M*M
  double* sM = sL + kMaxM * kMaxM;           // M*M
  double* row = sM + kMaxM * kMaxM;          // 2 M (pivot rows)
  double* vm = row + 2 * kMaxM;              // D mean
  double* vq = vm + kMaxD;                   // M
  double* red = vq + kMaxM;                  // 4
  __shared__ double sA[KT * KT], sQ[KT * KT];
  double* P = io.P + (int64_t)k * D * D;
  double* U = io.U + (int64_t)k * D * M;
  double* V = io.V + (int64_t)k * D * M;
  const double ls = hy.ls_t[k * hy.stride], var = hy.var_t[k * hy.stride];
  // P0 = Pinf (x) I, m0 = 0
  {
    double Ai[KT * KT], Pinf[KT * KT];
    (void)Ai;
    Model<KT, 0> mdl;
    mdl.init(ls, var, 1.0, 1.0);
    for (int i = 0; i < KT; ++i)
      for (int j = 0; j < KT; ++j) Pinf[i * KT + j] = mdl.Pinf[i][j];
    for (int64_t e = tid; e < (int64_t)D * D; e += kT) {
      const int i = (int)(e / D), j = (int)(e % D);
      P[e] = (i % M == j % M) ? Pinf[(i / M) * KT + (j / M)] : 0.0;
    }
    for (int e = tid; e < D; e += kT) vm[e] = 0.0;
  }
  __syncthreads();
  for (int t = 0; t < dm.T; ++t) {
    const int64_t kt = (int64_t)k * dm.T + t;
    if (tid == 0) temporal_AQ<KT>(ls, var, io.dt[kt], sA, sQ);
    __syncthreads();
    // predict: P- = (A (x) I) P (A (x) I)' + Q (x) I, m- = (A (x) I) m
    for (int e = tid; e < M * M; e += kT) {
      const int a = e / M, c = e % M;
      double blk[KT][KT], out[KT][KT];
      for (int i = 0; i < KT; ++i)
        for (int j = 0; j < KT; ++j) blk[i][j] = P[(int64_t)(i * M + a) * D + j * M + c];
      for (int i = 0; i < KT; ++i)
        for (int j = 0; j < KT; ++j) {
          double s = (a == c) ? sQ[i * KT + j] : 0.0;
          for (int p = 0; p < KT; ++p)
            for (int q = 0; q < KT; ++q) s = fma(sA[i * KT + p] * blk[p][q], sA[j * KT + q], s);
          out[i][j] = s;
        }
      for (int i = 0; i < KT; ++i)
        for (int j = 0; j < KT; ++j) P[(int64_t)(i * M + a) * D + j * M + c] = out[i][j];
    }
    if (tid < M) {
      double x[KT];
      for (int i = 0; i < KT; ++i) x[i] = vm[i * M + tid];
      for (int i = 0; i < KT; ++i) {
        double s = 0.0;
        for (int p = 0; p < KT; ++p) s = fma(sA[i * KT + p], x[p], s);
        vm[i * M + tid] = s;
      }
    }
    const double* L = io.Lw + kt * M * M;
    for (int e = tid; e < M * M; e += kT) sL[e] = L[e];
    __syncthreads();
    // U = P-[:, :M] L ;  Mm = I + L' Sh L = I + L' U[:M] ;  q = z - L' H m-
    wg_gemm(D, M, M, 1.0, MatRef{P, D, 1}, MatRef{sL, M, 1}, 0.0, U, M, pan);
    wg_gemm(M, M, M, 1.0, MatRef{sL, 1, M}, MatRef{U, M, 1}, 0.0, sM, M, pan);
    for (int e = tid; e < M; e += kT) sM[e * M + e] += 1.0;
    if (tid < M) {
      double u = io.zw[kt * M + tid];
      for (int a = tid; a < M; ++a) u = fma(-sL[a * M + tid], vm[a], u);
      vq[tid] = u;
    }
    const double ldM = wg_sweep_inv<kMaxM / 16, true>(sM, M, M, row);  // sM = -Mm^-1 (ldM: thread 0)
    // log N = -1/2 (q' Mm^-1 q + M log 2pi + log det Mm - 2 ldl)
    double qq = 0.0;
    if (tid < M) {
      double u = 0.0;
      for (int c = 0; c < M; ++c) u = fma(sM[tid * M + c], vq[c], u);
      qq = -vq[tid] * u;
    }
    const double quad = wg_sum(qq, red);
    if (tid == 0) io.ell[kt] = -0.5 * (quad + M * kLog2Pi + ldM - 2.0 * io.ldl[kt]);
    // V = U Mm^-1 ; m = m- + V q ; P = P- - V U'
    wg_gemm(D, M, M, -1.0, MatRef{U, M, 1}, MatRef{sM, M, 1}, 0.0, V, M, pan);
    for (int i = tid; i < D; i += kT) {
      double s2 = vm[i];
      for (int c = 0; c < M; ++c) s2 = fma(V[(int64_t)i * M + c], vq[c], s2);
      vm[i] = s2;
    }
    wg_gemm(D, D, M, -1.0, MatRef{V, M, 1}, MatRef{U, 1, M}, 1.0, P, D, pan);
    for (int i = tid; i < D; i += kT) io.fmean[kt * D + i] = vm[i];
    // P - V U' is symmetric only to rounding, and the asymmetry compounds through the next
    // steps' Mm (whose inverse assumes symmetry): P := (P + P')/2, each pair (i < j) by the
    // one lane that also writes its packed entry
    double* fc = io.fcov + kt * ((int64_t)D * (D + 1) / 2);
    for (int i = tid >> 6; i < D; i += kT / 64) {  // one wave per row
      const int64_t o = pidx(i, i, D) - i;
      for (int j = i + (tid & 63); j < D; j += 64) {
        const double v = 0.5 * (P[(int64_t)i * D + j] + P[(int64_t)j * D + i]);
        P[(int64_t)i * D + j] = v;
        P[(int64_t)j * D + i] = v;
        fc[o + j] = v;
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// RTS backward pass (ops.py:288-311) in two kernels.  The gain of step t,
//   G_t' = (P-_{t+1})^-1 A P_t,  P-_{t+1} = A P_t A' + Q  (A, Q of dt[t+1], basemodels.py:700),
// depends on the filtered covariances only, so k_st_gains computes the gains of a block
// of steps for every run at once, one workgroup per (step, run) (sweep inverse of P- in
// LDS, one D^3 product); k_st_smoother then walks the block backwards, one workgroup per
// run, with the two D^3 products left on the recursion:
//   m_s = m + G (m_s' - A m),  P_s = P + G (P_s' - P-) G'.
// Outputs H m_s and H P_s H' (the first M states), and with return_full the state, its
// covariance and the neighbours' cross covariance G_t P_s,t+1.
// Blocks of steps run from the end of the series; a block's gains and unpacked filtered
// covariances sit in a workspace buffer of gain_block_steps() steps per run.
// ---------------------------------------------------------------------------
struct StSmootherIO {
  const double* dt;
  const double* fmean;  // [K,T,D]
  const double* fcov;   // [K,T,D(D+1)/2]
  double* Ps;           // per-run scratch [K,D,D] each: P_s of the step after the current one
  double* Dl;           //   P_s' - P-
  double* GT;           //   G (P_s' - P-)
  double* ms;           // per-run [K,D]: m_s of the step after the current one
  double* buf;          // [K, tb, 2, D, D]: unpacked P_t and G_t' of the block's steps
  int tb;               // steps per block in buf
  double* pm;           // [K,T,M]
  double* pP;           // [K,T,M,M]
  double* smean;        // optional [K,T,D] smoothed state (return_full)
  double* scov;         // optional [K,T,D,D]
  double* scross;       // optional [K,T,D,D]: G_t P_s,t+1 = Cov(x_t, x_t+1 | y); Pinf at the last step
};

// bytes of the gain block buffer: 1953 steps of one d = 128 run, 9 steps each of 200 runs
constexpr int64_t kGainBudget = int64_t(512) << 20;

int gain_block_steps(const StDims& dm) {
  const int64_t per = (int64_t)dm.K * 2 * dm.D * dm.D * (int64_t)sizeof(double);
  int64_t tb = kGainBudget / (per > 0 ? per : 1);
  const int64_t need = dm.T > 1 ? dm.T - 1 : 1;
  if (tb > need) tb = need;
  if (tb < 1) tb = 1;
  return (int)tb;
}

template <int KT>
__global__ __launch_bounds__(kT) void k_st_gains(StDims dm, StHyp hy, StSmootherIO io, int t0) {
  extern __shared__ double smem[];
  const int M = dm.M, D = dm.D, tid = threadIdx.x;
  const int j = blockIdx.x, k = blockIdx.y, t = t0 + j;
  if (t >= dm.T - 1) return;  // uniform over the workgroup
  double* pan = smem;                        // 2*kPK*kPS
  double* sL = pan + 2 * kPK * kPS;          // D*D
  double* row = sL + kMaxD * kMaxD;          // 2 D (pivot rows)
  __shared__ double sA[KT * KT], sQ[KT * KT];
  const int64_t DD = (int64_t)D * D, NSD = (int64_t)D * (D + 1) / 2;
  const int64_t kt = (int64_t)k * dm.T + t;
  double* F = io.buf + ((int64_t)k * io.tb + j) * 2 * DD;
  double* G = F + DD;
  const double ls = hy.ls_t[k * hy.stride], var = hy.var_t[k * hy.stride];
  {
    const double* pk = io.fcov + kt * NSD;
    for (int i = tid >> 6; i < D; i += kT / 64)
      for (int c = tid & 63; c < D; c += 64) F[(int64_t)i * D + c] = pk[i <= c ? pidx(i, c, D) : pidx(c, i, D)];
  }
  if (tid == 0) temporal_AQ<KT>(ls, var, io.dt[kt + 1], sA, sQ);
  __syncthreads();
  // P- = (A (x) I) F (A (x) I)' + Q (x) I, block (a, c) of M x M position pairs at a time
  for (int e = tid; e < M * M; e += kT) {
    const int a = e / M, c = e % M;
    double f[KT][KT], af[KT][KT];
    for (int p = 0; p < KT; ++p)
      for (int q = 0; q < KT; ++q) f[p][q] = F[(int64_t)(p * M + a) * D + q * M + c];
    for (int i = 0; i < KT; ++i)
      for (int q = 0; q < KT; ++q) {
        double v = 0.0;
        for (int p = 0; p < KT; ++p) v = fma(sA[i * KT + p], f[p][q], v);
        af[i][q] = v;
      }
    for (int i = 0; i < KT; ++i)
      for (int jj = 0; jj < KT; ++jj) {
        double v = (a == c) ? sQ[i * KT + jj] : 0.0;
        for (int q = 0; q < KT; ++q) v = fma(af[i][q], sA[jj * KT + q], v);
        sL[(int64_t)(i * M + a) * D + jj * M + c] = v;
      }
  }
  (void)wg_sweep_inv<kMaxD / 16, false>(sL, D, D, row);  // sL = -(P-)^-1
  // X = sL (A (x) I): columns {p M + a} of a row mix by A' (in place, one (row, a) per thread)
  for (int e = tid; e < D * M; e += kT) {
    const int r = e / M, a = e % M;
    double x[KT];
    for (int p = 0; p < KT; ++p) x[p] = sL[(int64_t)r * D + p * M + a];
    for (int i = 0; i < KT; ++i) {
      double v = 0.0;
      for (int p = 0; p < KT; ++p) v = fma(x[p], sA[p * KT + i], v);
      sL[(int64_t)r * D + i * M + a] = v;
    }
  }
  // G' = (P-)^-1 (A (x) I) F = -X F
  wg_gemm(D, D, D, -1.0, MatRef{sL, D, 1}, MatRef{F, D, 1}, 0.0, G, D, pan);
}

template <int KT>
__global__ __launch_bounds__(kT) void k_st_smoother(StDims dm, StHyp hy, StSmootherIO io, int t1, int t0,
                                                    int terminal) {
  extern __shared__ double smem[];
  const int M = dm.M, D = dm.D, tid = threadIdx.x;
  const int k = blockIdx.x;
  double* pan = smem;                        // 2*kPK*kPS
  double* mf = pan + 2 * kPK * kPS;          // D
  double* mp = mf + kMaxD;                   // D
  double* dv = mp + kMaxD;                   // D
  double* ms = dv + kMaxD;                   // D
  __shared__ double sA[KT * KT], sQ[KT * KT];
  const int64_t DD = (int64_t)D * D, NSD = (int64_t)D * (D + 1) / 2;
  double* Ps = io.Ps + k * DD;
  double* Dl = io.Dl + k * DD;
  double* GT = io.GT + k * DD;
  double* msg = io.ms + (int64_t)k * D;
  const double ls = hy.ls_t[k * hy.stride], var = hy.var_t[k * hy.stride];
  auto emit = [&](int64_t kt) {
    for (int e = tid; e < M; e += kT) io.pm[kt * M + e] = ms[e];
    for (int e = tid; e < M * M; e += kT) io.pP[kt * M * M + e] = Ps[(int64_t)(e / M) * D + e % M];
    if (io.smean) {
      for (int e = tid; e < D; e += kT) io.smean[kt * D + e] = ms[e];
      for (int64_t e = tid; e < DD; e += kT) io.scov[kt * DD + e] = Ps[e];
    }
  };
  if (terminal) {
    // the last step: the smoothed state is the filtered one (ops.py:314-315)
    const int64_t kt = (int64_t)k * dm.T + dm.T - 1;
    const double* pk = io.fcov + kt * NSD;
    for (int i = tid >> 6; i < D; i += kT / 64)
      for (int c = tid & 63; c < D; c += 64) Ps[(int64_t)i * D + c] = pk[i <= c ? pidx(i, c, D) : pidx(c, i, D)];
    for (int e = tid; e < D; e += kT) ms[e] = io.fmean[kt * D + e];
    __syncthreads();
    emit(kt);
    if (io.scross) {
      Model<KT, 0> mdl;
      mdl.init(ls, var, 1.0, 1.0);
      for (int64_t e = tid; e < DD; e += kT) {
        const int i = (int)(e / D), j = (int)(e % D);
        io.scross[kt * DD + e] = (i % M == j % M) ? mdl.Pinf[i / M][j / M] : 0.0;
      }
    }
  } else {
    for (int e = tid; e < D; e += kT) ms[e] = msg[e];
  }
  __syncthreads();
  for (int t = t1; t >= t0; --t) {
    const int64_t kt = (int64_t)k * dm.T + t;
    const double* F = io.buf + ((int64_t)k * io.tb + (t - t0)) * 2 * DD;
    const double* G = F + DD;  // G_t' (row-major)
    for (int e = tid; e < D; e += kT) mf[e] = io.fmean[kt * D + e];
    if (tid == 0) temporal_AQ<KT>(ls, var, io.dt[kt + 1], sA, sQ);
    __syncthreads();
    // Dl = sym(P_s') - P-, P- = (A (x) I) F (A (x) I)' + Q (x) I
    for (int e = tid; e < M * M; e += kT) {
      const int a = e / M, c = e % M;
      double f[KT][KT], af[KT][KT];
      for (int p = 0; p < KT; ++p)
        for (int q = 0; q < KT; ++q) f[p][q] = F[(int64_t)(p * M + a) * D + q * M + c];
      for (int i = 0; i < KT; ++i)
        for (int q = 0; q < KT; ++q) {
          double v = 0.0;
          for (int p = 0; p < KT; ++p) v = fma(sA[i * KT + p], f[p][q], v);
          af[i][q] = v;
        }
      for (int i = 0; i < KT; ++i)
        for (int jj = 0; jj < KT; ++jj) {
          double v = (a == c) ? sQ[i * KT + jj] : 0.0;
          for (int q = 0; q < KT; ++q) v = fma(af[i][q], sA[jj * KT + q], v);
          const int64_t o = (int64_t)(i * M + a) * D + jj * M + c;
          const int64_t ot = (int64_t)(jj * M + c) * D + i * M + a;
          Dl[o] = 0.5 * (Ps[o] + Ps[ot]) - v;  // symmetric part of the next step's P_s
        }
    }
    if (tid < M) {
      double x[KT];
      for (int i = 0; i < KT; ++i) x[i] = mf[i * M + tid];
      for (int i = 0; i < KT; ++i) {
        double v = 0.0;
        for (int p = 0; p < KT; ++p) v = fma(sA[i * KT + p], x[p], v);
        mp[i * M + tid] = v;
      }
    }
    // scross = G P_s' ; GT = G (P_s' - P-) ; P_s = P + GT G'
    if (io.scross) wg_gemm(D, D, D, 1.0, MatRef{G, 1, D}, MatRef{Ps, D, 1}, 0.0, io.scross + kt * DD, D, pan);
    wg_gemm(D, D, D, 1.0, MatRef{G, 1, D}, MatRef{Dl, D, 1}, 0.0, GT, D, pan);
    for (int e = tid; e < D; e += kT) dv[e] = ms[e] - mp[e];
    wg_gemm_128(D, D, D, 1.0, MatRef{GT, D, 1}, MatRef{G, D, 1}, 1.0, F, D, Ps, D, pan);
    {  // m_s = m + G dv, the two halves of the workgroup split q; mp is free again
      const int i = tid & (kMaxD - 1), h = tid / kMaxD;
      double v = 0.0;
      if (i < D)
        for (int q = h; q < D; q += 2) v = fma(G[(int64_t)q * D + i], dv[q], v);
      if (h == 1 && i < D) mp[i] = v;
      __syncthreads();
      if (h == 0 && i < D) ms[i] = mf[i] + v + mp[i];
    }
    __syncthreads();
    emit(kt);
    __syncthreads();
  }
  for (int e = tid; e < D; e += kT) msg[e] = ms[e];
}

// per-run sums of per-step partials: out[k] = sum_t a[k*T + t]
__global__ __launch_bounds__(kT) void k_st_sum(int32_t T, const double* __restrict__ a, double* __restrict__ out) {
  __shared__ double red[4];
  const int k = blockIdx.x;
  double s = 0.0;
  for (int t = threadIdx.x; t < T; t += kT) s += a[(int64_t)k * T + t];
  s = wg_sum(s, red);
  if (threadIdx.x == 0) out[k] = s;
}

// ---------------------------------------------------------------------------
// Mean-field space-time filter and smoother (ops.py:429-706, the filter and smoother of
// MarkovMeanFieldGaussianProcess, basemodels.py:1158-1175).  The state is M independent
// temporal blocks (block i = the KT-dim state of latent i, kernels.py:526-575); each update
// keeps only the diagonal blocks of the state covariance.  With P block diagonal and
// H = I_M (x) e_0', the dense update of ops.py:452-461 reduces to
//   S = R + diag(d),  d_i = P-_i[0][0],  v = y - (m-_i[0])_i,  u = S^-1 v,
//   m_i = m-_i + P-_i e_0 u_i,  P_i = P-_i - [S^-1]_ii (P-_i e_0)(P-_i e_0)'
// (block i of K H P- is P-_i e_0 [S^-1]_ii e_0' P-_i: only the diagonal of S^-1 reaches it).
// The associative-scan form (ops.py:468-578) builds every element from S_t = R_t + diag(h Q_t h')
// and projects it to the blocks; per block its elements are those of an ordinary scalar-
// observation filter with noise r_i = 1/[S_t^-1]_ii - h Q_t,i h' and observation
// [S_t^-1 y_t]_i / [S_t^-1]_ii, which k_mf_blocks runs; its log likelihood is then taken
// from the predictions with the dense S (ops.py:572-576).
// Layouts: y [K,T,M], noise R [K,T,M,M] (symmetric), block means [K,T,M,KT], block
// covariances [K,T,M,KT,KT].
// ---------------------------------------------------------------------------

// Pseudo observation of compute_full_pseudo_lik (basemodels.py:676-687) from the whitened
// form: R = (L L')^-1, y = R L z.  Block per (step, run); L L' on the matrix cores, the
// inverse by the sweep operator in LDS.
__global__ __launch_bounds__(kT) void k_mf_pseudo(StDims dm, const double* __restrict__ Lw,
                                                  const double* __restrict__ zw, double* __restrict__ R,
                                                  double* __restrict__ y) {
  extern __shared__ double smem[];
  double* pan = smem;                 // 2*kPK*kPS
  double* sW = pan + 2 * kPK * kPS;   // M*M
  double* row = sW + kMaxM * kMaxM;   // 2*kMaxM
  double* sv = row + 2 * kMaxM;       // M
  const int t = blockIdx.x, k = blockIdx.y, M = dm.M, tid = threadIdx.x;
  const int64_t kt = (int64_t)k * dm.T + t;
  const double* L = Lw + kt * M * M;  // lower triangular, upper part zero
  wg_gemm_small(M, M, M, 1.0, MatRef{L, M, 1}, MatRef{L, 1, M}, 0.0, sW, M, pan);
  if (tid < M) {
    double s = 0.0;
    for (int q = 0; q <= tid; ++q) s = fma(L[tid * M + q], zw[kt * M + q], s);
    sv[tid] = s;
  }
  wg_sweep_inv<kMaxM / 16, false>(sW, M, M, row);  // sW = -(L L')^-1
  double* Rk = R + kt * M * M;
  for (int e = tid; e < M * M; e += kT) Rk[e] = -sW[e];
  if (tid < M) {
    double s = 0.0;
    for (int j = 0; j < M; ++j) s = fma(-sW[tid * M + j], sv[j], s);
    y[kt * M + tid] = s;
  }
}

struct MfFilterIO {
  const double* dt;   // [K,T] (dt[k,0] = 0 for the first step of a run)
  const double* y;    // [K,T,M]
  const double* R;    // [K,T,M,M]
  double* fmean;      // [K,T,M,KT]
  double* fcov;       // [K,T,M,KT,KT]
  double* ell;        // [K,T]
  double* reff;       // [K,T,M] (associative form)
  double* yeff;       // [K,T,M]
};

template <int KT>
__device__ inline void mf_predict(const double A[KT * KT], const double Q[KT * KT], const double m[KT],
                                  const double P[KT][KT], double mp[KT], double Pp[KT][KT]) {
  double AP[KT][KT];
  for (int i = 0; i < KT; ++i) {
    double s = 0.0;
    for (int q = 0; q < KT; ++q) s = fma(A[i * KT + q], m[q], s);
    mp[i] = s;
    for (int j = 0; j < KT; ++j) {
      double a = 0.0;
      for (int q = 0; q < KT; ++q) a = fma(A[i * KT + q], P[q][j], a);
      AP[i][j] = a;
    }
  }
  for (int i = 0; i < KT; ++i)
    for (int j = 0; j < KT; ++j) {
      double s = Q[i * KT + j];
      for (int q = 0; q < KT; ++q) s = fma(AP[i][q], A[j * KT + q], s);
      Pp[i][j] = s;
    }
  for (int i = 0; i < KT; ++i)
    for (int j = i + 1; j < KT; ++j) Pp[i][j] = Pp[j][i] = 0.5 * (Pp[i][j] + Pp[j][i]);
}

template <int KT>
__device__ inline void mf_pinf(double ls, double var, double P[KT][KT]) {
  Model<KT, 0> mdl;
  mdl.init(ls, var, 1.0, 1.0);
  for (int i = 0; i < KT; ++i)
    for (int j = 0; j < KT; ++j) P[i][j] = mdl.Pinf[i][j];
}

// Sequential form (ops.py:429-465): block per run, loop over T.  Thread i < M owns block i
// in registers; S = R_t + diag(d) is inverted in LDS by the sweep operator (log det from
// its pivots).
template <int KT>
__global__ __launch_bounds__(kT) void k_mf_filter(StDims dm, StHyp hy, MfFilterIO io) {
  __shared__ double sS[kMaxM * kMaxM];
  __shared__ double row[2 * kMaxM];
  __shared__ double sv[kMaxM];
  __shared__ double red[4];
  const int k = blockIdx.x, M = dm.M, tid = threadIdx.x;
  const double ls = hy.ls_t[k * hy.stride], var = hy.var_t[k * hy.stride];
  double m[KT], P[KT][KT], mp[KT], Pp[KT][KT];
  mf_pinf<KT>(ls, var, P);
  for (int i = 0; i < KT; ++i) m[i] = 0.0;
  for (int t = 0; t < dm.T; ++t) {
    const int64_t kt = (int64_t)k * dm.T + t;
    const double* Rt = io.R + kt * M * M;
    for (int e = tid; e < M * M; e += kT) sS[e] = Rt[e];
    if (tid < M) {
      double A[KT * KT], Q[KT * KT];
      temporal_AQ<KT>(ls, var, io.dt[kt], A, Q);
      mf_predict<KT>(A, Q, m, P, mp, Pp);
      sv[tid] = io.y[kt * M + tid] - mp[0];
    }
    __syncthreads();
    if (tid < M) sS[tid * M + tid] += Pp[0][0];
    const double ld = wg_sweep_inv<kMaxM / 16, true>(sS, M, M, row);  // sS = -S^-1
    double q = 0.0, u = 0.0, s = 0.0;
    if (tid < M) {
      for (int j = 0; j < M; ++j) u = fma(-sS[tid * M + j], sv[j], u);
      s = -sS[tid * M + tid];
      q = sv[tid] * u;
    }
    const double quad = wg_sum(q, red);
    if (tid == 0) io.ell[kt] = -0.5 * (quad + M * kLog2Pi + ld);
    if (tid < M) {
      double* fm = io.fmean + (kt * M + tid) * KT;
      double* fc = io.fcov + (kt * M + tid) * KT * KT;
      for (int i = 0; i < KT; ++i) {
        m[i] = fma(Pp[i][0], u, mp[i]);
        fm[i] = m[i];
        for (int j = 0; j < KT; ++j) {
          P[i][j] = fma(-s * Pp[i][0], Pp[j][0], Pp[i][j]);
          fc[i * KT + j] = P[i][j];
        }
      }
    }
  }
}

// Per-step dense part of the associative form, block per (step, run).
//   ELEMENTS: d = h Q_t h' (Q_0 = Pinf), reff_i = 1/[S^-1]_ii - d_i, yeff_i = [S^-1 y]_i / [S^-1]_ii
//   LOGLIK:   d, v from the prediction of the filtered blocks at t-1 (Pinf, 0 at t = 0),
//             ell_t = log N(v | 0, S)
template <int KT, bool LOGLIK>
__global__ __launch_bounds__(kT) void k_mf_sdiag(StDims dm, StHyp hy, MfFilterIO io) {
  __shared__ double sS[kMaxM * kMaxM];
  __shared__ double row[2 * kMaxM];
  __shared__ double sv[kMaxM];
  __shared__ double red[4];
  const int t = blockIdx.x, k = blockIdx.y, M = dm.M, tid = threadIdx.x;
  const int64_t kt = (int64_t)k * dm.T + t;
  const double ls = hy.ls_t[k * hy.stride], var = hy.var_t[k * hy.stride];
  const double* Rt = io.R + kt * M * M;
  for (int e = tid; e < M * M; e += kT) sS[e] = Rt[e];
  double d = 0.0;
  if (tid < M) {
    double P[KT][KT];
    mf_pinf<KT>(ls, var, P);
    double A[KT * KT], Q[KT * KT];
    temporal_AQ<KT>(ls, var, io.dt[kt], A, Q);
    if (!LOGLIK) {
      d = t == 0 ? P[0][0] : Q[0];
      sv[tid] = io.y[kt * M + tid];
    } else {
      double m[KT], mp[KT], Pp[KT][KT];
      for (int i = 0; i < KT; ++i) m[i] = 0.0;
      if (t > 0) {
        const double* fm = io.fmean + ((kt - 1) * M + tid) * KT;
        const double* fc = io.fcov + ((kt - 1) * M + tid) * KT * KT;
        for (int i = 0; i < KT; ++i) {
          m[i] = fm[i];
          for (int j = 0; j < KT; ++j) P[i][j] = fc[i * KT + j];
        }
      }
      mf_predict<KT>(A, Q, m, P, mp, Pp);
      d = Pp[0][0];
      sv[tid] = io.y[kt * M + tid] - mp[0];
    }
  }
  __syncthreads();
  if (tid < M) sS[tid * M + tid] += d;
  const double ld = wg_sweep_inv<kMaxM / 16, LOGLIK>(sS, M, M, row);  // sS = -S^-1
  double q = 0.0, u = 0.0;
  if (tid < M) {
    for (int j = 0; j < M; ++j) u = fma(-sS[tid * M + j], sv[j], u);
    q = sv[tid] * u;
  }
  if (LOGLIK) {
    const double quad = wg_sum(q, red);
    if (tid == 0) io.ell[kt] = -0.5 * (quad + M * kLog2Pi + ld);
  } else if (tid < M) {
    const double s = -sS[tid * M + tid];
    io.reff[kt * M + tid] = 1.0 / s - d;
    io.yeff[kt * M + tid] = u / s;
  }
}

// Scalar-observation filter of every block over the per-step (reff, yeff), lane per
// (run, block).
template <int KT>
__global__ __launch_bounds__(64) void k_mf_blocks(StDims dm, StHyp hy, MfFilterIO io) {
  const int64_t g = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (g >= (int64_t)dm.K * dm.M) return;
  const int k = (int)(g / dm.M), i = (int)(g % dm.M), M = dm.M;
  const double ls = hy.ls_t[k * hy.stride], var = hy.var_t[k * hy.stride];
  double m[KT], P[KT][KT], mp[KT], Pp[KT][KT];
  mf_pinf<KT>(ls, var, P);
  for (int a = 0; a < KT; ++a) m[a] = 0.0;
  for (int t = 0; t < dm.T; ++t) {
    const int64_t kt = (int64_t)k * dm.T + t;
    double A[KT * KT], Q[KT * KT];
    temporal_AQ<KT>(ls, var, io.dt[kt], A, Q);
    mf_predict<KT>(A, Q, m, P, mp, Pp);
    const double S = Pp[0][0] + io.reff[kt * M + i];
    const double v = io.yeff[kt * M + i] - mp[0];
    double* fm = io.fmean + (kt * M + i) * KT;
    double* fc = io.fcov + (kt * M + i) * KT * KT;
    for (int a = 0; a < KT; ++a) {
      const double Ka = Pp[a][0] / S;
      m[a] = fma(Ka, v, mp[a]);
      fm[a] = m[a];
      for (int b = 0; b < KT; ++b) {
        P[a][b] = fma(-Ka, Pp[b][0], Pp[a][b]);
        fc[a * KT + b] = P[a][b];
      }
    }
  }
}

struct MfSmootherIO {
  const double* dt;     // [K,T] step vector of the smoother (already the next step's dt)
  const double* fmean;  // [K,T,M,KT]
  const double* fcov;   // [K,T,M,KT,KT]
  double* hmean;        // optional [K,T,M]: h m_s
  double* hvar;         // optional: h P_s h' at hvar[kt*var_step + i*var_elem]
  int64_t var_step, var_elem;
  double* smean;        // optional [K,T,M,KT]
  double* scov;         // optional [K,T,M,KT,KT]
  double* gain;         // optional [K,T,M,KT,KT]
};

// RTS smoother of every block (ops.py:611-641; the associative form 644-678 gives the
// same result), lane per (run, block), from the last step back.
template <int KT>
__global__ __launch_bounds__(64) void k_mf_smoother(StDims dm, StHyp hy, MfSmootherIO io) {
  const int64_t g = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (g >= (int64_t)dm.K * dm.M) return;
  const int k = (int)(g / dm.M), i = (int)(g % dm.M), M = dm.M;
  const double ls = hy.ls_t[k * hy.stride], var = hy.var_t[k * hy.stride];
  double sm[KT], sP[KT][KT];
  {
    const int64_t kt = (int64_t)k * dm.T + dm.T - 1;
    for (int a = 0; a < KT; ++a) {
      sm[a] = io.fmean[(kt * M + i) * KT + a];
      for (int b = 0; b < KT; ++b) sP[a][b] = io.fcov[((kt * M + i) * KT + a) * KT + b];
    }
  }
  for (int t = dm.T - 1; t >= 0; --t) {
    const int64_t kt = (int64_t)k * dm.T + t;
    double fm[KT], fP[KT][KT], pm[KT], pP[KT][KT], A[KT * KT], Q[KT * KT];
    for (int a = 0; a < KT; ++a) {
      fm[a] = io.fmean[(kt * M + i) * KT + a];
      for (int b = 0; b < KT; ++b) fP[a][b] = io.fcov[((kt * M + i) * KT + a) * KT + b];
    }
    temporal_AQ<KT>(ls, var, io.dt[kt], A, Q);
    mf_predict<KT>(A, Q, fm, fP, pm, pP);
    // G = (A fP)' pP^-1 = fP A' pP^-1
    double AfP[KT][KT], iP[KT][KT], G[KT][KT];
    for (int a = 0; a < KT; ++a)
      for (int b = 0; b < KT; ++b) {
        double s = 0.0;
        for (int q = 0; q < KT; ++q) s = fma(A[a * KT + q], fP[q][b], s);
        AfP[a][b] = s;
      }
    if (KT == 1) {
      iP[0][0] = 1.0 / pP[0][0];
    } else {
      const double det = pP[0][0] * pP[KT - 1][KT - 1] - pP[0][KT - 1] * pP[KT - 1][0];
      iP[0][0] = pP[KT - 1][KT - 1] / det;
      iP[KT - 1][KT - 1] = pP[0][0] / det;
      iP[0][KT - 1] = -pP[0][KT - 1] / det;
      iP[KT - 1][0] = -pP[KT - 1][0] / det;
    }
    for (int a = 0; a < KT; ++a)
      for (int b = 0; b < KT; ++b) {
        double s = 0.0;
        for (int q = 0; q < KT; ++q) s = fma(AfP[q][a], iP[q][b], s);
        G[a][b] = s;
      }
    double dm_[KT], dP[KT][KT], GdP[KT][KT];
    for (int a = 0; a < KT; ++a) {
      dm_[a] = sm[a] - pm[a];
      for (int b = 0; b < KT; ++b) dP[a][b] = sP[a][b] - pP[a][b];
    }
    for (int a = 0; a < KT; ++a) {
      double s = fm[a];
      for (int q = 0; q < KT; ++q) s = fma(G[a][q], dm_[q], s);
      sm[a] = s;
      for (int b = 0; b < KT; ++b) {
        double w = 0.0;
        for (int q = 0; q < KT; ++q) w = fma(G[a][q], dP[q][b], w);
        GdP[a][b] = w;
      }
    }
    for (int a = 0; a < KT; ++a)
      for (int b = 0; b < KT; ++b) {
        double s = fP[a][b];
        for (int q = 0; q < KT; ++q) s = fma(GdP[a][q], G[b][q], s);
        sP[a][b] = s;
      }
    for (int a = 0; a < KT; ++a)
      for (int b = a + 1; b < KT; ++b) sP[a][b] = sP[b][a] = 0.5 * (sP[a][b] + sP[b][a]);
    if (io.hmean) io.hmean[kt * M + i] = sm[0];
    if (io.hvar) io.hvar[kt * io.var_step + i * io.var_elem] = sP[0][0];
    if (io.smean)
      for (int a = 0; a < KT; ++a) io.smean[(kt * M + i) * KT + a] = sm[a];
    if (io.scov)
      for (int a = 0; a < KT; ++a)
        for (int b = 0; b < KT; ++b) io.scov[((kt * M + i) * KT + a) * KT + b] = sP[a][b];
    if (io.gain)
      for (int a = 0; a < KT; ++a)
        for (int b = 0; b < KT; ++b) io.gain[((kt * M + i) * KT + a) * KT + b] = G[a][b];
  }
}

// ---------------------------------------------------------------------------
// Infinite-horizon space-time filter and smoother (ops.py:796-1068, the filter and
// smoother of InfiniteHorizonGaussianProcess, basemodels.py:1257-1301).  Equispaced steps:
// one A, Q per run (its step size).  Per run:
//   k_ih_tied_part, k_ih_tied
//                 R_tied = (mean_t B' diag(nat2_t) B)^-1 from the whitened form (L L' - 1e-12 I),
//                 partial sums over step ranges on the matrix cores
//   k_ih_dare     `iters` Riccati iterations X <- A (X - X H' S^-1 H X) A' + Q, S = H X H' + R_tied,
//                 from the previous solution; S_tied^-1 and cov = X - K H X (K = X H' S^-1)
//   k_ih_sinv     heteroscedastic: S_t = H X H' + R_t inverted per (step, run), log det S_t
//   k_ih_filter   m_t = A m_{t-1} + X H' S_t^-1 (y_t - H A m_{t-1}), log N(y_t | H A m_{t-1}, S_t)
//   k_ih_smoother G = cov A' X^-1, Qd = cov - G X G', `iters` iterations X_s <- G X_s G' + Qd
//                 (from the previous solution), sm_t = fm_t + G (sm_{t+1} - A fm_t) from
//                 sm = fm_{T-1}; H X_s H' at every step.
// State layout as in k_st_filter: index c*M + m, H picks the c = 0 block.
// ---------------------------------------------------------------------------
constexpr int kTiedParts = 32;  // step ranges summed in parallel per run

// part[k, p] = sum of L_t L_t' over the p-th of kTiedParts step ranges of run k.
__global__ __launch_bounds__(kT) void k_ih_tied_part(StDims dm, const double* __restrict__ Lw,
                                                     double* __restrict__ part) {
  extern __shared__ double smem[];
  double* pan = smem;                 // 2*kPK*kPS
  double* acc = pan + 2 * kPK * kPS;  // M*M
  const int p = blockIdx.x, k = blockIdx.y, M = dm.M, tid = threadIdx.x;
  const int t0 = (int)((int64_t)dm.T * p / kTiedParts), t1 = (int)((int64_t)dm.T * (p + 1) / kTiedParts);
  for (int e = tid; e < M * M; e += kT) acc[e] = 0.0;
  for (int t = t0; t < t1; ++t) {
    const double* L = Lw + ((int64_t)k * dm.T + t) * M * M;
    wg_gemm_small(M, M, M, 1.0, MatRef{L, M, 1}, MatRef{L, 1, M}, 1.0, acc, M, pan);
  }
  __syncthreads();
  double* out = part + ((int64_t)k * kTiedParts + p) * M * M;
  for (int e = tid; e < M * M; e += kT) out[e] = acc[e];
}

// R_tied = (mean_t L_t L_t' - 1e-12 I)^-1 = (mean_t B' diag(nat2_t) B)^-1, the partial
// sums added in a fixed order.
__global__ __launch_bounds__(kT) void k_ih_tied(StDims dm, const double* __restrict__ part, double* __restrict__ Rt) {
  __shared__ double sL[kMaxM * kMaxM];
  __shared__ double row[2 * kMaxM];
  const int k = blockIdx.x, M = dm.M, tid = threadIdx.x;
  const double it = 1.0 / dm.T;
  for (int e = tid; e < M * M; e += kT) {
    double s = 0.0;
    for (int p = 0; p < kTiedParts; ++p) s += part[((int64_t)k * kTiedParts + p) * M * M + e];
    sL[e] = s * it - ((e / M == e % M) ? kPseudoJitter : 0.0);
  }
  wg_sweep_inv<kMaxM / 16, false>(sL, M, M, row);  // -(tied precision)^-1
  for (int e = tid; e < M * M; e += kT) Rt[(int64_t)k * M * M + e] = -sL[e];
}

struct IhIO {
  const double* step;  // [K] step size of each run
  const double* y;     // [K,T,M]
  const double* R;     // [K,T,M,M] (heteroscedastic)
  const double* Rt;    // [K,M,M] tied noise
  double* X;           // [K,D,D] Riccati solution: start in, result out
  double* cov;         // [K,D,D]
  double* W1;          // [K,D,D] scratch
  double* W2;          // [K,D,D] scratch
  double* STn;         // [K,M,M] -S_tied^-1
  double* ldT;         // [K]
  double* Sn;          // [K,T,M,M] -S_t^-1 (heteroscedastic)
  double* ldS;         // [K,T]
  double* fmean;       // [K,T,D]
  double* ell;         // [K,T]
  int32_t iters, hetero;
};

template <int KT>
__global__ __launch_bounds__(kT) void k_ih_dare(StDims dm, StHyp hy, IhIO io) {
  extern __shared__ double smem[];
  double* pan = smem;                   // 2*kPK*kPS
  double* sS = pan + 2 * kPK * kPS;     // M*M
  double* sR = sS + kMaxM * kMaxM;      // M*M
  double* row = sR + kMaxM * kMaxM;     // 2*kMaxM
  __shared__ double sA[KT * KT], sQ[KT * KT];
  const int k = blockIdx.x, M = dm.M, D = dm.D, tid = threadIdx.x;
  const int64_t DD = (int64_t)D * D;
  double* X = io.X + k * DD;
  double* W1 = io.W1 + k * DD;
  double* W2 = io.W2 + k * DD;
  const double ls = hy.ls_t[k * hy.stride], var = hy.var_t[k * hy.stride];
  if (tid == 0) temporal_AQ<KT>(ls, var, io.step[k], sA, sQ);
  for (int e = tid; e < M * M; e += kT) sR[e] = io.Rt[(int64_t)k * M * M + e];
  __syncthreads();
  for (int it = 0; it < io.iters; ++it) {
    for (int e = tid; e < M * M; e += kT) sS[e] = X[(int64_t)(e / M) * D + e % M] + sR[e];
    wg_sweep_inv<kMaxM / 16, false>(sS, M, M, row);  // sS = -S^-1
    wg_gemm(D, M, M, -1.0, MatRef{X, D, 1}, MatRef{sS, M, 1}, 0.0, W1, M, pan);  // K = X H' S^-1
    for (int64_t e = tid; e < DD; e += kT) W2[e] = X[e];
    __syncthreads();
    wg_gemm(D, D, M, -1.0, MatRef{W1, M, 1}, MatRef{X, D, 1}, 1.0, W2, D, pan);  // X - K H X
    __syncthreads();
    // X = A sym(W2) A' + Q over the KT x KT blocks of M x M
    for (int e = tid; e < M * M; e += kT) {
      const int a = e / M, c = e % M;
      double w[KT][KT];
      for (int p = 0; p < KT; ++p)
        for (int q = 0; q < KT; ++q)
          w[p][q] = 0.5 * (W2[(int64_t)(p * M + a) * D + q * M + c] + W2[(int64_t)(q * M + c) * D + p * M + a]);
      for (int i = 0; i < KT; ++i)
        for (int j = 0; j < KT; ++j) {
          double v = (a == c) ? sQ[i * KT + j] : 0.0;
          for (int p = 0; p < KT; ++p)
            for (int q = 0; q < KT; ++q) v = fma(sA[i * KT + p] * w[p][q], sA[j * KT + q], v);
          X[(int64_t)(i * M + a) * D + j * M + c] = v;
        }
    }
    __syncthreads();
  }
  // tied innovation covariance, gain and the infinite-horizon filtered covariance
  for (int e = tid; e < M * M; e += kT) sS[e] = X[(int64_t)(e / M) * D + e % M] + sR[e];
  const double ld = wg_sweep_inv<kMaxM / 16, true>(sS, M, M, row);
  for (int e = tid; e < M * M; e += kT) io.STn[(int64_t)k * M * M + e] = sS[e];
  if (tid == 0) io.ldT[k] = ld;
  wg_gemm(D, M, M, -1.0, MatRef{X, D, 1}, MatRef{sS, M, 1}, 0.0, W1, M, pan);
  double* cv = io.cov + k * DD;
  for (int64_t e = tid; e < DD; e += kT) cv[e] = X[e];
  __syncthreads();
  wg_gemm(D, D, M, -1.0, MatRef{W1, M, 1}, MatRef{X, D, 1}, 1.0, cv, D, pan);
}

__global__ __launch_bounds__(kT) void k_ih_sinv(StDims dm, IhIO io) {
  __shared__ double sS[kMaxM * kMaxM];
  __shared__ double row[2 * kMaxM];
  const int t = blockIdx.x, k = blockIdx.y, M = dm.M, D = dm.D, tid = threadIdx.x;
  const int64_t kt = (int64_t)k * dm.T + t;
  const double* X = io.X + (int64_t)k * D * D;
  const double* Rt = io.R + kt * M * M;
  for (int e = tid; e < M * M; e += kT) sS[e] = Rt[e] + X[(int64_t)(e / M) * D + e % M];
  const double ld = wg_sweep_inv<kMaxM / 16, true>(sS, M, M, row);
  for (int e = tid; e < M * M; e += kT) io.Sn[kt * M * M + e] = sS[e];
  if (tid == 0) io.ldS[kt] = ld;
}

template <int KT>
__global__ __launch_bounds__(kT) void k_ih_filter(StDims dm, StHyp hy, IhIO io) {
  __shared__ double sm[kMaxD], mp[kMaxD], sv[kMaxM], su[kMaxM], red[4];
  __shared__ double sA[KT * KT], sQ[KT * KT];
  const int k = blockIdx.x, M = dm.M, D = dm.D, tid = threadIdx.x;
  const double* X = io.X + (int64_t)k * D * D;
  const double ls = hy.ls_t[k * hy.stride], var = hy.var_t[k * hy.stride];
  if (tid == 0) temporal_AQ<KT>(ls, var, io.step[k], sA, sQ);
  for (int e = tid; e < D; e += kT) sm[e] = 0.0;
  __syncthreads();
  for (int t = 0; t < dm.T; ++t) {
    const int64_t kt = (int64_t)k * dm.T + t;
    if (tid < D) {
      const int c = tid / M, a = tid % M;
      double v = 0.0;
      for (int p = 0; p < KT; ++p) v = fma(sA[c * KT + p], sm[p * M + a], v);
      mp[tid] = v;
    }
    __syncthreads();
    if (tid < M) sv[tid] = io.y[kt * M + tid] - mp[tid];
    __syncthreads();
    const double* Sn = io.hetero ? io.Sn + kt * M * M : io.STn + (int64_t)k * M * M;  // -S^-1
    double q = 0.0;
    if (tid < M) {
      double u = 0.0;
      for (int j = 0; j < M; ++j) u = fma(-Sn[j * M + tid], sv[j], u);
      su[tid] = u;
      q = sv[tid] * u;
    }
    const double quad = wg_sum(q, red);
    if (tid == 0) io.ell[kt] = -0.5 * (quad + M * kLog2Pi + (io.hetero ? io.ldS[kt] : io.ldT[k]));
    if (tid < D) {
      // X H' u: column i of X H' is row i of the (symmetric) X
      double v = mp[tid];
      for (int i = 0; i < M; ++i) v = fma(X[(int64_t)i * D + tid], su[i], v);
      sm[tid] = v;
      io.fmean[kt * D + tid] = v;
    }
    __syncthreads();
  }
}

struct IhSmIO {
  const double* step;   // [K]
  const double* fmean;  // [K,T,D]
  const double* X;      // [K,D,D] filter Riccati solution
  const double* cov;    // [K,D,D]
  double* Xs;           // [K,D,D] smoother Riccati solution: start in, result out
  double* G;            // [K,D,D] gain
  double* W1;           // [K,D,D] scratch
  double* W2;           // [K,D,D] scratch
  double* pm;           // optional [K,T,M]
  double* pP;           // optional [K,T,M,M]
  double* smean;        // optional [K,T,D]
  int32_t iters;
};

template <int KT>
__global__ __launch_bounds__(kT) void k_ih_smoother(StDims dm, StHyp hy, IhSmIO io) {
  extern __shared__ double smem[];
  const int k = blockIdx.x, M = dm.M, D = dm.D, tid = threadIdx.x;
  double* pan = smem;                    // 2*kPK*kPS
  double* sL = pan + 2 * kPK * kPS;      // D*D
  double* row = sL + kMaxD * kMaxD;      // 2 D
  double* ms = row + 2 * kMaxD;          // D
  double* mf = ms + kMaxD;               // D
  double* dv = mf + kMaxD;               // D
  double* tmp = dv + kMaxD;              // D
  __shared__ double sA[KT * KT], sQ[KT * KT];
  const int64_t DD = (int64_t)D * D;
  const double* X = io.X + k * DD;
  const double* cov = io.cov + k * DD;
  double* Xs = io.Xs + k * DD;
  double* G = io.G + k * DD;
  double* W1 = io.W1 + k * DD;
  double* W2 = io.W2 + k * DD;
  const double ls = hy.ls_t[k * hy.stride], var = hy.var_t[k * hy.stride];
  if (tid == 0) temporal_AQ<KT>(ls, var, io.step[k], sA, sQ);
  for (int64_t e = tid; e < DD; e += kT) sL[e] = X[e];
  (void)wg_sweep_inv<kMaxD / 16, false>(sL, D, D, row);  // sL = -X^-1
  // W1 = A' X^-1 ; G = cov W1
  for (int64_t e = tid; e < DD; e += kT) {
    const int r = (int)(e / D), j = (int)(e % D), c = r / M, m = r % M;
    double v = 0.0;
    for (int a = 0; a < KT; ++a) v = fma(-sA[a * KT + c], sL[(int64_t)(a * M + m) * D + j], v);
    W1[e] = v;
  }
  wg_gemm(D, D, D, 1.0, MatRef{cov, D, 1}, MatRef{W1, D, 1}, 0.0, G, D, pan);
  // Qd = cov - G X G' -> W2
  wg_gemm(D, D, D, 1.0, MatRef{G, D, 1}, MatRef{X, D, 1}, 0.0, W1, D, pan);
  for (int64_t e = tid; e < DD; e += kT) W2[e] = cov[e];
  wg_gemm(D, D, D, -1.0, MatRef{W1, D, 1}, MatRef{G, 1, D}, 1.0, W2, D, pan);
  for (int it = 0; it < io.iters; ++it) {
    wg_gemm(D, D, D, 1.0, MatRef{G, D, 1}, MatRef{Xs, D, 1}, 0.0, W1, D, pan);
    for (int64_t e = tid; e < DD; e += kT) Xs[e] = W2[e];
    wg_gemm(D, D, D, 1.0, MatRef{W1, D, 1}, MatRef{G, 1, D}, 1.0, Xs, D, pan);
  }
  // G' into LDS for the mean recursion
  for (int64_t e = tid; e < DD; e += kT) sL[(e % D) * D + e / D] = G[e];
  {
    const int64_t kt = (int64_t)k * dm.T + dm.T - 1;
    for (int e = tid; e < D; e += kT) ms[e] = io.fmean[kt * D + e];
  }
  __syncthreads();
  for (int t = dm.T - 1; t >= 0; --t) {
    const int64_t kt = (int64_t)k * dm.T + t;
    for (int e = tid; e < D; e += kT) mf[e] = io.fmean[kt * D + e];
    __syncthreads();
    if (tid < D) {
      const int c = tid / M, a = tid % M;
      double v = 0.0;
      for (int p = 0; p < KT; ++p) v = fma(sA[c * KT + p], mf[p * M + a], v);
      dv[tid] = ms[tid] - v;
    }
    __syncthreads();
    if (tid < D) {
      double v = mf[tid];
      for (int j = 0; j < D; ++j) v = fma(sL[(int64_t)j * D + tid], dv[j], v);
      tmp[tid] = v;
    }
    __syncthreads();
    if (tid < D) {
      ms[tid] = tmp[tid];
      if (io.smean) io.smean[kt * D + tid] = tmp[tid];
      if (io.pm && tid < M) io.pm[kt * M + tid] = tmp[tid];
    }
    if (io.pP)
      for (int e = tid; e < M * M; e += kT) io.pP[kt * M * M + e] = Xs[(int64_t)(e / M) * D + e % M];
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// k_st_predict: MarkovGaussianProcess.predict on the space-time branch
// (basemodels.py:766-816 with temporal_conditional, utils.py:125-141, 185-215), block per
// (test time, run).  The prior is Kronecker in (temporal state) x (latent), so the
// conditional of x(t*) on the two neighbouring smoothed states is P_t (x) I_M, T_t (x) I_M
// with the KT-dim statistics P_t [KT x 2KT], T_t of the temporal kernel (the reference's
// 1e-8 jitter is per diagonal entry either way).  Only H x(t*) is needed:
//   H m* = sum_u P_t[0][u] m_j[u],  S = sum_uv P_t[0][u] P_t[0][v] Sig_j[u][v] + T_t[0][0] I
// over the 2KT M x M blocks of the joint (m_j, Sig_j) of the neighbours (cross block
// G_t P_s,t+1 from the smoother); then f(t*, r) has mean b_r' H m* and variance
// b_r' S b_r + c_r (the diagonal of W P W' + C).  Sentinels at -+1e10 carry (0, Pinf).
// ---------------------------------------------------------------------------
struct StPredIO {
  const double* x;       // [K,T] training times
  const double* xs;      // [K,NS] test times
  const double* B;       // [K,NS,RS,M] spatial conditional rows of the test locations
  const double* cdiag;   // [K,NS,RS]
  const double* smean;   // [K,T,D]
  const double* scov;    // [K,T,D,D]
  const double* scross;  // [K,T,D,D]
  const double* steady;  // [K,D,D] or nullptr: the infinite-horizon state covariance at every test time
  double* mean;          // [K,NS,RS]
  double* var;           // [K,NS,RS]
  int32_t ns, rs;
};

template <int KT>
__global__ __launch_bounds__(kT) void k_st_predict(StDims dm, StHyp hy, StPredIO io) {
  __shared__ double sS[kMaxM * kMaxM];
  __shared__ double hm[kMaxM];
  __shared__ double pc[2 * KT + 1];  // P_t[0][u], then T_t[0][0]
  __shared__ int sidx;
  const int i = blockIdx.x, k = blockIdx.y, M = dm.M, D = dm.D, T = dm.T, tid = threadIdx.x;
  const int64_t DD = (int64_t)D * D;
  const double ls = hy.ls_t[k * hy.stride], var = hy.var_t[k * hy.stride];
  const double* xk = io.x + (int64_t)k * T;
  const double ts = io.xs[(int64_t)k * io.ns + i];
  Model<KT, 0> mdl;
  mdl.init(ls, var, 1.0, 1.0);
  if (tid == 0) {
    // augmented grid [-1e10, x_0 .. x_T-1, 1e10]; ind = searchsorted(left) - 1
    int lo = 0, hi = T;  // first training index with x >= ts
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (xk[mid] < ts) lo = mid + 1;
      else hi = mid;
    }
    const int ind = lo;  // interval (aug[ind], aug[ind + 1]]
    sidx = ind;
    const double xl = ind == 0 ? -1e10 : xk[ind - 1];
    const double xr = ind == T ? 1e10 : xk[ind];
    double Af[KT][KT], Qf[KT][KT], Ab[KT][KT], Qb[KT][KT];
    mdl.discretise(ts - xl, Af, Qf);
    mdl.discretise(xr - ts, Ab, Qb);
    // Q-+ = Qb + Ab Qf Ab' + 1e-8 I ; V = Q-+^-1 Ab ; T = Qf - (Ab Qf)' V Qf ; W = Qf V'
    double AQ[KT][KT], Qm[KT][KT], V[KT][KT], W[KT][KT], Pt[KT][2 * KT];
    for (int a = 0; a < KT; ++a)
      for (int b = 0; b < KT; ++b) {
        double s = 0.0;
        for (int q = 0; q < KT; ++q) s = fma(Ab[a][q], Qf[q][b], s);
        AQ[a][b] = s;
      }
    for (int a = 0; a < KT; ++a)
      for (int b = 0; b < KT; ++b) {
        double s = Qb[a][b] + (a == b ? 1e-8 : 0.0);
        for (int q = 0; q < KT; ++q) s = fma(AQ[a][q], Ab[b][q], s);
        Qm[a][b] = s;
      }
    double iQ[KT][KT];
    if (KT == 1) {
      iQ[0][0] = 1.0 / Qm[0][0];
    } else {
      const double det = Qm[0][0] * Qm[KT - 1][KT - 1] - Qm[0][KT - 1] * Qm[KT - 1][0];
      iQ[0][0] = Qm[KT - 1][KT - 1] / det;
      iQ[KT - 1][KT - 1] = Qm[0][0] / det;
      iQ[0][KT - 1] = -Qm[0][KT - 1] / det;
      iQ[KT - 1][0] = -Qm[KT - 1][0] / det;
    }
    for (int a = 0; a < KT; ++a)
      for (int b = 0; b < KT; ++b) {
        double s = 0.0;
        for (int q = 0; q < KT; ++q) s = fma(iQ[a][q], Ab[q][b], s);
        V[a][b] = s;
      }
    for (int a = 0; a < KT; ++a)
      for (int b = 0; b < KT; ++b) {
        double s = 0.0;
        for (int q = 0; q < KT; ++q) s = fma(Qf[a][q], V[b][q], s);
        W[a][b] = s;
      }
    double WAb[KT][KT];
    for (int a = 0; a < KT; ++a)
      for (int b = 0; b < KT; ++b) {
        double s = 0.0;
        for (int q = 0; q < KT; ++q) s = fma(W[a][q], Ab[q][b], s);
        WAb[a][b] = s;
      }
    for (int a = 0; a < KT; ++a)
      for (int b = 0; b < KT; ++b) {
        double s = Af[a][b];
        for (int q = 0; q < KT; ++q) s = fma(-WAb[a][q], Af[q][b], s);
        Pt[a][b] = s;
        Pt[a][KT + b] = W[a][b];
      }
    // T[0][0] = Qf[0][0] - (Ab Qf)[:,0]' V Qf[:,0]
    double VQ[KT];
    for (int a = 0; a < KT; ++a) {
      double s = 0.0;
      for (int q = 0; q < KT; ++q) s = fma(V[a][q], Qf[q][0], s);
      VQ[a] = s;
    }
    double t00 = Qf[0][0];
    for (int a = 0; a < KT; ++a) t00 = fma(-AQ[a][0], VQ[a], t00);
    for (int u = 0; u < 2 * KT; ++u) pc[u] = Pt[0][u];
    pc[2 * KT] = t00;
  }
  __syncthreads();
  const int ind = sidx;
  const bool lsent = ind == 0, rsent = ind == T;
  const int64_t tl = (int64_t)k * T + (lsent ? 0 : ind - 1), tr = (int64_t)k * T + (rsent ? T - 1 : ind);
  const double* mL = io.smean + tl * D;
  const double* mR = io.smean + tr * D;
  const double* PL = io.steady ? nullptr : io.scov + tl * DD;
  const double* PR = io.steady ? nullptr : io.scov + tr * DD;
  const double* Cx = io.steady ? nullptr : io.scross + tl * DD;  // Cov(left, right); 0 after the left sentinel
  for (int a = tid; a < M; a += kT) {
    double s = 0.0;
    for (int c = 0; c < KT; ++c) {
      if (!lsent) s = fma(pc[c], mL[c * M + a], s);
      if (!rsent) s = fma(pc[KT + c], mR[c * M + a], s);
    }
    hm[a] = s;
  }
  if (io.steady) {
    // temporal_conditional_infinite_horizon (utils.py:163-177): the smoother's steady
    // covariance replaces the conditional one
    for (int e = tid; e < M * M; e += kT) sS[e] = io.steady[k * DD + (int64_t)(e / M) * D + e % M];
  } else
  for (int e = tid; e < M * M; e += kT) {
    const int a = e / M, b = e % M;
    double s = (a == b) ? pc[2 * KT] : 0.0;
    for (int c = 0; c < KT; ++c)
      for (int c2 = 0; c2 < KT; ++c2) {
        const int64_t ia = (int64_t)(c * M + a) * D, ib = (int64_t)(c2 * M + b) * D;
        const double wl = pc[c], wr = pc[KT + c], vl = pc[c2], vr = pc[KT + c2];
        const double pl = lsent ? (a == b ? mdl.Pinf[c][c2] : 0.0) : PL[ia + c2 * M + b];
        const double pr = rsent ? (a == b ? mdl.Pinf[c][c2] : 0.0) : PR[ia + c2 * M + b];
        const double xlr = lsent ? 0.0 : Cx[ia + c2 * M + b];   // Cov(left_a, right_b)
        const double xrl = lsent ? 0.0 : Cx[ib + c * M + a];    // Cov(right_a, left_b)
        s = fma(wl * vl, pl, s);
        s = fma(wr * vr, pr, s);
        s = fma(wl * vr, xlr, s);
        s = fma(wr * vl, xrl, s);
      }
    sS[e] = s;
  }
  __syncthreads();
  const int64_t base = ((int64_t)k * io.ns + i) * io.rs;
  for (int r = tid; r < io.rs; r += kT) {
    const double* b = io.B + (base + r) * M;
    double mu = 0.0, v = io.cdiag[base + r];
    for (int a = 0; a < M; ++a) {
      const double ba = b[a];
      mu = fma(ba, hm[a], mu);
      double u = 0.0;
      for (int c = 0; c < M; ++c) u = fma(sS[a * M + c], b[c], u);
      v = fma(ba, u, v);
    }
    io.mean[base + r] = mu;
    io.var[base + r] = v;
  }
}

constexpr size_t kFilterSmem = sizeof(double) * (2 * kPK * kPS + 2 * kMaxM * kMaxM + 2 * kMaxM + kMaxD + kMaxM + 4);
constexpr size_t kSmootherSmem = sizeof(double) * (2 * kPK * kPS + kMaxD * kMaxD + 6 * kMaxD);
constexpr size_t kGainsSmem = sizeof(double) * (2 * kPK * kPS + kMaxD * kMaxD + 2 * kMaxD);
constexpr size_t kSmootherSeqSmem = sizeof(double) * (2 * kPK * kPS + 4 * kMaxD);
constexpr size_t kPseudoSmem = sizeof(double) * (2 * kPK * kPS + kMaxM * kMaxM + 3 * kMaxM);
constexpr size_t kTiedSmem = sizeof(double) * (2 * kPK * kPS + kMaxM * kMaxM);
constexpr size_t kIhDareSmem = sizeof(double) * (2 * kPK * kPS + 2 * kMaxM * kMaxM + 2 * kMaxM);
constexpr size_t kSitesSmem =
    sizeof(double) * (2 * kPK * kPS + 2 * kRC * kLC + 2 * kMaxM + kRC + 2 * kMaxM + 4);
constexpr size_t kCavitySmem = sizeof(double) * (2 * kPK * kPS + 2 * kMaxM * kMaxM + 4 * kMaxM + 4);

int check_dims(const bn_st_dims* d, StDims& dm) {
  if (d == nullptr) return st_fail(BN_EINVAL, "dims is NULL");
  if (d->n_runs < 1 || d->n_steps < 1 || d->n_loc < 1) return st_fail(BN_EINVAL, "empty space-time problem");
  if (d->n_inducing < 1 || d->n_inducing > kMaxM)
    return st_fail(BN_EUNSUPPORTED, "number of spatial inducing points must be in [1, 64]");
  if (d->temporal_kind != BN_MATERN12 && d->temporal_kind != BN_MATERN32)
    return st_fail(BN_EUNSUPPORTED, "space-time temporal kernel must be Matern12 or Matern32");
  dm.K = d->n_runs;
  dm.T = d->n_steps;
  dm.R = d->n_loc;
  dm.M = d->n_inducing;
  dm.KT = d->temporal_kind;  // state dimension of the temporal component
  dm.D = dm.KT * dm.M;
  return BN_OK;
}

StHyp make_hyp(const bn_st_hypers* h) { return StHyp{h->ls_t, h->var_t, h->var_s, h->stride}; }

// Dynamic LDS above the default limit is opted into once per kernel.
int set_smem_limits() {
  static int done = 0;
  if (done) return BN_OK;
  ST_HIP_CHECK(hipFuncSetAttribute((const void*)k_st_filter<1>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)kFilterSmem));
  ST_HIP_CHECK(hipFuncSetAttribute((const void*)k_st_filter<2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)kFilterSmem));
  ST_HIP_CHECK(hipFuncSetAttribute((const void*)k_st_gains<1>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)kGainsSmem));
  ST_HIP_CHECK(hipFuncSetAttribute((const void*)k_st_gains<2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)kGainsSmem));
#define ST_SITES_ATTR1(MT, MD, KD)                                                                             \
  ST_HIP_CHECK(hipFuncSetAttribute((const void*)k_st_sites<MT, MD, KD>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                                   (int)kSitesSmem))
#define ST_SITES_ATTR(MT, MD)                 \
  ST_SITES_ATTR1(MT, MD, LIK_GAUSSIAN);        \
  ST_SITES_ATTR1(MT, MD, LIK_BERNOULLI_PROBIT); \
  ST_SITES_ATTR1(MT, MD, LIK_POISSON);         \
  ST_SITES_ATTR1(MT, MD, LIK_STUDENTS_T)
  ST_HIP_CHECK(hipFuncSetAttribute((const void*)k_mf_pseudo, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)kPseudoSmem));
  ST_HIP_CHECK(hipFuncSetAttribute((const void*)k_ih_dare<1>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)kIhDareSmem));
  ST_HIP_CHECK(hipFuncSetAttribute((const void*)k_ih_dare<2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)kIhDareSmem));
  ST_HIP_CHECK(hipFuncSetAttribute((const void*)k_ih_smoother<1>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)kSmootherSmem));
  ST_HIP_CHECK(hipFuncSetAttribute((const void*)k_ih_smoother<2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)kSmootherSmem));
  ST_SITES_ATTR(M_VI, BN_ST_UPDATE);
  ST_SITES_ATTR(M_VI, BN_ST_ENERGY);
  ST_SITES_ATTR1(M_VI, BN_ST_PROJECT, 0);
  ST_SITES_ATTR(M_NEWTON, BN_ST_UPDATE);
  ST_SITES_ATTR(M_NEWTON, BN_ST_ENERGY);
  ST_SITES_ATTR(M_EP, BN_ST_UPDATE);
  ST_SITES_ATTR(M_EP, BN_ST_ENERGY);
#undef ST_SITES_ATTR
#undef ST_SITES_ATTR1
  ST_HIP_CHECK(hipFuncSetAttribute((const void*)k_st_ep_cavity<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)kCavitySmem));
  ST_HIP_CHECK(hipFuncSetAttribute((const void*)k_st_ep_cavity<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)kCavitySmem));
  done = 1;
  return BN_OK;
}

// k_st_sites<MT, MD> for the likelihood kind L.kind (checked by check_lik).
template <int MT, int MD>
void launch_st_sites(dim3 grid, hipStream_t st, StDims dm, LikParams L, GHTable gh, StSiteIO io, double lr,
                     int ensure_psd, double power, int mask_pseudo) {
  switch (L.kind) {
    case LIK_GAUSSIAN:
      hipLaunchKernelGGL((k_st_sites<MT, MD, LIK_GAUSSIAN>), grid, dim3(kT), kSitesSmem, st, dm, L, gh, io, lr,
                         ensure_psd, power, mask_pseudo);
      break;
    case LIK_BERNOULLI_PROBIT:
      hipLaunchKernelGGL((k_st_sites<MT, MD, LIK_BERNOULLI_PROBIT>), grid, dim3(kT), kSitesSmem, st, dm, L, gh, io,
                         lr, ensure_psd, power, mask_pseudo);
      break;
    case LIK_POISSON:
      hipLaunchKernelGGL((k_st_sites<MT, MD, LIK_POISSON>), grid, dim3(kT), kSitesSmem, st, dm, L, gh, io, lr,
                         ensure_psd, power, mask_pseudo);
      break;
    default:
      hipLaunchKernelGGL((k_st_sites<MT, MD, LIK_STUDENTS_T>), grid, dim3(kT), kSitesSmem, st, dm, L, gh, io, lr,
                         ensure_psd, power, mask_pseudo);
      break;
  }
}

}  // namespace

extern "C" {

int bn_st_conditional(const bn_st_dims* dims, const bn_st_hypers* hyp, int32_t spatial_kind, double ls_s, const double* z,
                      const double* Lzz, const float* r, double* B, double* cdiag, void* stream) {
  StDims dm;
  if (int rc = check_dims(dims, dm)) return rc;
  if (!hyp || !z || !Lzz || !r || !B || !cdiag) return st_fail(BN_EINVAL, "NULL argument to bn_st_conditional");
  if (!(ls_s > 0)) return st_fail(BN_EINVAL, "spatial lengthscale must be > 0");
  if (spatial_kind < BN_MATERN12 || spatial_kind > BN_MATERN52)
    return st_fail(BN_EUNSUPPORTED, "spatial kernel must be Matern12, Matern32 or Matern52");
  hipStream_t st = (hipStream_t)stream;
  BN_TIMED(KID_ST_CONDITIONAL, st);
  hipLaunchKernelGGL(k_st_conditional, dim3(dm.T, dm.K), dim3(kT), 0, st, dm, make_hyp(hyp), Lzz, z, spatial_kind, ls_s, r, B,
                     cdiag);
  ST_HIP_CHECK(hipGetLastError());
  return BN_OK;
}

int bn_st_sites(const bn_st_dims* dims, const bn_lik_desc* lik, int32_t method, int32_t mode, int32_t gh_points,
                double lr, int32_t ensure_psd, const float* y, const double* B, const double* cdiag,
                const double* post_mean, const double* post_cov, const float* batch, double* nat1, double* nat2,
                double* Lw, double* zw, double* ldl, double* part0, double* part1, float* out_mean, float* out_jac,
                float* out_hess, void* stream) {
  StDims dm;
  if (int rc = check_dims(dims, dm)) return rc;
  if (!lik || !y || !B || !cdiag || !post_mean || !post_cov || !nat1 || !nat2 || !Lw || !zw || !ldl || !part0 ||
      !part1)
    return st_fail(BN_EINVAL, "NULL argument to bn_st_sites");
  if (method != BN_VI && method != BN_NEWTON)
    return st_fail(BN_EUNSUPPORTED, "space-time sites support BN_VI and BN_NEWTON");
  if (mode != BN_ST_UPDATE && mode != BN_ST_ENERGY && mode != BN_ST_PROJECT)
    return st_fail(BN_EINVAL, "mode must be BN_ST_UPDATE, BN_ST_ENERGY or BN_ST_PROJECT");
  LikParams L;
  if (int rc = check_lik(lik, L)) return rc;
  GHTable gh;
  if (int rc = make_gh_table(gh_points, gh)) return rc;
  if (int rc = set_smem_limits()) return rc;
  StSiteIO io{y,    B,    cdiag, post_mean, post_cov, nat1,    nat2,     Lw,
              zw,   ldl,  part0, part1,     out_mean, out_jac, out_hess, batch};
  hipStream_t st = (hipStream_t)stream;
  const dim3 grid(dm.T, dm.K);
  BN_TIMED(KID_ST_SITES, st);
  if (mode == BN_ST_PROJECT) {
    // the projection is method- and likelihood-free
    hipLaunchKernelGGL((k_st_sites<M_VI, BN_ST_PROJECT, 0>), grid, dim3(kT), kSitesSmem, st, dm, L, gh, io, lr,
                       ensure_psd, 1.0, 0);
  } else if (method == BN_VI) {
    if (mode == BN_ST_UPDATE) launch_st_sites<M_VI, BN_ST_UPDATE>(grid, st, dm, L, gh, io, lr, ensure_psd, 1.0, 0);
    else launch_st_sites<M_VI, BN_ST_ENERGY>(grid, st, dm, L, gh, io, lr, ensure_psd, 1.0, 0);
  } else {
    if (mode == BN_ST_UPDATE) launch_st_sites<M_NEWTON, BN_ST_UPDATE>(grid, st, dm, L, gh, io, lr, ensure_psd, 1.0, 0);
    else launch_st_sites<M_NEWTON, BN_ST_ENERGY>(grid, st, dm, L, gh, io, lr, ensure_psd, 1.0, 0);
  }
  ST_HIP_CHECK(hipGetLastError());
  return BN_OK;
}

int bn_st_ep_sites(const bn_st_dims* dims, const bn_lik_desc* lik, int32_t mode, int32_t gh_points, double lr,
                   int32_t ensure_psd, double power, const float* y, const double* B, const double* cdiag,
                   const double* post_mean, const double* post_cov, double* nat1, double* nat2, double* Lw,
                   double* zw, double* ldl, double* cav_mean, double* cav_cov, double* part0, double* part1,
                   float* out_mean, float* out_jac, float* out_hess, void* stream) {
  StDims dm;
  if (int rc = check_dims(dims, dm)) return rc;
  if (!lik || !y || !B || !cdiag || !post_mean || !post_cov || !nat1 || !nat2 || !Lw || !zw || !ldl || !cav_mean ||
      !cav_cov || !part0 || !part1)
    return st_fail(BN_EINVAL, "NULL argument to bn_st_ep_sites");
  if (mode != BN_ST_UPDATE && mode != BN_ST_ENERGY)
    return st_fail(BN_EINVAL, "mode must be BN_ST_UPDATE or BN_ST_ENERGY (the projection is method-free: bn_st_sites)");
  if (!(power > 0.0 && power <= 1.0)) return st_fail(BN_EINVAL, "EP power must be in (0, 1]");
  LikParams L;
  if (int rc = check_lik(lik, L)) return rc;
  if (dm.R > 1 && L.kind != LIK_GAUSSIAN)
    return st_fail(BN_EUNSUPPORTED,
                   "EP with several locations per step moment-matches all of them jointly; only the Gaussian "
                   "likelihood has that in closed form (one location per step takes any likelihood)");
  if (dm.R > 1 && mode == BN_ST_ENERGY)
    return st_fail(BN_EUNSUPPORTED, "masking in spatio-temporal models not implemented for EP");
  GHTable gh;
  if (int rc = make_gh_table(gh_points, gh)) return rc;
  if (int rc = set_smem_limits()) return rc;
  // basemodels.py:138-144, 653: the pseudo observations are masked only when their
  // dimension (M) equals the data's (R)
  const int mask_pseudo = dm.M == dm.R ? 1 : 0;
  hipStream_t st = (hipStream_t)stream;
  const dim3 grid(dm.T, dm.K);
  StCavityIO cio{y, post_mean, post_cov, Lw, zw, ldl, cav_mean, cav_cov, part1};
  StSiteIO io{y,  B,   cdiag, cav_mean, cav_cov,  nat1,    nat2,     Lw,
              zw, ldl, part0, part1,    out_mean, out_jac, out_hess, nullptr};
  BN_TIMED(KID_ST_SITES, st);
  if (mode == BN_ST_UPDATE) {
    hipLaunchKernelGGL((k_st_ep_cavity<false>), grid, dim3(kT), kCavitySmem, st, dm, cio, power, mask_pseudo);
    launch_st_sites<M_EP, BN_ST_UPDATE>(grid, st, dm, L, gh, io, lr, ensure_psd, power, mask_pseudo);
  } else {
    hipLaunchKernelGGL((k_st_ep_cavity<true>), grid, dim3(kT), kCavitySmem, st, dm, cio, power, mask_pseudo);
    launch_st_sites<M_EP, BN_ST_ENERGY>(grid, st, dm, L, gh, io, lr, ensure_psd, power, mask_pseudo);
  }
  ST_HIP_CHECK(hipGetLastError());
  return BN_OK;
}

int64_t bn_st_workspace_bytes(const bn_st_dims* dims) {
  StDims dm;
  if (check_dims(dims, dm) != BN_OK) return -1;
  // per run: filter P (D x D) and V (D x M); smoother Ps, Dl, GT (3 x D x D) and m_s (D),
  // then the smoother's block of gains and unpacked covariances (gain_block_steps)
  const int64_t D = dm.D;
  return (int64_t)dm.K * (5 * D * D + 2 * D * D * gain_block_steps(dm)) * (int64_t)sizeof(double) + 4096;
}

int bn_st_filter(const bn_st_dims* dims, const bn_st_hypers* hyp, const double* dt, const double* Lw,
                 const double* zw, const double* ldl, double* filter_mean, double* filter_cov, double* ell_steps,
                 double* loglik, void* workspace, void* stream) {
  StDims dm;
  if (int rc = check_dims(dims, dm)) return rc;
  if (!hyp || !dt || !Lw || !zw || !ldl || !filter_mean || !filter_cov || !ell_steps || !loglik || !workspace)
    return st_fail(BN_EINVAL, "NULL argument to bn_st_filter");
  double* ws = (double*)workspace;
  const int64_t D = dm.D;
  if (int rc = set_smem_limits()) return rc;
  StFilterIO io{dt, Lw, zw, ldl, ws, ws + dm.K * D * D, ws + dm.K * (D * D + D * dm.M), filter_mean, filter_cov,
                ell_steps};
  hipStream_t st = (hipStream_t)stream;
  {
    BN_TIMED(KID_ST_FILTER, st);
    if (dm.KT == 1)
      hipLaunchKernelGGL(k_st_filter<1>, dim3(dm.K), dim3(kT), kFilterSmem, st, dm, make_hyp(hyp), io);
    else
      hipLaunchKernelGGL(k_st_filter<2>, dim3(dm.K), dim3(kT), kFilterSmem, st, dm, make_hyp(hyp), io);
  }
  ST_HIP_CHECK(hipGetLastError());
  {
    BN_TIMED(KID_ST_SUM, st);
    hipLaunchKernelGGL(k_st_sum, dim3(dm.K), dim3(kT), 0, st, dm.T, ell_steps, loglik);
  }
  ST_HIP_CHECK(hipGetLastError());
  return BN_OK;
}

int bn_st_smoother(const bn_st_dims* dims, const bn_st_hypers* hyp, const double* dt, const double* filter_mean,
                   const double* filter_cov, double* post_mean, double* post_cov, void* workspace, void* stream) {
  return bn_st_smoother_full(dims, hyp, dt, filter_mean, filter_cov, post_mean, post_cov, nullptr, nullptr, nullptr,
                             workspace, stream);
}

int bn_st_smoother_full(const bn_st_dims* dims, const bn_st_hypers* hyp, const double* dt, const double* filter_mean,
                        const double* filter_cov, double* post_mean, double* post_cov, double* smooth_mean,
                        double* smooth_cov, double* smooth_cross, void* workspace, void* stream) {
  StDims dm;
  if (int rc = check_dims(dims, dm)) return rc;
  if (!hyp || !dt || !filter_mean || !filter_cov || !post_mean || !post_cov || !workspace)
    return st_fail(BN_EINVAL, "NULL argument to bn_st_smoother");
  if ((smooth_mean == nullptr) != (smooth_cov == nullptr) || (smooth_mean == nullptr) != (smooth_cross == nullptr))
    return st_fail(BN_EINVAL, "bn_st_smoother_full: smooth_mean, smooth_cov and smooth_cross go together");
  if (int rc = set_smem_limits()) return rc;
  double* ws = (double*)workspace;
  const int64_t D = dm.D, K = dm.K, DD = D * D;
  const int tb = gain_block_steps(dm);
  StSmootherIO io{dt, filter_mean, filter_cov, ws, ws + K * DD, ws + 2 * K * DD, ws + 3 * K * DD,
                  ws + 5 * K * DD, tb, post_mean, post_cov, smooth_mean, smooth_cov, smooth_cross};
  hipStream_t st = (hipStream_t)stream;
  const StHyp hy = make_hyp(hyp);
  BN_TIMED(KID_ST_SMOOTHER, st);
  // blocks of steps from the end: gains of the block in parallel, then the recursion over it
  int t1 = dm.T - 2;
  bool terminal = true;
  do {
    const int t0 = t1 - tb + 1 > 0 ? t1 - tb + 1 : 0;
    if (t1 >= 0) {
      const dim3 g((unsigned)(t1 - t0 + 1), (unsigned)dm.K);
      if (dm.KT == 1)
        hipLaunchKernelGGL(k_st_gains<1>, g, dim3(kT), kGainsSmem, st, dm, hy, io, t0);
      else
        hipLaunchKernelGGL(k_st_gains<2>, g, dim3(kT), kGainsSmem, st, dm, hy, io, t0);
    }
    if (dm.KT == 1)
      hipLaunchKernelGGL(k_st_smoother<1>, dim3(dm.K), dim3(kT), kSmootherSeqSmem, st, dm, hy, io, t1, t0,
                         terminal ? 1 : 0);
    else
      hipLaunchKernelGGL(k_st_smoother<2>, dim3(dm.K), dim3(kT), kSmootherSeqSmem, st, dm, hy, io, t1, t0,
                         terminal ? 1 : 0);
    ST_HIP_CHECK(hipGetLastError());
    terminal = false;
    t1 = t0 - 1;
  } while (t1 >= 0);
  return BN_OK;
}

int bn_st_marginals(const bn_st_dims* dims, const double* B, const double* cdiag, const double* post_mean,
                    const double* post_cov, double* mean_f, double* var_f, void* stream) {
  StDims dm;
  if (int rc = check_dims(dims, dm)) return rc;
  if (!B || !cdiag || !post_mean || !post_cov || !mean_f || !var_f)
    return st_fail(BN_EINVAL, "NULL argument to bn_st_marginals");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_st_marginals, dim3(dm.T, dm.K), dim3(kT), 0, st, dm, B, cdiag, post_mean, post_cov, mean_f,
                     var_f);
  ST_HIP_CHECK(hipGetLastError());
  return BN_OK;
}

int bn_st_sum(int32_t n_runs, int32_t n_steps, const double* per_step, double* per_run, void* stream) {
  if (n_runs < 1 || n_steps < 1 || !per_step || !per_run) return st_fail(BN_EINVAL, "bad argument to bn_st_sum");
  hipStream_t st = (hipStream_t)stream;
  BN_TIMED(KID_ST_SUM, st);
  hipLaunchKernelGGL(k_st_sum, dim3(n_runs), dim3(kT), 0, st, n_steps, per_step, per_run);
  ST_HIP_CHECK(hipGetLastError());
  return BN_OK;
}

int64_t bn_mf_workspace_bytes(const bn_st_dims* dims) {
  StDims dm;
  if (check_dims(dims, dm) != BN_OK) return -1;
  return 2 * (int64_t)dm.K * dm.T * dm.M * (int64_t)sizeof(double) + 256;
}

int bn_mf_pseudo(const bn_st_dims* dims, const double* Lw, const double* zw, double* noise, double* y,
                 void* stream) {
  StDims dm;
  if (int rc = check_dims(dims, dm)) return rc;
  if (!Lw || !zw || !noise || !y) return st_fail(BN_EINVAL, "NULL argument to bn_mf_pseudo");
  if (int rc = set_smem_limits()) return rc;
  hipStream_t st = (hipStream_t)stream;
  BN_TIMED(KID_MF_PSEUDO, st);
  hipLaunchKernelGGL(k_mf_pseudo, dim3(dm.T, dm.K), dim3(kT), kPseudoSmem, st, dm, Lw, zw, noise, y);
  ST_HIP_CHECK(hipGetLastError());
  return BN_OK;
}

int bn_mf_filter(const bn_st_dims* dims, const bn_st_hypers* hyp, int32_t parallel, const double* dt,
                 const double* y, const double* noise, double* filter_mean, double* filter_cov, double* ell_steps,
                 double* loglik, void* workspace, void* stream) {
  StDims dm;
  if (int rc = check_dims(dims, dm)) return rc;
  if (!hyp || !dt || !y || !noise || !filter_mean || !filter_cov || !ell_steps || !loglik)
    return st_fail(BN_EINVAL, "NULL argument to bn_mf_filter");
  if (parallel && !workspace) return st_fail(BN_EINVAL, "the associative form needs bn_mf_workspace_bytes of workspace");
  double* ws = (double*)workspace;
  const int64_t ktm = (int64_t)dm.K * dm.T * dm.M;
  MfFilterIO io{dt, y, noise, filter_mean, filter_cov, ell_steps, ws, ws ? ws + ktm : nullptr};
  const StHyp hy = make_hyp(hyp);
  hipStream_t st = (hipStream_t)stream;
  if (!parallel) {
    BN_TIMED(KID_MF_FILTER, st);
    if (dm.KT == 1) hipLaunchKernelGGL(k_mf_filter<1>, dim3(dm.K), dim3(kT), 0, st, dm, hy, io);
    else hipLaunchKernelGGL(k_mf_filter<2>, dim3(dm.K), dim3(kT), 0, st, dm, hy, io);
    ST_HIP_CHECK(hipGetLastError());
  } else {
    const dim3 grid(dm.T, dm.K);
    const int nb = (int)((dm.K * (int64_t)dm.M + 63) / 64);
    {
      BN_TIMED(KID_MF_SDIAG, st);
      if (dm.KT == 1) hipLaunchKernelGGL((k_mf_sdiag<1, false>), grid, dim3(kT), 0, st, dm, hy, io);
      else hipLaunchKernelGGL((k_mf_sdiag<2, false>), grid, dim3(kT), 0, st, dm, hy, io);
      ST_HIP_CHECK(hipGetLastError());
    }
    {
      BN_TIMED(KID_MF_BLOCKS, st);
      if (dm.KT == 1) hipLaunchKernelGGL(k_mf_blocks<1>, dim3(nb), dim3(64), 0, st, dm, hy, io);
      else hipLaunchKernelGGL(k_mf_blocks<2>, dim3(nb), dim3(64), 0, st, dm, hy, io);
      ST_HIP_CHECK(hipGetLastError());
    }
    {
      BN_TIMED(KID_MF_SDIAG, st);
      if (dm.KT == 1) hipLaunchKernelGGL((k_mf_sdiag<1, true>), grid, dim3(kT), 0, st, dm, hy, io);
      else hipLaunchKernelGGL((k_mf_sdiag<2, true>), grid, dim3(kT), 0, st, dm, hy, io);
      ST_HIP_CHECK(hipGetLastError());
    }
  }
  {
    BN_TIMED(KID_ST_SUM, st);
    hipLaunchKernelGGL(k_st_sum, dim3(dm.K), dim3(kT), 0, st, dm.T, ell_steps, loglik);
  }
  ST_HIP_CHECK(hipGetLastError());
  return BN_OK;
}

int bn_mf_smoother(const bn_st_dims* dims, const bn_st_hypers* hyp, const double* dt, const double* filter_mean,
                   const double* filter_cov, double* post_mean, double* post_var, int64_t var_step_stride,
                   int64_t var_elem_stride, double* smooth_mean, double* smooth_cov, double* gains, void* stream) {
  StDims dm;
  if (int rc = check_dims(dims, dm)) return rc;
  if (!hyp || !dt || !filter_mean || !filter_cov) return st_fail(BN_EINVAL, "NULL argument to bn_mf_smoother");
  if (!post_mean && !post_var && !smooth_mean && !smooth_cov && !gains)
    return st_fail(BN_EINVAL, "bn_mf_smoother: no output requested");
  if (post_var && (var_step_stride < 0 || var_elem_stride < 0 ||
                   (dm.M - 1) * var_elem_stride >= (var_step_stride > 0 ? var_step_stride : INT64_MAX)))
    return st_fail(BN_EINVAL, "bn_mf_smoother: post_var strides overlap");
  MfSmootherIO io{dt, filter_mean, filter_cov, post_mean, post_var, var_step_stride, var_elem_stride,
                  smooth_mean, smooth_cov, gains};
  hipStream_t st = (hipStream_t)stream;
  const int nb = (int)((dm.K * (int64_t)dm.M + 63) / 64);
  BN_TIMED(KID_MF_SMOOTHER, st);
  if (dm.KT == 1) hipLaunchKernelGGL(k_mf_smoother<1>, dim3(nb), dim3(64), 0, st, dm, make_hyp(hyp), io);
  else hipLaunchKernelGGL(k_mf_smoother<2>, dim3(nb), dim3(64), 0, st, dm, make_hyp(hyp), io);
  ST_HIP_CHECK(hipGetLastError());
  return BN_OK;
}

int64_t bn_ih_workspace_bytes(const bn_st_dims* dims, int32_t heteroscedastic) {
  StDims dm;
  if (check_dims(dims, dm) != BN_OK) return -1;
  const int64_t K = dm.K, D = dm.D, M = dm.M, T = dm.T;
  int64_t n = 3 * K * D * D + K * M * M + K + K * kTiedParts * M * M;
  if (heteroscedastic) n += K * T * M * M + K * T;
  return n * (int64_t)sizeof(double) + 256;
}

int bn_ih_tied(const bn_st_dims* dims, const double* Lw, double* noise_tied, void* workspace, void* stream) {
  StDims dm;
  if (int rc = check_dims(dims, dm)) return rc;
  if (!Lw || !noise_tied || !workspace) return st_fail(BN_EINVAL, "NULL argument to bn_ih_tied");
  // partial sums after bn_ih_filter's regions (see bn_ih_workspace_bytes)
  const int64_t K = dm.K, D = dm.D, M = dm.M;
  double* part = (double*)workspace + 3 * K * D * D + K * M * M + K;
  hipStream_t st = (hipStream_t)stream;
  BN_TIMED(KID_IH_TIED, st);
  hipLaunchKernelGGL(k_ih_tied_part, dim3(kTiedParts, dm.K), dim3(kT), kTiedSmem, st, dm, Lw, part);
  ST_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_ih_tied, dim3(dm.K), dim3(kT), 0, st, dm, part, noise_tied);
  ST_HIP_CHECK(hipGetLastError());
  return BN_OK;
}

int bn_ih_filter(const bn_st_dims* dims, const bn_st_hypers* hyp, const double* step, int32_t iters,
                 int32_t heteroscedastic, const double* y, const double* noise, const double* noise_tied, double* dare,
                 double* filter_cov, double* filter_mean, double* ell_steps, double* loglik, void* workspace,
                 void* stream) {
  StDims dm;
  if (int rc = check_dims(dims, dm)) return rc;
  if (!hyp || !step || !y || !noise_tied || !dare || !filter_cov || !filter_mean || !ell_steps || !loglik ||
      !workspace || (heteroscedastic && !noise))
    return st_fail(BN_EINVAL, "NULL argument to bn_ih_filter");
  if (iters < 0) return st_fail(BN_EINVAL, "bn_ih_filter: iters must be >= 0");
  if (int rc = set_smem_limits()) return rc;
  const int64_t K = dm.K, D = dm.D, M = dm.M;
  double* ws = (double*)workspace;
  IhIO io{};
  io.step = step;
  io.y = y;
  io.R = noise;
  io.Rt = noise_tied;
  io.X = dare;
  io.cov = filter_cov;
  io.W1 = ws;
  io.W2 = ws + K * D * D;
  io.STn = ws + 3 * K * D * D;
  io.ldT = io.STn + K * M * M;
  io.Sn = heteroscedastic ? io.ldT + K + K * kTiedParts * M * M : nullptr;
  io.ldS = heteroscedastic ? io.Sn + K * dm.T * M * M : nullptr;
  io.fmean = filter_mean;
  io.ell = ell_steps;
  io.iters = iters;
  io.hetero = heteroscedastic ? 1 : 0;
  const StHyp hy = make_hyp(hyp);
  hipStream_t st = (hipStream_t)stream;
  {
    BN_TIMED(KID_IH_DARE, st);
    if (dm.KT == 1) hipLaunchKernelGGL(k_ih_dare<1>, dim3(dm.K), dim3(kT), kIhDareSmem, st, dm, hy, io);
    else hipLaunchKernelGGL(k_ih_dare<2>, dim3(dm.K), dim3(kT), kIhDareSmem, st, dm, hy, io);
    ST_HIP_CHECK(hipGetLastError());
  }
  if (heteroscedastic) {
    BN_TIMED(KID_IH_SINV, st);
    hipLaunchKernelGGL(k_ih_sinv, dim3(dm.T, dm.K), dim3(kT), 0, st, dm, io);
    ST_HIP_CHECK(hipGetLastError());
  }
  {
    BN_TIMED(KID_IH_FILTER, st);
    if (dm.KT == 1) hipLaunchKernelGGL(k_ih_filter<1>, dim3(dm.K), dim3(kT), 0, st, dm, hy, io);
    else hipLaunchKernelGGL(k_ih_filter<2>, dim3(dm.K), dim3(kT), 0, st, dm, hy, io);
    ST_HIP_CHECK(hipGetLastError());
  }
  {
    BN_TIMED(KID_ST_SUM, st);
    hipLaunchKernelGGL(k_st_sum, dim3(dm.K), dim3(kT), 0, st, dm.T, ell_steps, loglik);
  }
  ST_HIP_CHECK(hipGetLastError());
  return BN_OK;
}

int bn_ih_smoother(const bn_st_dims* dims, const bn_st_hypers* hyp, const double* step, int32_t iters,
                   const double* filter_mean, const double* dare, const double* filter_cov, double* dare_smoother,
                   double* gain, double* post_mean, double* post_cov, double* smooth_mean, void* workspace,
                   void* stream) {
  StDims dm;
  if (int rc = check_dims(dims, dm)) return rc;
  if (!hyp || !step || !filter_mean || !dare || !filter_cov || !dare_smoother || !workspace)
    return st_fail(BN_EINVAL, "NULL argument to bn_ih_smoother");
  if (iters < 0) return st_fail(BN_EINVAL, "bn_ih_smoother: iters must be >= 0");
  if (int rc = set_smem_limits()) return rc;
  const int64_t K = dm.K, D = dm.D;
  double* ws = (double*)workspace;
  IhSmIO io{};
  io.step = step;
  io.fmean = filter_mean;
  io.X = dare;
  io.cov = filter_cov;
  io.Xs = dare_smoother;
  io.G = gain ? gain : ws + 2 * K * D * D;
  io.W1 = ws;
  io.W2 = ws + K * D * D;
  io.pm = post_mean;
  io.pP = post_cov;
  io.smean = smooth_mean;
  io.iters = iters;
  hipStream_t st = (hipStream_t)stream;
  BN_TIMED(KID_IH_SMOOTHER, st);
  if (dm.KT == 1)
    hipLaunchKernelGGL(k_ih_smoother<1>, dim3(dm.K), dim3(kT), kSmootherSmem, st, dm, make_hyp(hyp), io);
  else
    hipLaunchKernelGGL(k_ih_smoother<2>, dim3(dm.K), dim3(kT), kSmootherSmem, st, dm, make_hyp(hyp), io);
  ST_HIP_CHECK(hipGetLastError());
  return BN_OK;
}

int bn_st_predict(const bn_st_dims* dims, const bn_st_hypers* hyp, const double* t_train, int32_t n_test,
                  int32_t n_loc_test, const double* t_test, const double* B_test, const double* cdiag_test,
                  const double* smooth_mean, const double* smooth_cov, const double* smooth_cross,
                  const double* steady_cov, double* mean, double* var, void* stream) {
  StDims dm;
  if (int rc = check_dims(dims, dm)) return rc;
  if (!hyp || !t_train || !t_test || !B_test || !cdiag_test || !smooth_mean || !mean || !var ||
      (!steady_cov && (!smooth_cov || !smooth_cross)))
    return st_fail(BN_EINVAL, "NULL argument to bn_st_predict");
  if (n_test < 1 || n_loc_test < 1) return st_fail(BN_EINVAL, "bn_st_predict: empty test set");
  StPredIO io{t_train, t_test, B_test, cdiag_test, smooth_mean, smooth_cov, smooth_cross, steady_cov, mean, var, n_test,
              n_loc_test};
  hipStream_t st = (hipStream_t)stream;
  if (dm.KT == 1)
    hipLaunchKernelGGL(k_st_predict<1>, dim3(n_test, dm.K), dim3(kT), 0, st, dm, make_hyp(hyp), io);
  else
    hipLaunchKernelGGL(k_st_predict<2>, dim3(n_test, dm.K), dim3(kT), 0, st, dm, make_hyp(hyp), io);
  ST_HIP_CHECK(hipGetLastError());
  return BN_OK;
}

}  // extern "C"
