// Per-lane math of the Markov-GP hot path: state-space discretisation, the
// Kalman filter / RTS smoother step, and the associative elements with their
// combine operators.  Everything here is __host__ __device__ so the same code
// runs in the HIP kernels (one lane = one chunk of time steps) and in a host
// build used to check the kernels' data flow on the CPU.
//
// State-space models are sums of Matern components (kernels.py:123-300,
// Sum 1624-1663): the component "kind" is its state dimension (1 = Matern12,
// 2 = Matern32, 3 = Matern52).  A, Q and Pinf are block diagonal, so every
// product below skips the zero blocks at compile time.  Observations are
// scalar per step (p = 1): H sums the first state of each component.
#pragma once
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define BN_HD __host__ __device__ __forceinline__
#else
#include <cmath>
#define BN_HD inline
#endif
#include <stdint.h>

#if defined(__HIPCC__)
#define BN_QTAB __constant__ static const
#else
#define BN_QTAB static const
#endif
#include "bn_qtaylor.h"

namespace bn {

#if defined(__HIPCC__)
#define BN_UNROLL _Pragma("unroll")
#else
#define BN_UNROLL
#endif

static constexpr double kLog2Pi = 1.8378770664093453;
static constexpr double kLn2 = 0.6931471805599453;
#ifndef BN_SEMICHOL_RTOL
#define BN_SEMICHOL_RTOL 1e-9
#endif

// ---------------------------------------------------------------------------
// Model structure
// ---------------------------------------------------------------------------
template <int K1, int K2>
struct Spec {
  static constexpr int D = K1 + K2;
  static constexpr int NS = D * (D + 1) / 2;  // packed symmetric size
  static constexpr BN_HD int blk(int i) { return i < K1 ? 0 : 1; }
  static constexpr BN_HD bool nz(int i, int j) { return blk(i) == blk(j); }
  static constexpr BN_HD bool h(int i) { return i == 0 || (K2 > 0 && i == K1); }
  static constexpr BN_HD int sidx(int i, int j) {  // packed upper index, i <= j
    return i * D - (i * (i - 1)) / 2 + (j - i);
  }
};

// True when `p` holds on every active lane of the wavefront (the host build: `p`).
BN_HD bool lanes_all(bool p) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __all(p) != 0;
#else
  return p;
#endif
}
BN_HD bool lanes_any(bool p) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __any(p) != 0;
#else
  return p;
#endif
}
// a * x + c for a wave-uniform c (held in scalar registers): one VOP3 FMA reading c
// from its scalar registers, where the compiler's own choice copies c into vector
// registers first to use the two-operand accumulate form.
BN_HD double fma_uniform_addend(double a, double x, double c) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r;
  asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(x), "s"(c));
  return r;
#else
  return std::fma(a, x, c);
#endif
}
// log x.  For finite x > 0: x = m 2^e with m in [1/sqrt 2, sqrt 2), log m = 2 atanh s
// with s = (m - 1) / (m + 1), |s| < 0.172, summed by the atanh Taylor series (eleven
// terms reach double precision there), plus e ln 2 with ln 2 split so that e * hi is
// exact.  About 30 instructions against the library's 100 or so, within a few ulp of it;
// 0, negative, infinite and NaN x give log's values by selects (no branch, so nothing
// for the compiler to evaluate speculatively).
BN_HD double log_fast(double x) {
  constexpr double kLn2Hi = 0.6931471803691238;      // ln 2 to 32 bits: e * kLn2Hi is exact
  constexpr double kLn2Lo = 1.9082149292705877e-10;  // ln 2 - kLn2Hi
  int e;
  double m = frexp(x, &e);  // m in [0.5, 1) for finite x > 0
  if (m < 0.70710678118654752440) {
    m += m;
    --e;
  }
#if defined(__HIP_DEVICE_COMPILE__)
  // m + 1 is in [1.7, 2.5]: the hardware reciprocal refined by two Newton steps, in
  // place of the full division sequence (which guards ranges that cannot occur here)
  const double den = m + 1.0;
  double rc = __builtin_amdgcn_rcp(den);
  rc = fma(fma(-den, rc, 1.0), rc, rc);
  rc = fma(fma(-den, rc, 1.0), rc, rc);
  const double s = (m - 1.0) * rc;
#else
  const double s = (m - 1.0) / (m + 1.0);
#endif
  const double z = s * s;
  double p = 1.0 / 23.0;  // sum_k z^k / (2k + 3), k = 0..10
  BN_UNROLL for (int k = 9; k >= 0; --k) p = fma_uniform_addend(p, z, 1.0 / (2 * k + 3));
  const double r = 2.0 * s + (2.0 * s) * (z * p);
  const double v = (double)e * kLn2Hi + ((double)e * kLn2Lo + r);
  const double inf = __builtin_huge_val();
  const double special = x == 0.0 ? -inf : (x > 0.0 ? x : __builtin_nan(""));
  return (x > 0.0 && x < inf) ? v : special;
}
// 1 / sqrt(x): the device's reciprocal square root (one refined sequence in place of a
// square root and a division; within an ulp or two of the two), 1 / sqrt on the host.
BN_HD double recip_sqrt(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return rsqrt(x);
#else
  return 1.0 / std::sqrt(x);
#endif
}

// Closed-form A = expm(F dt) of one component (kernels.py:156-163, 216-224, 273-286),
// given e = exp(-lam dt).
template <int K>
BN_HD void comp_transition_e(double lam, double dt, double e, double* A) {
  if (K == 1) {
    A[0] = e;
  } else if (K == 2) {
    A[0] = e * (dt * lam + 1.0);
    A[1] = e * dt;
    A[2] = e * (-dt * lam * lam);
    A[3] = e * (-dt * lam + 1.0);
  } else {
    const double dl = dt * lam;
    A[0] = e * (dt * lam * (0.5 * dl + 1.0) + 1.0);
    A[1] = e * (dt * (dl + 1.0));
    A[2] = e * (dt * 0.5 * dt);
    A[3] = e * (dt * (-0.5 * dl * lam * lam));
    A[4] = e * (dt * lam * (1.0 - dl) + 1.0);
    A[5] = e * (dt * (1.0 - 0.5 * dl));
    A[6] = e * (dt * lam * lam * lam * (0.5 * dl - 1.0));
    A[7] = e * (dt * lam * lam * (dl - 3.0));
    A[8] = e * (dt * lam * (0.5 * dl - 2.0) + 1.0);
  }
}

template <int K>
BN_HD void comp_transition(double lam, double dt, double* A) {
  comp_transition_e<K>(lam, dt, exp(-(lam * dt)), A);
}

// Stationary covariance of one component (kernels.py:144-146, 207-210, 288-293).
template <int K>
BN_HD void comp_pinf(double ls, double var, double* P) {
  if (K == 1) {
    P[0] = var;
  } else if (K == 2) {
    P[0] = var; P[1] = 0.0; P[2] = 0.0; P[3] = 3.0 * var / (ls * ls);
  } else {
    double kap = 5.0 / 3.0 * var / (ls * ls);
    P[0] = var; P[1] = 0.0; P[2] = -kap;
    P[3] = 0.0; P[4] = kap; P[5] = 0.0;
    P[6] = -kap; P[7] = 0.0; P[8] = 25.0 * var / (ls * ls * ls * ls);
  }
}

template <int K>
BN_HD double comp_lambda(double ls) {
  return K == 1 ? 1.0 / ls : (K == 2 ? sqrt(3.0) / ls : sqrt(5.0) / ls);
}

// Row `idx` (packed upper index) of the small-x series or direct-form polynomial table.
template <int K>
BN_HD const double* qtab(int idx, bool small) {
  if (K == 1) return small ? kQSeries1[idx] : kQPoly1[idx];
  if (K == 2) return small ? kQSeries2[idx] : kQPoly2[idx];
  return small ? kQSeries3[idx] : kQPoly3[idx];
}
template <int K>
BN_HD int qlen(int idx) {
  return K == 1 ? kQLen1[idx] : (K == 2 ? kQLen2[idx] : kQLen3[idx]);
}
template <int K>
BN_HD double qc(int idx) {
  return K == 1 ? kQC1[idx] : (K == 2 ? kQC2[idx] : kQC3[idx]);
}
template <int K>
BN_HD constexpr int qpolydeg() {
  return K == 1 ? kQPolyDeg1 : (K == 2 ? kQPolyDeg2 : kQPolyDeg3);
}
// declared row width of the small-x series tables (rows shorter than it are zero-padded)
template <int K>
BN_HD constexpr int qser_width() {
  return K == 1 ? (int)(sizeof(kQSeries1[0]) / sizeof(double))
                : (K == 2 ? (int)(sizeof(kQSeries2[0]) / sizeof(double)) : (int)(sizeof(kQSeries3[0]) / sizeof(double)));
}

// Per-series model: hypers and Pinf, in FP64.
template <int K1, int K2>
struct Model {
  using S = Spec<K1, K2>;
  static constexpr int D = S::D;
  double lam1, lam2, var1, var2;
  double Pinf[D][D];

  BN_HD void init(double ls1, double v1, double ls2, double v2) {
    var1 = v1;
    var2 = v2;
    lam1 = comp_lambda<K1>(ls1);
    lam2 = K2 > 0 ? comp_lambda<(K2 > 0 ? K2 : 1)>(ls2) : 0.0;
    double b1[9], b2[9];
    comp_pinf<K1>(ls1, v1, b1);
    if (K2 > 0) comp_pinf<(K2 > 0 ? K2 : 1)>(ls2, v2, b2);
    BN_UNROLL for (int i = 0; i < D; ++i)
      BN_UNROLL for (int j = 0; j < D; ++j) {
        double v = 0.0;
        if (i < K1 && j < K1) v = b1[i * K1 + j];
        else if (i >= K1 && j >= K1) v = b2[(i - K1) * (K2 > 0 ? K2 : 1) + (j - K1)];
        Pinf[i][j] = v;
      }
  }

  // A(dt) and Q(dt) = Pinf - A Pinf A' (ops.py:149-151) in FP64, rounded to T.
  // Q comes from the cancellation-free form Q_ij = var lam^(i+j) F_ij(lam dt)
  // (bn_qtaylor.h): evaluating Pinf - A Pinf A' as written leaves only
  // rounding noise in the O(dt^5) entries when dt << lengthscale, which breaks
  // one-step scan elements.  Only the nonzero blocks are meaningful.
  template <class T>
  BN_HD void discretise(double dt, T A[D][D], T Q[D][D]) const {
    BN_UNROLL for (int i = 0; i < D; ++i)
      BN_UNROLL for (int j = 0; j < D; ++j) { A[i][j] = T(0); Q[i][j] = T(0); }
    comp_block<K1, 0>(lam1, var1, dt, A, Q);
    if (K2 > 0) comp_block<(K2 > 0 ? K2 : 1), K1>(lam2, var2, dt, A, Q);
  }

  template <int K, int O, class T>
  BN_HD static void comp_block(double lam, double var, double dt, T A[D][D], T Q[D][D]) {
    const double x = lam * dt;
    const double e = exp(-x);  // shared by A and Q's direct form
    double a[9];
    comp_transition_e<K>(lam, dt, e, a);
    BN_UNROLL for (int i = 0; i < K; ++i)
      BN_UNROLL for (int j = 0; j < K; ++j) A[O + i][O + j] = (T)a[i * K + j];
    const bool small = x < kQX0;
    // which Q forms the wavefront needs: a form no active lane uses is not evaluated
    const bool need_s = lanes_any(small), need_p = lanes_any(!small);
    const double e2 = e * e;
    // An opaque zero offset keeps the coefficient loads at their use: hoisted out of a
    // caller's step loop, the tables outgrow the scalar registers and spill into
    // vector-register lanes, read back with one v_readlane per use.
    int zo = 0;
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+s"(zo));
#endif
    // The forms some lane of the wavefront needs, from coefficient rows every lane reads
    // at the same address (scalar loads, independent of the lane's x), fully unrolled,
    // then a per-lane select.  Small x: F_ij = x^(2K-1-i-j) G_ij(x); else
    // F_ij = C_ij - e^{-2x} P_ij(x).  A per-lane choice of table turns every coefficient
    // into a dependent vector load on the recurrence's critical path.  The series rows
    // are zero-padded to their declared width, which leaves the Horner value unchanged.
    // The entries' Horner chains are interleaved (independent FMAs back to back).
    constexpr int NQ = K * (K + 1) / 2;
    constexpr int WS = qser_width<K>(), WP = qpolydeg<K>();
    auto series = [&](double (&o)[NQ]) {
      BN_UNROLL for (int q = 0; q < NQ; ++q) o[q] = qtab<K>(q, true)[zo + WS - 1];
      BN_UNROLL for (int n = WS - 2; n >= 0; --n)
        BN_UNROLL for (int q = 0; q < NQ; ++q) o[q] = fma_uniform_addend(o[q], x, qtab<K>(q, true)[zo + n]);
      double xp[2 * K];  // x^m
      xp[0] = 1.0;
      BN_UNROLL for (int m = 1; m < 2 * K; ++m) xp[m] = (m & 1) ? xp[m - 1] * x : xp[m / 2] * xp[m / 2];
      int q = 0;
      BN_UNROLL for (int i = 0; i < K; ++i)
        BN_UNROLL for (int j = i; j < K; ++j) { o[q] *= xp[2 * K - 1 - i - j]; ++q; }
    };
    auto direct = [&](double (&o)[NQ]) {
      BN_UNROLL for (int q = 0; q < NQ; ++q) o[q] = qtab<K>(q, false)[zo + WP - 1];
      BN_UNROLL for (int n = WP - 2; n >= 0; --n)
        BN_UNROLL for (int q = 0; q < NQ; ++q) o[q] = fma_uniform_addend(o[q], x, qtab<K>(q, false)[zo + n]);
      BN_UNROLL for (int q = 0; q < NQ; ++q) o[q] = qc<K>(q) - e2 * o[q];
    };
    double v[NQ];
    if (need_p) {
      direct(v);
      if (need_s) {
        double w[NQ];
        series(w);
        BN_UNROLL for (int q = 0; q < NQ; ++q) v[q] = small ? w[q] : v[q];
      }
    } else {
      series(v);
    }
    int idx = 0;
    double lp_i = var;  // var * lam^(i + j), built incrementally
    BN_UNROLL for (int i = 0; i < K; ++i) {
      double lp = lp_i;
      BN_UNROLL for (int j = i; j < K; ++j) {
        const T q = (T)(lp * v[idx]);
        Q[O + i][O + j] = q;
        Q[O + j][O + i] = q;
        lp *= lam;
        ++idx;
      }
      lp_i *= lam * lam;
    }
  }
};

// ---------------------------------------------------------------------------
// Small dense helpers (fixed size, fully unrolled)
// ---------------------------------------------------------------------------
template <int D, class T>
BN_HD void set_identity(T A[D][D]) {
  BN_UNROLL for (int i = 0; i < D; ++i)
    BN_UNROLL for (int j = 0; j < D; ++j) A[i][j] = (i == j) ? T(1) : T(0);
}

template <int D, class T>
BN_HD void set_zero(T A[D][D]) {
  BN_UNROLL for (int i = 0; i < D; ++i)
    BN_UNROLL for (int j = 0; j < D; ++j) A[i][j] = T(0);
}

// C = A B
template <int D, class T>
BN_HD void mat_mul(const T A[D][D], const T B[D][D], T C[D][D]) {
  BN_UNROLL for (int i = 0; i < D; ++i)
    BN_UNROLL for (int j = 0; j < D; ++j) {
      T s = T(0);
      BN_UNROLL for (int k = 0; k < D; ++k) s += A[i][k] * B[k][j];
      C[i][j] = s;
    }
}

// C = A B', symmetric result assumed: fills upper and mirrors
template <int D, class T>
BN_HD void mat_mul_bt_sym(const T A[D][D], const T B[D][D], T C[D][D]) {
  BN_UNROLL for (int i = 0; i < D; ++i)
    BN_UNROLL for (int j = i; j < D; ++j) {
      T s = T(0);
      BN_UNROLL for (int k = 0; k < D; ++k) s += A[i][k] * B[j][k];
      C[i][j] = s;
      C[j][i] = s;
    }
}

// Semidefinite Cholesky in FP64: A = U U' with U lower; a pivot at or below
// rtol * A[j][j] (rounding noise of an FP32 PSD matrix) becomes a zero column.
template <int D>
BN_HD void semichol(double A[D][D], double rtol) {
  BN_UNROLL for (int j = 0; j < D; ++j) {
    double s = A[j][j];
    double d0 = s;
    BN_UNROLL for (int k = 0; k < j; ++k) s -= A[j][k] * A[j][k];
    bool ok = s > rtol * (d0 > 0.0 ? d0 : 0.0) && s > 0.0;
    double l = ok ? sqrt(s) : 0.0;
    double il = ok ? 1.0 / l : 0.0;
    A[j][j] = l;
    BN_UNROLL for (int i = j + 1; i < D; ++i) {
      double t = A[i][j];
      BN_UNROLL for (int k = 0; k < j; ++k) t -= A[i][k] * A[j][k];
      A[i][j] = t * il;
    }
    BN_UNROLL for (int i = 0; i < j; ++i) A[i][j] = 0.0;
  }
}

// Cholesky of an SPD matrix in FP64 (lower, in place).  Returns false when a
// pivot is not positive (the reference then yields NaN; callers propagate it).
template <int D>
BN_HD bool chol_spd(double A[D][D]) {
  bool ok = true;
  BN_UNROLL for (int j = 0; j < D; ++j) {
    double s = A[j][j];
    BN_UNROLL for (int k = 0; k < j; ++k) s -= A[j][k] * A[j][k];
    ok = ok && (s > 0.0);
    double il = recip_sqrt(s);
    double l = s * il;
    A[j][j] = l;
    BN_UNROLL for (int i = j + 1; i < D; ++i) {
      double t = A[i][j];
      BN_UNROLL for (int k = 0; k < j; ++k) t -= A[i][k] * A[j][k];
      A[i][j] = t * il;
    }
  }
  return ok;
}

// X <- L^-1 X (forward substitution), L lower FP64, X has D rows and NC columns.
template <int D, int NC, class T>
BN_HD void lower_solve(const double L[D][D], T X[D][NC]) {
  // one division per pivot, not per right-hand side (an FP64 division is a ten
  // instruction sequence; the product by the reciprocal differs by at most an ulp)
  double il[D];
  BN_UNROLL for (int i = 0; i < D; ++i) il[i] = 1.0 / L[i][i];
  BN_UNROLL for (int c = 0; c < NC; ++c)
    BN_UNROLL for (int i = 0; i < D; ++i) {
      double s = (double)X[i][c];
      BN_UNROLL for (int k = 0; k < i; ++k) s -= L[i][k] * (double)X[k][c];
      X[i][c] = (T)(s * il[i]);
    }
}

// X <- L'^-1 X (backward substitution with L').
template <int D, int NC, class T>
BN_HD void upper_t_solve(const double L[D][D], T X[D][NC]) {
  double il[D];
  BN_UNROLL for (int i = 0; i < D; ++i) il[i] = 1.0 / L[i][i];
  BN_UNROLL for (int c = 0; c < NC; ++c)
    BN_UNROLL for (int i = D - 1; i >= 0; --i) {
      double s = (double)X[i][c];
      BN_UNROLL for (int k = i + 1; k < D; ++k) s -= L[k][i] * (double)X[k][c];
      X[i][c] = (T)(s * il[i]);
    }
}

// ---------------------------------------------------------------------------
// Associative filtering element (ops.py:183-219), kept as
//   x_e | x_{s-1} ~ N(Phi x + beta, Sig),  log p(y_s..e | x_{s-1}) = -x'Jx/2 + eta'x + c
// A "state" (a filtered marginal) is an element with Phi = 0, J = 0, eta = 0.
// ---------------------------------------------------------------------------
template <int D, class T>
struct FElem {
  T Phi[D][D];
  T beta[D];
  T Sig[D][D];
  T J[D][D];
  T eta[D];
  static constexpr int NF = D * D + 2 * D + D * (D + 1);  // stored floats

  BN_HD void set_identity_elem() {
    set_identity<D>(Phi);
    set_zero<D>(Sig);
    set_zero<D>(J);
    BN_UNROLL for (int i = 0; i < D; ++i) { beta[i] = T(0); eta[i] = T(0); }
  }
  // field f of the packed storage order: Phi, beta, Sig(upper), J(upper), eta
  template <class Store>
  BN_HD void store(Store st) const {
    int f = 0;
    BN_UNROLL for (int i = 0; i < D; ++i) BN_UNROLL for (int j = 0; j < D; ++j) st(f++, Phi[i][j]);
    BN_UNROLL for (int i = 0; i < D; ++i) st(f++, beta[i]);
    BN_UNROLL for (int i = 0; i < D; ++i) BN_UNROLL for (int j = i; j < D; ++j) st(f++, Sig[i][j]);
    BN_UNROLL for (int i = 0; i < D; ++i) BN_UNROLL for (int j = i; j < D; ++j) st(f++, J[i][j]);
    BN_UNROLL for (int i = 0; i < D; ++i) st(f++, eta[i]);
  }
  template <class Load>
  BN_HD void load(Load ld) {
    int f = 0;
    BN_UNROLL for (int i = 0; i < D; ++i) BN_UNROLL for (int j = 0; j < D; ++j) Phi[i][j] = ld(f++);
    BN_UNROLL for (int i = 0; i < D; ++i) beta[i] = ld(f++);
    BN_UNROLL for (int i = 0; i < D; ++i) BN_UNROLL for (int j = i; j < D; ++j) { T v = ld(f++); Sig[i][j] = v; Sig[j][i] = v; }
    BN_UNROLL for (int i = 0; i < D; ++i) BN_UNROLL for (int j = i; j < D; ++j) { T v = ld(f++); J[i][j] = v; J[j][i] = v; }
    BN_UNROLL for (int i = 0; i < D; ++i) eta[i] = ld(f++);
  }
};

// Shared factorisation of T = (I + Sig1 J2)^-1 (the reference's
// solve(inv(C1) + J2, inv(C1)), ops.py:208-209) without inverting Sig1:
// Sig1 = U U', W = I + U' J2 U = Wc Wc', Z = Wc^-1 U', Y = Wc^-1 U' J2.  Then
//   T X = X - Z'(Y X),  T' x = x - Y'(Z x),  T Sig1 = Z'Z,  J2 T = J2 - Y'Y.
// W has eigenvalues >= 1, so both factorisations are well conditioned.
template <int D, class T>
BN_HD void combine_factor(const T Sig1[D][D], const T J2[D][D], T Z[D][D], T Y[D][D]) {
  double U[D][D];
  BN_UNROLL for (int i = 0; i < D; ++i) BN_UNROLL for (int j = 0; j < D; ++j) U[i][j] = (double)Sig1[i][j];
  semichol<D>(U, BN_SEMICHOL_RTOL);
  // V = J2 U ; W = I + U' V
  double V[D][D];
  BN_UNROLL for (int i = 0; i < D; ++i)
    BN_UNROLL for (int j = 0; j < D; ++j) {
      double s = 0.0;
      BN_UNROLL for (int k = j; k < D; ++k) s += (double)J2[i][k] * U[k][j];
      V[i][j] = s;
    }
  double W[D][D];
  BN_UNROLL for (int i = 0; i < D; ++i)
    BN_UNROLL for (int j = i; j < D; ++j) {
      double s = (i == j) ? 1.0 : 0.0;
      BN_UNROLL for (int k = i; k < D; ++k) s += U[k][i] * V[k][j];
      W[i][j] = s;
      W[j][i] = s;
    }
  chol_spd<D>(W);
  // Z = Wc^-1 U', Y = Wc^-1 V'
  double Zd[D][D], Yd[D][D];
  BN_UNROLL for (int i = 0; i < D; ++i)
    BN_UNROLL for (int j = 0; j < D; ++j) { Zd[i][j] = U[j][i]; Yd[i][j] = V[j][i]; }
  lower_solve<D, D, double>(W, Zd);
  lower_solve<D, D, double>(W, Yd);
  BN_UNROLL for (int i = 0; i < D; ++i)
    BN_UNROLL for (int j = 0; j < D; ++j) { Z[i][j] = (T)Zd[i][j]; Y[i][j] = (T)Yd[i][j]; }
}

// Sum of a[t], a[t + nt], a[t + 2 nt], ... below n: eight independent loads and partial
// sums per round, so a thread striding over a long series waits one memory latency per
// eight elements rather than per element.  For n <= t + nt the result is a[t] exactly.
BN_HD double strided_sum(const double* __restrict__ a, int32_t n, int32_t t, int32_t nt) {
  double acc[8];
  BN_UNROLL for (int u = 0; u < 8; ++u) acc[u] = 0.0;
  int32_t i = t;
  for (; i + 7 * nt < n; i += 8 * nt)
    BN_UNROLL for (int u = 0; u < 8; ++u) acc[u] += a[i + u * nt];
  for (; i < n; i += nt) acc[0] += a[i];
  return ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
}

// Bitwise or of the same strided elements of a flag array, eight loads per round.
BN_HD int32_t strided_or(const int32_t* __restrict__ a, int32_t n, int32_t t, int32_t nt) {
  int32_t acc[8];
  BN_UNROLL for (int u = 0; u < 8; ++u) acc[u] = 0;
  int32_t i = t;
  for (; i + 7 * nt < n; i += 8 * nt)
    BN_UNROLL for (int u = 0; u < 8; ++u) acc[u] |= a[i + u * nt];
  for (; i < n; i += nt) acc[0] |= a[i];
  return (acc[0] | acc[1] | acc[2] | acc[3]) | (acc[4] | acc[5] | acc[6] | acc[7]);
}

// Scheduling fence between the phases of a long unrolled body: keeps the compiler from
// hoisting a later phase's loads above it (which lengthens their live ranges).
BN_HD void sched_fence() {
#if defined(__HIP_DEVICE_COMPILE__)
  __builtin_amdgcn_sched_barrier(0);
#endif
}

// The earlier operand of fcombine_from held in registers.  An operand type gives the
// fields of an element by name: sig and jm as full symmetric matrices.
template <int D, class T>
struct RegOperand {
  const FElem<D, T>& e;
  BN_HD void phi(T P[D][D]) const { BN_UNROLL for (int i = 0; i < D; ++i) BN_UNROLL for (int j = 0; j < D; ++j) P[i][j] = e.Phi[i][j]; }
  BN_HD void beta(T b[D]) const { BN_UNROLL for (int i = 0; i < D; ++i) b[i] = e.beta[i]; }
  BN_HD void sig(T S[D][D]) const { BN_UNROLL for (int i = 0; i < D; ++i) BN_UNROLL for (int j = 0; j < D; ++j) S[i][j] = e.Sig[i][j]; }
  BN_HD void jm(T J[D][D]) const { BN_UNROLL for (int i = 0; i < D; ++i) BN_UNROLL for (int j = 0; j < D; ++j) J[i][j] = e.J[i][j]; }
  BN_HD void eta(T v[D]) const { BN_UNROLL for (int i = 0; i < D; ++i) v[i] = e.eta[i]; }
};

// e = e1 (earlier) followed by e2 (later): parallel_filtering_operator, ops.py:203-219.
// Three phases, each reading only the fields of e1 it needs (through the operand type
// L), so a caller that streams e1 from memory holds about half an element less:
//   1. Z, Y from Sig1 and J2; w = beta1 + Sig1 eta2; q = eta2 - J2 beta1;
//   2. from Phi1, J1, eta1: T Phi1, Tw, eta = Phi1' Ttq + eta1 and
//      J = Phi1' J2 Phi1 - (Y Phi1)'(Y Phi1) + J1 (e1 and J2 are then dead);
//   3. beta = Phi2 Tw + beta2, Phi = Phi2 T Phi1, Sig = (Phi2 Z')(Phi2 Z')' + Sig2.
// out must not alias e2.
template <int D, class T, class L>
BN_HD void fcombine_from(const L& e1, const FElem<D, T>& e2, FElem<D, T>& out) {
  T Z[D][D], Y[D][D], w[D], q[D];
  {
    T Sig1[D][D], beta1[D];
    e1.sig(Sig1);
    e1.beta(beta1);
    combine_factor<D, T>(Sig1, e2.J, Z, Y);
    BN_UNROLL for (int i = 0; i < D; ++i) {
      T s = beta1[i];
      BN_UNROLL for (int k = 0; k < D; ++k) s += Sig1[i][k] * e2.eta[k];
      w[i] = s;
      T u = e2.eta[i];
      BN_UNROLL for (int k = 0; k < D; ++k) u -= e2.J[i][k] * beta1[k];
      q[i] = u;
    }
  }
  sched_fence();
  T TP[D][D], Tw[D];
  {
    T Phi1[D][D];
    e1.phi(Phi1);
    // T Phi1 = Phi1 - Z'(Y Phi1)
    T YP[D][D];
    mat_mul<D, T>(Y, Phi1, YP);
    BN_UNROLL for (int i = 0; i < D; ++i)
      BN_UNROLL for (int j = 0; j < D; ++j) {
        T s = Phi1[i][j];
        BN_UNROLL for (int k = 0; k < D; ++k) s -= Z[k][i] * YP[k][j];
        TP[i][j] = s;
      }
    // Tw = w - Z'(Y w)
    T yw[D];
    BN_UNROLL for (int i = 0; i < D; ++i) {
      T s = T(0);
      BN_UNROLL for (int k = 0; k < D; ++k) s += Y[i][k] * w[k];
      yw[i] = s;
    }
    BN_UNROLL for (int i = 0; i < D; ++i) {
      T s = w[i];
      BN_UNROLL for (int k = 0; k < D; ++k) s -= Z[k][i] * yw[k];
      Tw[i] = s;
    }
    // Ttq = q - Y'(Z q) ; eta = Phi1' Ttq + eta1
    T zq[D], Ttq[D], eta1[D];
    BN_UNROLL for (int i = 0; i < D; ++i) {
      T s = T(0);
      BN_UNROLL for (int k = 0; k < D; ++k) s += Z[i][k] * q[k];
      zq[i] = s;
    }
    BN_UNROLL for (int i = 0; i < D; ++i) {
      T s = q[i];
      BN_UNROLL for (int k = 0; k < D; ++k) s -= Y[k][i] * zq[k];
      Ttq[i] = s;
    }
    e1.eta(eta1);
    BN_UNROLL for (int i = 0; i < D; ++i) {
      T s = eta1[i];
      BN_UNROLL for (int k = 0; k < D; ++k) s += Phi1[k][i] * Ttq[k];
      out.eta[i] = s;
    }
    // J = Phi1' J2 Phi1 - (Y Phi1)'(Y Phi1) + J1
    T JP[D][D], J1[D][D];
    mat_mul<D, T>(e2.J, Phi1, JP);
    e1.jm(J1);
    BN_UNROLL for (int i = 0; i < D; ++i)
      BN_UNROLL for (int j = i; j < D; ++j) {
        T s = J1[i][j];
        BN_UNROLL for (int k = 0; k < D; ++k) s += Phi1[k][i] * JP[k][j] - YP[k][i] * YP[k][j];
        out.J[i][j] = s;
        out.J[j][i] = s;
      }
  }
  sched_fence();
  // beta = Phi2 Tw + beta2 ; Phi = Phi2 TP
  BN_UNROLL for (int i = 0; i < D; ++i) {
    T s = e2.beta[i];
    BN_UNROLL for (int k = 0; k < D; ++k) s += e2.Phi[i][k] * Tw[k];
    out.beta[i] = s;
  }
  mat_mul<D, T>(e2.Phi, TP, out.Phi);
  // R = Phi2 Z' ; Sig = R R' + Sig2
  T R[D][D];
  BN_UNROLL for (int i = 0; i < D; ++i)
    BN_UNROLL for (int j = 0; j < D; ++j) {
      T s = T(0);
      BN_UNROLL for (int k = 0; k < D; ++k) s += e2.Phi[i][k] * Z[j][k];
      R[i][j] = s;
    }
  BN_UNROLL for (int i = 0; i < D; ++i)
    BN_UNROLL for (int j = i; j < D; ++j) {
      T s = e2.Sig[i][j];
      BN_UNROLL for (int k = 0; k < D; ++k) s += R[i][k] * R[j][k];
      out.Sig[i][j] = s;
      out.Sig[j][i] = s;
    }
}

template <int D, class T>
BN_HD void fcombine(const FElem<D, T>& e1, const FElem<D, T>& e2, FElem<D, T>& out) {
  fcombine_from<D, T>(RegOperand<D, T>{e1}, e2, out);
}

// (m, P) <- state (m, P) followed by element e: the combine with Phi1 = 0.
template <int D, class T>
BN_HD void fapply(T m[D], T P[D][D], const FElem<D, T>& e) {
  T Z[D][D], Y[D][D];
  combine_factor<D, T>(P, e.J, Z, Y);
  T w[D], yw[D], Tw[D];
  BN_UNROLL for (int i = 0; i < D; ++i) {
    T s = m[i];
    BN_UNROLL for (int k = 0; k < D; ++k) s += P[i][k] * e.eta[k];
    w[i] = s;
  }
  BN_UNROLL for (int i = 0; i < D; ++i) {
    T s = T(0);
    BN_UNROLL for (int k = 0; k < D; ++k) s += Y[i][k] * w[k];
    yw[i] = s;
  }
  BN_UNROLL for (int i = 0; i < D; ++i) {
    T s = w[i];
    BN_UNROLL for (int k = 0; k < D; ++k) s -= Z[k][i] * yw[k];
    Tw[i] = s;
  }
  T R[D][D];
  BN_UNROLL for (int i = 0; i < D; ++i)
    BN_UNROLL for (int j = 0; j < D; ++j) {
      T s = T(0);
      BN_UNROLL for (int k = 0; k < D; ++k) s += e.Phi[i][k] * Z[j][k];
      R[i][j] = s;
    }
  BN_UNROLL for (int i = 0; i < D; ++i) {
    T s = e.beta[i];
    BN_UNROLL for (int k = 0; k < D; ++k) s += e.Phi[i][k] * Tw[k];
    m[i] = s;
  }
  BN_UNROLL for (int i = 0; i < D; ++i)
    BN_UNROLL for (int j = i; j < D; ++j) {
      T s = e.Sig[i][j];
      BN_UNROLL for (int k = 0; k < D; ++k) s += R[i][k] * R[j][k];
      P[i][j] = s;
      P[j][i] = s;
    }
}

// ---------------------------------------------------------------------------
// Associative smoothing element (ops.py:318-335):
//   x_s | x_{e+1}, y_{1:N} ~ N(E x_{e+1} + g, L).   A state has E = 0.
// ---------------------------------------------------------------------------
template <int D, class T>
struct SElem {
  T E[D][D];
  T g[D];
  T L[D][D];
  static constexpr int NF = D * D + D + D * (D + 1) / 2;

  BN_HD void set_identity_elem() {
    set_identity<D>(E);
    set_zero<D>(L);
    BN_UNROLL for (int i = 0; i < D; ++i) g[i] = T(0);
  }
  template <class Store>
  BN_HD void store(Store st) const {
    int f = 0;
    BN_UNROLL for (int i = 0; i < D; ++i) BN_UNROLL for (int j = 0; j < D; ++j) st(f++, E[i][j]);
    BN_UNROLL for (int i = 0; i < D; ++i) st(f++, g[i]);
    BN_UNROLL for (int i = 0; i < D; ++i) BN_UNROLL for (int j = i; j < D; ++j) st(f++, L[i][j]);
  }
  template <class Load>
  BN_HD void load(Load ld) {
    int f = 0;
    BN_UNROLL for (int i = 0; i < D; ++i) BN_UNROLL for (int j = 0; j < D; ++j) E[i][j] = ld(f++);
    BN_UNROLL for (int i = 0; i < D; ++i) g[i] = ld(f++);
    BN_UNROLL for (int i = 0; i < D; ++i) BN_UNROLL for (int j = i; j < D; ++j) { T v = ld(f++); L[i][j] = v; L[j][i] = v; }
  }
};

// out = a (earlier in time) composed with b (later): parallel_smoothing_operator
// (ops.py:328-335) with the reference's elem1 = b, elem2 = a.
template <int D, class T>
BN_HD void scombine(const SElem<D, T>& a, const SElem<D, T>& b, SElem<D, T>& out) {
  T EL[D][D];
  mat_mul<D, T>(a.E, b.L, EL);
  BN_UNROLL for (int i = 0; i < D; ++i)
    BN_UNROLL for (int j = i; j < D; ++j) {
      T s = a.L[i][j];
      BN_UNROLL for (int k = 0; k < D; ++k) s += EL[i][k] * a.E[j][k];
      out.L[i][j] = s;
      out.L[j][i] = s;
    }
  BN_UNROLL for (int i = 0; i < D; ++i) {
    T s = a.g[i];
    BN_UNROLL for (int k = 0; k < D; ++k) s += a.E[i][k] * b.g[k];
    out.g[i] = s;
  }
  mat_mul<D, T>(a.E, b.E, out.E);
}

// (m, P) <- element a applied to the later state (m, P).
template <int D, class T>
BN_HD void sapply(const SElem<D, T>& a, T m[D], T P[D][D]) {
  T EL[D][D], nm[D];
  mat_mul<D, T>(a.E, P, EL);
  BN_UNROLL for (int i = 0; i < D; ++i) {
    T s = a.g[i];
    BN_UNROLL for (int k = 0; k < D; ++k) s += a.E[i][k] * m[k];
    nm[i] = s;
  }
  BN_UNROLL for (int i = 0; i < D; ++i)
    BN_UNROLL for (int j = i; j < D; ++j) {
      T s = a.L[i][j];
      BN_UNROLL for (int k = 0; k < D; ++k) s += EL[i][k] * a.E[j][k];
      P[i][j] = s;
      P[j][i] = s;
    }
  BN_UNROLL for (int i = 0; i < D; ++i) m[i] = nm[i];
}

// ---------------------------------------------------------------------------
// One step of the filter / smoother on the block-structured model
// ---------------------------------------------------------------------------

// mp = A m ; Pp = A P A' + Q   (ops.py:160-161)
template <int K1, int K2, class T>
BN_HD void predict(const T A[Spec<K1, K2>::D][Spec<K1, K2>::D], const T Q[Spec<K1, K2>::D][Spec<K1, K2>::D],
                   const T m[Spec<K1, K2>::D], const T P[Spec<K1, K2>::D][Spec<K1, K2>::D],
                   T mp[Spec<K1, K2>::D], T Pp[Spec<K1, K2>::D][Spec<K1, K2>::D]) {
  using S = Spec<K1, K2>;
  constexpr int D = S::D;
  T AP[D][D];
  BN_UNROLL for (int i = 0; i < D; ++i) {
    T s = T(0);
    BN_UNROLL for (int k = 0; k < D; ++k) if (S::nz(i, k)) s += A[i][k] * m[k];
    mp[i] = s;
  }
  BN_UNROLL for (int i = 0; i < D; ++i)
    BN_UNROLL for (int j = 0; j < D; ++j) {
      T s = T(0);
      BN_UNROLL for (int k = 0; k < D; ++k) if (S::nz(i, k)) s += A[i][k] * P[k][j];
      AP[i][j] = s;
    }
  BN_UNROLL for (int i = 0; i < D; ++i)
    BN_UNROLL for (int j = i; j < D; ++j) {
      T s = S::nz(i, j) ? Q[i][j] : T(0);
      BN_UNROLL for (int k = 0; k < D; ++k) if (S::nz(j, k)) s += AP[i][k] * A[j][k];
      Pp[i][j] = s;
      Pp[j][i] = s;
    }
}

// Measurement update with a scalar observation y, variance r (ops.py:163-171).
// Returns the innovation log-density log N(y; H mp, S) in FP64 (utils.py:376-396);
// the caller drops it for masked steps.
template <int K1, int K2, class T>
BN_HD double update(const T mp[Spec<K1, K2>::D], const T Pp[Spec<K1, K2>::D][Spec<K1, K2>::D], T y, T r,
                    T m[Spec<K1, K2>::D], T P[Spec<K1, K2>::D][Spec<K1, K2>::D]) {
  using S = Spec<K1, K2>;
  constexpr int D = S::D;
  T hP[D];
  BN_UNROLL for (int j = 0; j < D; ++j) {
    T s = T(0);
    BN_UNROLL for (int i = 0; i < D; ++i) if (S::h(i)) s += Pp[i][j];
    hP[j] = s;
  }
  double Sd = (double)r;
  T hm = T(0);
  BN_UNROLL for (int j = 0; j < D; ++j) if (S::h(j)) { Sd += (double)hP[j]; hm += mp[j]; }
  double v = (double)y - (double)hm;
  double iS = 1.0 / Sd;
  T vS = (T)(v * iS);
  T iSt = (T)iS;
  BN_UNROLL for (int i = 0; i < D; ++i) m[i] = mp[i] + hP[i] * vS;
  BN_UNROLL for (int i = 0; i < D; ++i)
    BN_UNROLL for (int j = i; j < D; ++j) {
      T s = Pp[i][j] - hP[i] * hP[j] * iSt;
      P[i][j] = s;
      P[j][i] = s;
    }
  return -0.5 * (v * v * iS + kLog2Pi + log(Sd));
}

// RTS gain G = fP A' pP^-1 from AfP = A fP and pP (ops.py:294-298), with the
// Cholesky of pP in FP64.  Writes G; returns false if pP is not SPD.
template <int D, class T>
BN_HD bool rts_gain(const T AfP[D][D], const T pP[D][D], T G[D][D]) {
  double Lc[D][D];
  BN_UNROLL for (int i = 0; i < D; ++i) BN_UNROLL for (int j = 0; j < D; ++j) Lc[i][j] = (double)pP[i][j];
  bool ok = chol_spd<D>(Lc);
  double X[D][D];  // X = pP^-1 AfP = G'
  BN_UNROLL for (int i = 0; i < D; ++i) BN_UNROLL for (int j = 0; j < D; ++j) X[i][j] = (double)AfP[i][j];
  lower_solve<D, D, double>(Lc, X);
  upper_t_solve<D, D, double>(Lc, X);
  BN_UNROLL for (int i = 0; i < D; ++i) BN_UNROLL for (int j = 0; j < D; ++j) G[i][j] = (T)X[j][i];
  return ok;
}

}  // namespace bn
