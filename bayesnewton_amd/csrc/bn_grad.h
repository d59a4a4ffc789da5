// Derivative of the filter log-likelihood log Z = sum_k log N(y_k; H m-_k, S_k)
// (ops.py:154-180, compute_log_lik basemodels.py:726-741) with respect to the
// kernel hyperparameters (lengthscale, variance of every component), the
// gradient objax.GradValues(model.energy) takes through the filter with the
// sites and posterior held fixed (energy = ... - log Z, inference.py:197-222).
//
// Forward mode inside the chunked scan: tangents (dm, dP) obey a recursion
// that is affine given the primal trajectory,
//   z_k = J_k z_{k-1} + s_k^(theta),   dlogZ_k = g_k . z_{k-1} + c_k^(theta),
// with one linear part J_k for every hyperparameter.  A lane propagates, over
// its chunk, the NZ = D + D(D+1)/2 unit tangents (the chunk's linear map M and
// functional f) and the NH hyperparameter tangents started at zero (sources o,
// constants c).  Applying the chunks in order, z <- M_j z + o_j, grad += f_j . z + c_j,
// is a composition of affine maps, which bn_markov.hip evaluates as an associative
// reduction (k_grad_group, k_grad_total).  Nothing is stored per time step.
#pragma once
#include "bn_lanes.h"

namespace bn {

// lam^p for a small integer p (|p| < 2K), ilam = 1 / lam: a product chain (p is a
// constant once the callers' loops unroll) instead of a pow call per matrix entry.
BN_HD double int_pow(double lam, double ilam, int p) {
  const double f = p < 0 ? ilam : lam;
  const int n = p < 0 ? -p : p;
  double r = 1.0;
  for (int q = 0; q < n; ++q) r *= f;
  return r;
}

// dA/dlambda of one component: A_ij = lam^(i-j) e^{-x} a_ij(x), x = lam dt, so
//   dA_ij/dlam = (i-j) lam^(i-j-1) e^{-x} a_ij + lam^(i-j) e^{-x} dt (a_ij' - a_ij)
// e = exp(-x) from the caller.
template <int K>
BN_HD void comp_dA_dlam(double lam, double ilam, double dt, double e, double* dA) {
  const double x = lam * dt;
  double a[9], da[9];
  if (K == 1) {
    a[0] = 1.0; da[0] = 0.0;
  } else if (K == 2) {
    a[0] = 1.0 + x; a[1] = x; a[2] = -x; a[3] = 1.0 - x;
    da[0] = 1.0; da[1] = 1.0; da[2] = -1.0; da[3] = -1.0;
  } else {
    a[0] = 1.0 + x + 0.5 * x * x; a[1] = x + x * x;     a[2] = 0.5 * x * x;
    a[3] = -0.5 * x * x;          a[4] = 1.0 + x - x * x; a[5] = x - 0.5 * x * x;
    a[6] = 0.5 * x * x - x;       a[7] = x * x - 3.0 * x; a[8] = 1.0 - 2.0 * x + 0.5 * x * x;
    da[0] = 1.0 + x; da[1] = 1.0 + 2.0 * x; da[2] = x;
    da[3] = -x;      da[4] = 1.0 - 2.0 * x; da[5] = 1.0 - x;
    da[6] = x - 1.0; da[7] = 2.0 * x - 3.0; da[8] = -2.0 + x;
  }
  BN_UNROLL for (int i = 0; i < K; ++i)
    BN_UNROLL for (int j = 0; j < K; ++j) {
      const int p = i - j;
      const double lp = int_pow(lam, ilam, p);
      dA[i * K + j] = (p != 0 ? p * lp * ilam * e * a[i * K + j] : 0.0) + lp * e * dt * (da[i * K + j] - a[i * K + j]);
    }
}

// dQ/dlambda of one component (packed upper, K(K+1)/2 entries) from
// Q_ij = var lam^(i+j) F_ij(x):  dQ_ij/dlam = var (i+j) lam^(i+j-1) F + var lam^(i+j) dt F'(x),
// with F = x^m G(x) (series, x < X0) or C - e^{-2x} P(x) (above), as in
// Model::comp_block.  Also returns Q itself.  e = exp(-x) from the caller.
template <int K>
BN_HD void comp_dQ_dlam(double lam, double ilam, double var, double dt, double e, double* Qp, double* dQp) {
  const double x = lam * dt;
  const bool small = x < kQX0;
  const double e2 = e * e;
  int idx = 0;
  BN_UNROLL for (int i = 0; i < K; ++i)
    BN_UNROLL for (int j = i; j < K; ++j) {
      // both forms from wave-uniform coefficient rows, selected per lane (as in
      // Model::comp_block): no per-lane table choice, no dependent vector loads
      const double* ts = qtab<K>(idx, true);
      const double* tp = qtab<K>(idx, false);
      constexpr int WS = qser_width<K>(), WP = qpolydeg<K>();
      double vs = ts[WS - 1], dvs = 0.0;
      for (int n = WS - 2; n >= 0; --n) {
        dvs = dvs * x + vs;
        vs = vs * x + ts[n];
      }
      double vp = tp[WP - 1], dvp = 0.0;
      for (int n = WP - 2; n >= 0; --n) {
        dvp = dvp * x + vp;
        vp = vp * x + tp[n];
      }
      const int m = 2 * K - 1 - i - j;
      double xm1 = 1.0;
      for (int q = 0; q < m - 1; ++q) xm1 *= x;
      const double F = small ? xm1 * x * vs : qc<K>(idx) - e2 * vp;
      const double dF = small ? xm1 * (m * vs + x * dvs) : e2 * (2.0 * vp - dvp);
      const double lp = int_pow(lam, ilam, i + j);
      Qp[idx] = var * lp * F;
      dQp[idx] = var * ((i + j) * lp * ilam * F + lp * dt * dF);
      ++idx;
    }
}

// One linearised filter step applied to a tangent (dm, dP), with the primal
// quantities of the step: A, m (previous filtered mean), P (previous filtered
// covariance), mp = A m, Pp = A P A' + Q, s, v, b = Pp h.  Source terms
// (dA, dQ) are added when `src` is set; at the first step of a series the
// predicted moments are (0, Pinf) and the tangent is (0, dPinf).
// Returns the step's d logZ; masked steps contribute 0 but still update.
template <int K1, int K2>
BN_HD double tangent_step(const double A[Spec<K1, K2>::D][Spec<K1, K2>::D], const double m[Spec<K1, K2>::D],
                          const double P[Spec<K1, K2>::D][Spec<K1, K2>::D], const double b[Spec<K1, K2>::D],
                          double s, double v, bool masked, bool first, const double* dA, const double* dQ,
                          const double* dPinf, double* z) {
  using S = Spec<K1, K2>;
  constexpr int D = S::D;
  double dm[D], dP[D][D];
  // unpack z = (dm, packed upper dP)
  {
    int f = D;
    BN_UNROLL for (int i = 0; i < D; ++i) dm[i] = z[i];
    BN_UNROLL for (int i = 0; i < D; ++i)
      BN_UNROLL for (int j = i; j < D; ++j) { dP[i][j] = z[f]; dP[j][i] = z[f]; ++f; }
  }
  double dmp[D], dPp[D][D];
  if (first) {
    BN_UNROLL for (int i = 0; i < D; ++i) {
      dmp[i] = 0.0;
      BN_UNROLL for (int j = 0; j < D; ++j) dPp[i][j] = dPinf ? dPinf[i * D + j] : 0.0;
    }
  } else {
    double AdP[D][D];
    BN_UNROLL for (int i = 0; i < D; ++i) {
      double t = 0.0;
      BN_UNROLL for (int k = 0; k < D; ++k) if (S::nz(i, k)) t += A[i][k] * dm[k];
      if (dA) BN_UNROLL for (int k = 0; k < D; ++k) if (S::nz(i, k)) t += dA[i * D + k] * m[k];
      dmp[i] = t;
      BN_UNROLL for (int c = 0; c < D; ++c) {
        double u = 0.0;
        BN_UNROLL for (int k = 0; k < D; ++k) if (S::nz(i, k)) u += A[i][k] * dP[k][c];
        AdP[i][c] = u;
      }
    }
    // dPp = A dP A' + (dA P A' + A P dA') + dQ; X = dA P A' so the middle term is X + X'
    double X[D][D];
    if (dA) {
      double dAP[D][D];
      BN_UNROLL for (int i = 0; i < D; ++i)
        BN_UNROLL for (int c = 0; c < D; ++c) {
          double u = 0.0;
          BN_UNROLL for (int k = 0; k < D; ++k) if (S::nz(i, k)) u += dA[i * D + k] * P[k][c];
          dAP[i][c] = u;
        }
      BN_UNROLL for (int i = 0; i < D; ++i)
        BN_UNROLL for (int c = 0; c < D; ++c) {
          double u = 0.0;
          BN_UNROLL for (int k = 0; k < D; ++k) if (S::nz(c, k)) u += dAP[i][k] * A[c][k];
          X[i][c] = u;
        }
    }
    BN_UNROLL for (int i = 0; i < D; ++i)
      BN_UNROLL for (int c = i; c < D; ++c) {
        double t = dQ ? dQ[i * D + c] : 0.0;
        BN_UNROLL for (int k = 0; k < D; ++k) if (S::nz(c, k)) t += AdP[i][k] * A[c][k];
        if (dA) t += X[i][c] + X[c][i];
        dPp[i][c] = t;
        dPp[c][i] = t;
      }
  }
  // ds = h' dPp h, dv = -h' dmp, db = dPp h
  double db[D];
  double ds = 0.0, dv = 0.0;
  BN_UNROLL for (int i = 0; i < D; ++i) {
    double t = 0.0;
    BN_UNROLL for (int q = 0; q < D; ++q) if (S::h(q)) t += dPp[i][q];
    db[i] = t;
  }
  BN_UNROLL for (int q = 0; q < D; ++q) if (S::h(q)) { ds += db[q]; dv -= dmp[q]; }
  const double is = 1.0 / s;
  BN_UNROLL for (int i = 0; i < D; ++i) z[i] = dmp[i] + (db[i] * v + b[i] * dv) * is - b[i] * v * ds * is * is;
  {
    int f = D;
    BN_UNROLL for (int i = 0; i < D; ++i)
      BN_UNROLL for (int j = i; j < D; ++j) {
        z[f++] = dPp[i][j] - (db[i] * b[j] + b[i] * db[j]) * is + b[i] * b[j] * ds * is * is;
      }
  }
  if (masked) return 0.0;
  return -v * dv * is + 0.5 * (v * v * is - 1.0) * ds * is;
}

}  // namespace bn

namespace bn {

template <int K1, int K2>
struct GradDims {
  static constexpr int D = K1 + K2;
  static constexpr int NZ = D + D * (D + 1) / 2;  // tangent size (dm, packed dP)
  static constexpr int NH = K2 > 0 ? 4 : 2;       // (lengthscale, variance) per component
  static constexpr int NF = NZ * NZ + NH * NZ + NZ + NH;  // stored doubles per chunk
};

struct GradOut {
  double* agg;     // field-major per chunk: M (NZ*NZ, column-major), o (NH*NZ), f (NZ), c (NH)
  int64_t stride;  // n_chunks_pad
  double* ell;     // per-chunk log-likelihood partial (primal)
};

// A, Q of one component block at one step and, for kind 0 (lengthscale) or 1 (variance),
// that hyperparameter's source terms dA, dQ, dPinf on the block (full D x D arrays,
// row-major; kind -1: none).  lam = c / ls, so d/dls = (-lam / ls) d/dlam; Pinf and Q are
// proportional to var, A does not depend on it.
template <int K, int O, int D>
BN_HD void comp_step_col(double lam, double var, double ls, double dt, const double* Pinf_full, int kind,
                         double A[D][D], double Q[D][D], double* dA, double* dQ, double* dPinf) {
  double a[9], q[6], dq[6];
  const double e = exp(-(lam * dt)), ilam = 1.0 / lam;
  comp_transition_e<K>(lam, dt, e, a);
  comp_dQ_dlam<K>(lam, ilam, var, dt, e, q, dq);
  const double dlam_dls = -lam / ls;
  int idx = 0;
  for (int i = 0; i < K; ++i)
    for (int j = i; j < K; ++j) {
      Q[O + i][O + j] = Q[O + j][O + i] = q[idx];
      if (kind >= 0) {
        const double v = kind == 0 ? dq[idx] * dlam_dls : q[idx] / var;
        dQ[(O + i) * D + O + j] = dQ[(O + j) * D + O + i] = v;
      }
      ++idx;
    }
  double da[9];
  if (kind == 0) comp_dA_dlam<K>(lam, ilam, dt, e, da);
  for (int i = 0; i < K; ++i)
    for (int j = 0; j < K; ++j) {
      A[O + i][O + j] = a[i * K + j];
      if (kind >= 0) {
        const double p = Pinf_full[(O + i) * D + O + j];
        dA[(O + i) * D + O + j] = kind == 0 ? da[i * K + j] * dlam_dls : 0.0;
        dPinf[(O + i) * D + O + j] = kind == 0 ? (i + j) * p * ilam * dlam_dls : p / var;
      }
    }
}

// One lane: column `col` of chunk j's tangent aggregate (see the header comment) -- a
// unit tangent (col < NZ: column col of M and entry col of f) or a hyperparameter
// tangent started at zero (h = col - NZ: column h of O and entry h of c).  The columns
// of a chunk are independent given the primal trajectory, which every lane recomputes.
// HK is the column's class, fixed at compile time so that each instance carries only its
// own source terms: -1 a unit tangent (no sources), 0 a lengthscale, 1 a variance
// (h = 2 comp + HK).
template <int K1, int K2, int HK>
BN_HD void lane_grad_col(const PlanView& pl, const HyperView& hv, const float* __restrict__ dt,
                         const float* __restrict__ y, const float* __restrict__ r, const uint8_t* __restrict__ mask,
                         CSoA fagg, CSoA fcarry, const GradOut& out, int32_t j, int col) {
  using S = Spec<K1, K2>;
  using GD = GradDims<K1, K2>;
  constexpr int D = S::D, NZ = GD::NZ, NH = GD::NH;
  constexpr int FF = NZ * NZ + NH * NZ, FC = FF + NZ;
  const int64_t st = out.stride;
  const int32_t b = pl.chunk_series[j];
  const int32_t L = pl.chunk_len_[j];
  const bool first_chunk = (pl.chunk_flags[j] & CH_FIRST) != 0;
  const int h = HK < 0 ? -1 : col - NZ;
  if (HK < 0 && first_chunk) {  // a series' first chunk has no incoming tangent: M = 0, f = 0
    for (int q = 0; q < NZ; ++q) out.agg[(int64_t)(col * NZ + q) * st + j] = 0.0;
    out.agg[(int64_t)(FF + col) * st + j] = 0.0;
    return;
  }
  const int comp = HK >= 0 ? (K2 > 0 ? h / 2 : 0) : -1;
  Model<K1, K2> mdl = load_model<K1, K2>(hv, b);
  const int64_t ho = (int64_t)b * hv.stride;
  const double ls1 = hv.ls[ho], var1 = hv.var[ho];
  const double ls2 = K2 > 0 ? hv.ls[ho + 1] : 1.0, var2 = K2 > 0 ? hv.var[ho + 1] : 1.0;
  double Pinf_full[D * D];
  for (int i = 0; i < D; ++i)
    for (int c = 0; c < D; ++c) Pinf_full[i * D + c] = mdl.Pinf[i][c];
  double m[D], P[D][D];
  if (!first_chunk) {
    filter_incoming<D, double>(pl, fagg, fcarry, j, m, P);
  } else {
    for (int i = 0; i < D; ++i) { m[i] = 0.0; for (int c = 0; c < D; ++c) P[i][c] = 0.0; }
  }
  double z[NZ];
  for (int q = 0; q < NZ; ++q) z[q] = (q == col) ? 1.0 : 0.0;
  double acc = 0.0, ell = 0.0;
  for (int32_t k = 0; k < L; ++k) {
    const int64_t s = pl.slot(j, k);
    const bool first = (k == 0) && first_chunk;
    double A[D][D], Q[D][D], dA[D * D], dQ[D * D], dPinf[D * D];
    for (int i = 0; i < D; ++i)
      for (int c = 0; c < D; ++c) { A[i][c] = 0.0; Q[i][c] = 0.0; }
    for (int q = 0; q < D * D; ++q) { dA[q] = 0.0; dQ[q] = 0.0; dPinf[q] = 0.0; }
    const double dts = first ? 0.0 : (double)dt[s];
    comp_step_col<K1, 0, D>(mdl.lam1, var1, ls1, dts, Pinf_full, comp == 0 ? HK : -1, A, Q, dA, dQ, dPinf);
    if (K2 > 0)
      comp_step_col<(K2 > 0 ? K2 : 1), K1, D>(mdl.lam2, var2, ls2, dts, Pinf_full, comp == 1 ? HK : -1, A, Q, dA,
                                              dQ, dPinf);
    // primal prediction
    double mp[D], Pp[D][D];
    if (first) {
      for (int i = 0; i < D; ++i) { mp[i] = 0.0; for (int c = 0; c < D; ++c) Pp[i][c] = mdl.Pinf[i][c]; }
    } else {
      double AP[D][D];
      for (int i = 0; i < D; ++i) {
        double t = 0.0;
        for (int q = 0; q < D; ++q) t += A[i][q] * m[q];
        mp[i] = t;
        for (int c = 0; c < D; ++c) {
          double u = 0.0;
          for (int q = 0; q < D; ++q) u += A[i][q] * P[q][c];
          AP[i][c] = u;
        }
      }
      for (int i = 0; i < D; ++i)
        for (int c = i; c < D; ++c) {
          double t = Q[i][c];
          for (int q = 0; q < D; ++q) t += AP[i][q] * A[c][q];
          Pp[i][c] = Pp[c][i] = t;
        }
    }
    double bb[D];
    double sd = (double)r[s], hm = 0.0;
    for (int i = 0; i < D; ++i) {
      double t = 0.0;
      for (int q = 0; q < D; ++q) if (S::h(q)) t += Pp[i][q];
      bb[i] = t;
    }
    for (int q = 0; q < D; ++q) if (S::h(q)) { sd += bb[q]; hm += mp[q]; }
    const double v = (double)y[s] - hm;
    const bool masked = mask != nullptr && mask[s];
    if (h == 0 && !masked) ell += -0.5 * (v * v / sd + kLog2Pi + log(sd));  // one column stores it
    acc += tangent_step<K1, K2>(A, m, P, bb, sd, v, masked, first, HK >= 0 ? dA : nullptr,
                                HK >= 0 ? dQ : nullptr, HK >= 0 ? dPinf : nullptr, z);
    // primal update
    for (int i = 0; i < D; ++i) m[i] = mp[i] + bb[i] * v / sd;
    for (int i = 0; i < D; ++i)
      for (int c = i; c < D; ++c) P[i][c] = P[c][i] = Pp[i][c] - bb[i] * bb[c] / sd;
  }
  if (HK < 0) {
    for (int q = 0; q < NZ; ++q) out.agg[(int64_t)(col * NZ + q) * st + j] = z[q];
    out.agg[(int64_t)(FF + col) * st + j] = acc;
  } else {
    for (int q = 0; q < NZ; ++q) out.agg[(int64_t)(NZ * NZ + h * NZ + q) * st + j] = z[q];
    out.agg[(int64_t)(FC + h) * st + j] = acc;
    if (h == 0) out.ell[j] = ell;
  }
}

}  // namespace bn
