// Sparse Markov GP (bayesnewton/basemodels.py:928-1152): the per-transition work of
// one inference step or one energy evaluation, on gfx950.
//
// The inducing states u_0..u_{M+1} sit at the sorted inducing times with sentinels at
// -/+1e10 (z_aug[M + 2]); transition i joins (u_i, u_{i+1}) and carries a 2d-dimensional
// Gaussian site (nat1 [2d], nat2 [2d x 2d]).  The pairs filter (k_pairs_filter) and the
// RTS smoother over the M inducing states (k_smoother_pass) run first; this kernel takes
// the smoothed states sm / sP and gains G (contiguous [M, d], [M, d, d]) and, one lane per
// transition:
//   * builds the pair posterior (build_joint, utils.py:544-553) with (0, Pinf) at the
//     sentinels and cross covariance G_i P_{i+1};
//   * for each data point t_n in [z_aug[i], z_aug[i+1]) (CSR tr_off, data sorted by time)
//     forms the conditional statistics of x(t_n) given the pair (utils.py:185-215, jitter
//     1e-8) reduced to W_n = H P_n and nu_n = H T_n H', and q(f_n) = N(W_n m, W_n C W_n' + nu_n)
//     (conditional_posterior_to_data, basemodels.py:1074-1105);
//   * UPDATE: the VI / Newton site terms at q(f_n) (ensure_psd on the scalar Hessian),
//     mapped to the pair (conditional_data_to_posterior, 1107-1113), newton_update with
//     the missing-data rule (inference.py:21-39), summed per transition with the old site
//     kept where a transition has no data and 1e-8 I added (group_natural_params,
//     1115-1140), damped by lr (inference.py:83-86), then reparametrised to the site mean
//     and covariance the pairs filter reads (basemodels.py:85-90); per-transition |d nat|
//     partial sums for diff1 / diff2 (inference.py:79-81);
//   * ENERGY: per transition the sum of the data terms (VI E_q log p, Newton log p at the
//     mean; masked points skipped as nansum does) and E_q log N(site mean | pair, site cov)
//     of the KL (expected_density_pseudo, basemodels.py:214-222).
// Everything is FP64; 2d <= 10, so the pair algebra lives in registers.
#include <hip/hip_runtime.h>

#include <string>

#include "bn_internal.h"
#include "bn_sites.h"

using namespace bn;

namespace {
thread_local std::string g_sp_err;

int sp_fail(int code, const std::string& msg) {
  g_sp_err = msg;
  return set_error(code, msg);
}

constexpr int kSpBlock = 64;
constexpr double kEdge = 1e10;     // basemodels.py:946 sentinel times
constexpr double kJitter = 1e-8;   // utils.py:207 (conditional statistics), basemodels.py:1138 (grouping)

struct SparseIO {
  int32_t n_z;            // inducing states (transitions: n_z + 1)
  const double* z_aug;    // [n_z + 2]
  const double* sm;       // [n_z, d]
  const double* sP;       // [n_z, d, d]
  const double* gain;     // [n_z, d, d]
  const int64_t* tr_off;  // [n_z + 2]
  const double* t;        // [N] data times, sorted
  const double* y;        // [N] (NaN = missing)
  double* nat1;           // [n_z + 1, 2d]
  double* nat2;           // [n_z + 1, 2d, 2d]
  double* site_mean;      // [n_z + 1, 2d]
  double* site_cov;       // [n_z + 1, 2d, 2d]
  double* pair_mean;      // [n_z + 1, 2d] or nullptr
  double* pair_cov;       // [n_z + 1, 2d, 2d] or nullptr
  double* part1;          // [n_z + 1]
  double* part2;          // [n_z + 1]
};

template <int D2>
__device__ bool chol_inplace(double L[D2][D2]) {
  bool ok = true;
  for (int j = 0; j < D2; ++j) {
    double s = L[j][j];
    for (int k = 0; k < j; ++k) s -= L[j][k] * L[j][k];
    ok = ok && s > 0.0;
    const double dj = sqrt(s);
    L[j][j] = dj;
    for (int i = j + 1; i < D2; ++i) {
      double t = L[i][j];
      for (int k = 0; k < j; ++k) t -= L[i][k] * L[j][k];
      L[i][j] = t / dj;
    }
    for (int i = 0; i < j; ++i) L[i][j] = 0.0;
  }
  return ok;
}

// x := (L L')^-1 x
template <int D2>
__device__ void chol_solve_vec(const double L[D2][D2], double x[D2]) {
  for (int i = 0; i < D2; ++i) {
    double s = x[i];
    for (int k = 0; k < i; ++k) s -= L[i][k] * x[k];
    x[i] = s / L[i][i];
  }
  for (int i = D2 - 1; i >= 0; --i) {
    double s = x[i];
    for (int k = i + 1; k < D2; ++k) s -= L[k][i] * x[k];
    x[i] = s / L[i][i];
  }
}

template <int K1, int K2, int METHOD, bool UPDATE>
__global__ __launch_bounds__(kSpBlock) void k_sparse_sites(HyperView hv, LikParams L, GHTable gh, SparseIO io,
                                                           double lr, int ensure_psd) {
  using S = Spec<K1, K2>;
  constexpr int D = S::D;
  constexpr int D2 = 2 * D;
  const int32_t i = blockIdx.x * kSpBlock + threadIdx.x;
  const int32_t n_tr = io.n_z + 1;
  if (i >= n_tr) return;
  const Model<K1, K2> mdl = load_model<K1, K2>(hv, 0);
  // ---- pair posterior (u_i, u_{i+1}) ----
  double pm[D2], pc[D2][D2];
  {
    double G[D][D], Pr[D][D];
    for (int p = 0; p < D; ++p) {
      pm[p] = i == 0 ? 0.0 : io.sm[(int64_t)(i - 1) * D + p];
      pm[D + p] = i == io.n_z ? 0.0 : io.sm[(int64_t)i * D + p];
      for (int q = 0; q < D; ++q) {
        pc[p][q] = i == 0 ? mdl.Pinf[p][q] : io.sP[((int64_t)(i - 1) * D + p) * D + q];
        Pr[p][q] = i == io.n_z ? mdl.Pinf[p][q] : io.sP[((int64_t)i * D + p) * D + q];
        pc[D + p][D + q] = Pr[p][q];
        G[p][q] = i == 0 ? 0.0 : io.gain[((int64_t)(i - 1) * D + p) * D + q];
      }
    }
    for (int p = 0; p < D; ++p)
      for (int q = 0; q < D; ++q) {
        double c = 0.0;
        for (int k = 0; k < D; ++k) c = fma(G[p][k], Pr[k][q], c);
        pc[p][D + q] = c;
        pc[D + q][p] = c;
      }
  }
  if (io.pair_mean != nullptr) {
    for (int p = 0; p < D2; ++p) {
      io.pair_mean[(int64_t)i * D2 + p] = pm[p];
      for (int q = 0; q < D2; ++q) io.pair_cov[((int64_t)i * D2 + p) * D2 + q] = pc[p][q];
    }
  }
  const double zl = io.z_aug[i], zr = io.z_aug[i + 1];
  double S1[D2], S2[D2][D2];
  for (int p = 0; p < D2; ++p) {
    S1[p] = 0.0;
    for (int q = 0; q < D2; ++q) S2[p][q] = 0.0;
  }
  double e_data = 0.0;
  const int64_t n0 = io.tr_off[i], n1 = io.tr_off[i + 1];
  for (int64_t n = n0; n < n1; ++n) {
    const double tn = io.t[n];
    // conditional statistics (utils.py:185-215) reduced by H: W = h P (1 x 2d), nu = h T h'
    double Af[D][D], Qf[D][D], Ab[D][D], Qb[D][D];
    mdl.discretise(tn - zl, Af, Qf);
    mdl.discretise(zr - tn, Ab, Qb);
    double ABQ[D][D], Lq[D][D];
    for (int p = 0; p < D; ++p)
      for (int q = 0; q < D; ++q) {
        double s = 0.0;
        for (int k = 0; k < D; ++k) s = fma(Ab[p][k], Qf[k][q], s);
        ABQ[p][q] = s;
      }
    for (int p = 0; p < D; ++p)
      for (int q = 0; q < D; ++q) {
        double s = Qb[p][q] + (p == q ? kJitter : 0.0);
        for (int k = 0; k < D; ++k) s = fma(ABQ[p][k], Ab[q][k], s);
        Lq[p][q] = s;
      }
    chol_spd<D>(Lq);
    double V[D][D];
    for (int p = 0; p < D; ++p)
      for (int q = 0; q < D; ++q) V[p][q] = Ab[p][q];
    lower_solve<D, D, double>(Lq, V);
    upper_t_solve<D, D, double>(Lq, V);
    double hQ[D], hW[D];
    for (int q = 0; q < D; ++q) {
      double s = 0.0;
      for (int p = 0; p < D; ++p) if (S::h(p)) s += Qf[p][q];
      hQ[q] = s;
    }
    for (int q = 0; q < D; ++q) {
      double s = 0.0;
      for (int k = 0; k < D; ++k) s = fma(hQ[k], V[q][k], s);
      hW[q] = s;
    }
    double nu = 0.0;
    for (int p = 0; p < D; ++p) if (S::h(p)) nu += hQ[p];
    for (int p = 0; p < D; ++p) {
      double vq = 0.0, ab = 0.0;
      for (int k = 0; k < D; ++k) {
        vq = fma(V[p][k], hQ[k], vq);
        if (S::h(k)) ab += ABQ[p][k];
      }
      nu -= ab * vq;
    }
    double W[D2];
    {
      double hWAb[D];
      for (int q = 0; q < D; ++q) {
        double s = 0.0;
        for (int k = 0; k < D; ++k) s = fma(hW[k], Ab[k][q], s);
        hWAb[q] = s;
      }
      for (int q = 0; q < D; ++q) {
        double s = 0.0;
        for (int k = 0; k < D; ++k) s += ((S::h(k) ? 1.0 : 0.0) - hWAb[k]) * Af[k][q];
        W[q] = s;
        W[D + q] = hW[q];
      }
    }
    double mf = 0.0, vf = nu;
    for (int p = 0; p < D2; ++p) {
      mf = fma(W[p], pm[p], mf);
      double s = 0.0;
      for (int q = 0; q < D2; ++q) s = fma(pc[p][q], W[q], s);
      vf = fma(W[p], s, vf);
    }
    const double yn = io.y[n];
    const bool miss = yn != yn;
    if (UPDATE) {
      double dE, d2E;
      if (METHOD == M_VI) {
        double E;
        var_exp(L, gh, miss ? mf : yn, mf, vf, E, dE, d2E);
      } else {
        double ll;
        log_lik_derivs(L, miss ? mf : yn, mf, ll, dE, d2E);
      }
      if (ensure_psd && -d2E < 0.0) d2E = -1e-2;  // utils.py:89-96 on a 1 x 1 precision
      if (miss) {
        // newton_update: H := -1e-6 everywhere, J := H m, so nat1 += 0, nat2 += 1e-6
        for (int p = 0; p < D2; ++p)
          for (int q = 0; q < D2; ++q) S2[p][q] += 1e-6;
      } else {
        // nat1 += W' (dE - d2E W m) = W' dE - (W' d2E W) m ; nat2 += -d2E W' W
        const double g = dE - d2E * mf;
        for (int p = 0; p < D2; ++p) {
          S1[p] = fma(W[p], g, S1[p]);
          for (int q = 0; q < D2; ++q) S2[p][q] = fma(-d2E * W[p], W[q], S2[p][q]);
        }
      }
    } else if (!miss) {
      if (METHOD == M_VI) {
        double E, dE, d2E;
        var_exp(L, gh, yn, mf, vf, E, dE, d2E);
        e_data += E;
      } else {
        double ll, J, H;
        log_lik_derivs(L, yn, mf, ll, J, H);
        e_data += ll;
      }
    }
  }
  double* n1p = io.nat1 + (int64_t)i * D2;
  double* n2p = io.nat2 + (int64_t)i * D2 * D2;
  if (UPDATE) {
    const bool keep = n1 == n0;  // (1 - counter / max(num_neighbours, 1)) = 1 only without data
    double d1 = 0.0, d2 = 0.0;
    double nn1[D2], nn2[D2][D2];
    for (int p = 0; p < D2; ++p) {
      const double o1 = n1p[p];
      const double g1 = S1[p] + (keep ? o1 : 0.0);
      d1 += fabs(g1 - o1);
      nn1[p] = (1.0 - lr) * o1 + lr * g1;
      for (int q = 0; q < D2; ++q) {
        const double o2 = n2p[p * D2 + q];
        const double g2 = S2[p][q] + (keep ? o2 : 0.0) + (p == q ? kJitter : 0.0);
        d2 += fabs(g2 - o2);
        nn2[p][q] = (1.0 - lr) * o2 + lr * g2;
      }
    }
    io.part1[i] = d1;
    io.part2[i] = d2;
    // reparametrise: site_cov = nat2^-1, site_mean = site_cov nat1 (basemodels.py:85-90)
    double Lc[D2][D2], cov[D2][D2], mean[D2];
    for (int p = 0; p < D2; ++p) {
      n1p[p] = nn1[p];
      mean[p] = nn1[p];
      for (int q = 0; q < D2; ++q) {
        n2p[p * D2 + q] = nn2[p][q];
        Lc[p][q] = 0.5 * (nn2[p][q] + nn2[q][p]);
      }
    }
    chol_inplace<D2>(Lc);
    chol_solve_vec<D2>(Lc, mean);
    for (int c = 0; c < D2; ++c) {
      double col[D2];
      for (int p = 0; p < D2; ++p) col[p] = p == c ? 1.0 : 0.0;
      chol_solve_vec<D2>(Lc, col);
      for (int p = 0; p < D2; ++p) cov[p][c] = col[p];
    }
    for (int p = 0; p < D2; ++p) {
      io.site_mean[(int64_t)i * D2 + p] = mean[p];
      for (int q = 0; q < D2; ++q) io.site_cov[((int64_t)i * D2 + p) * D2 + q] = 0.5 * (cov[p][q] + cov[q][p]);
    }
  } else {
    // E_q log N(site_mean | f, site_cov) with q = N(pm, pc) (utils.py:510-531, no mask)
    double Lc[D2][D2], r[D2];
    for (int p = 0; p < D2; ++p) {
      r[p] = io.site_mean[(int64_t)i * D2 + p] - pm[p];
      for (int q = 0; q < D2; ++q) Lc[p][q] = io.site_cov[((int64_t)i * D2 + p) * D2 + q];
    }
    chol_inplace<D2>(Lc);
    double logdet = 0.0;
    for (int p = 0; p < D2; ++p) logdet += 2.0 * log(fabs(Lc[p][p]));
    double x[D2];
    for (int p = 0; p < D2; ++p) x[p] = r[p];
    chol_solve_vec<D2>(Lc, x);
    double dist = 0.0;
    for (int p = 0; p < D2; ++p) dist += r[p] * x[p];
    double tr = 0.0;
    for (int c = 0; c < D2; ++c) {
      double col[D2];
      for (int p = 0; p < D2; ++p) col[p] = pc[p][c];
      chol_solve_vec<D2>(Lc, col);
      tr += col[c];
    }
    io.part1[i] = e_data;
    io.part2[i] = -0.5 * (dist + D2 * kLog2Pi + logdet) - 0.5 * tr;
  }
}

int sp_state_dim(const bn_kernel_desc* k) {
  int d = 0;
  for (int c = 0; c < k->n_comp; ++c) d += k->kind[c];
  return d;
}

template <int K1, int K2, int METHOD>
int launch_sparse(HyperView hv, const LikParams& L, const GHTable& gh, const SparseIO& io, int mode, double lr,
                  int ensure_psd, hipStream_t st) {
  const dim3 grid((unsigned)((io.n_z + 1 + kSpBlock - 1) / kSpBlock));
  if (mode == BN_SPARSE_UPDATE)
    hipLaunchKernelGGL((k_sparse_sites<K1, K2, METHOD, true>), grid, dim3(kSpBlock), 0, st, hv, L, gh, io, lr,
                       ensure_psd);
  else
    hipLaunchKernelGGL((k_sparse_sites<K1, K2, METHOD, false>), grid, dim3(kSpBlock), 0, st, hv, L, gh, io, lr,
                       ensure_psd);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return sp_fail(BN_EHIP, std::string("k_sparse_sites: ") + hipGetErrorString(e));
  return BN_OK;
}

template <int K1, int K2>
int dispatch_method(int32_t method, HyperView hv, const LikParams& L, const GHTable& gh, const SparseIO& io, int mode,
                    double lr, int ensure_psd, hipStream_t st) {
  if (method == BN_VI) return launch_sparse<K1, K2, M_VI>(hv, L, gh, io, mode, lr, ensure_psd, st);
  if (method == BN_NEWTON) return launch_sparse<K1, K2, M_NEWTON>(hv, L, gh, io, mode, lr, ensure_psd, st);
  return sp_fail(BN_EUNSUPPORTED, "the sparse Markov sites support VI and Newton");
}
}  // namespace

extern "C" {

int bn_sparse_sites(const bn_kernel_desc* kern, const double* lengthscale, const double* variance,
                    const bn_lik_desc* lik, int32_t method, int32_t mode, int32_t gh_points, double lr,
                    int32_t ensure_psd, int32_t n_z, const double* z_aug, const double* sm, const double* sP,
                    const double* gain, const int64_t* tr_off, const double* t, const double* y, double* nat1,
                    double* nat2, double* site_mean, double* site_cov, double* pair_mean, double* pair_cov,
                    double* part1, double* part2, void* stream) {
  if (kern == nullptr || kern->n_comp < 1 || kern->n_comp > 2) return sp_fail(BN_EUNSUPPORTED, "kernel must have 1 or 2 Matern components");
  for (int c = 0; c < kern->n_comp; ++c)
    if (kern->kind[c] < 1 || kern->kind[c] > 3) return sp_fail(BN_EUNSUPPORTED, "component kind must be Matern12/32/52");
  if (mode != BN_SPARSE_UPDATE && mode != BN_SPARSE_ENERGY) return sp_fail(BN_EINVAL, "mode must be BN_SPARSE_UPDATE or BN_SPARSE_ENERGY");
  if (n_z < 1) return sp_fail(BN_EINVAL, "n_z must be >= 1");
  if (!lengthscale || !variance || !z_aug || !sm || !sP || !gain || !tr_off || !t || !y || !nat1 || !nat2 ||
      !site_mean || !site_cov || !part1 || !part2)
    return sp_fail(BN_EINVAL, "NULL array argument to bn_sparse_sites");
  if ((pair_mean == nullptr) != (pair_cov == nullptr)) return sp_fail(BN_EINVAL, "pair_mean and pair_cov go together");
  LikParams L;
  if (int rc = check_lik(lik, L)) return rc;
  GHTable gh;
  if (int rc = make_gh_table(gh_points, gh)) return rc;
  HyperView hv{lengthscale, variance, 0};
  SparseIO io{n_z, z_aug, sm, sP, gain, tr_off, t, y, nat1, nat2, site_mean, site_cov, pair_mean, pair_cov, part1,
              part2};
  hipStream_t st = (hipStream_t)stream;
  const int key = kern->kind[0] * 4 + (kern->n_comp > 1 ? kern->kind[1] : 0);
  (void)sp_state_dim;
  switch (key) {
    case 1 * 4 + 0: return dispatch_method<1, 0>(method, hv, L, gh, io, mode, lr, ensure_psd, st);
    case 2 * 4 + 0: return dispatch_method<2, 0>(method, hv, L, gh, io, mode, lr, ensure_psd, st);
    case 3 * 4 + 0: return dispatch_method<3, 0>(method, hv, L, gh, io, mode, lr, ensure_psd, st);
    case 1 * 4 + 1: return dispatch_method<1, 1>(method, hv, L, gh, io, mode, lr, ensure_psd, st);
    case 1 * 4 + 2: return dispatch_method<1, 2>(method, hv, L, gh, io, mode, lr, ensure_psd, st);
    case 1 * 4 + 3: return dispatch_method<1, 3>(method, hv, L, gh, io, mode, lr, ensure_psd, st);
    case 2 * 4 + 1: return dispatch_method<2, 1>(method, hv, L, gh, io, mode, lr, ensure_psd, st);
    case 2 * 4 + 2: return dispatch_method<2, 2>(method, hv, L, gh, io, mode, lr, ensure_psd, st);
    case 2 * 4 + 3: return dispatch_method<2, 3>(method, hv, L, gh, io, mode, lr, ensure_psd, st);
    case 3 * 4 + 1: return dispatch_method<3, 1>(method, hv, L, gh, io, mode, lr, ensure_psd, st);
    case 3 * 4 + 2: return dispatch_method<3, 2>(method, hv, L, gh, io, mode, lr, ensure_psd, st);
    case 3 * 4 + 3: return dispatch_method<3, 3>(method, hv, L, gh, io, mode, lr, ensure_psd, st);
    default: return sp_fail(BN_EUNSUPPORTED, "unsupported kernel combination");
  }
}

}  // extern "C"
