// HIP kernels and C ABI of the site updates and energy terms of the Markov-GP
// inference step (inference.py:21-325) on the plan's tiled layout.  A workgroup
// takes 64 chunks, one per lane, and its four wavefronts take every fourth step of
// them (the same coalesced access pattern as the filter kernels, four times the
// wavefronts); the per-chunk FP64 partial sums are combined in LDS, and a block per
// series reduces them in a fixed order.  The kernels are instantiated per likelihood:
// with only its branch compiled, the site step needs a third of the registers of one
// compiled for all likelihoods, and the wavefronts fit several to a SIMD.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <type_traits>
#include <vector>

#include "bn_internal.h"
#include "bn_sites.h"

using namespace bn;

namespace {

constexpr int kBlock = 256;
constexpr int kLanes = 64;                // chunks per workgroup, one per lane
constexpr int kPhases = kBlock / kLanes;  // wavefronts per workgroup, each every kPhases-th step

// Sum over the workgroup's wavefronts of a per-chunk partial (fixed order); the result
// is valid in the first wavefront.  Every thread of the workgroup calls it.
__device__ double phase_sum(double v, double* red) {
  const int lane = threadIdx.x % kLanes, w = threadIdx.x / kLanes;
  red[w * kLanes + lane] = v;
  __syncthreads();
  double t = 0.0;
  if (w == 0)
    for (int q = 0; q < kPhases; ++q) t += red[q * kLanes + lane];
  __syncthreads();  // red is reused by the next sum
  return t;
}

// Runs f with the likelihood kind as a compile-time constant.
template <class F>
int with_kind(int32_t kind, F&& f) {
  switch (kind) {
    case LIK_GAUSSIAN: f(std::integral_constant<int, LIK_GAUSSIAN>{}); return BN_OK;
    case LIK_BERNOULLI_PROBIT: f(std::integral_constant<int, LIK_BERNOULLI_PROBIT>{}); return BN_OK;
    case LIK_POISSON: f(std::integral_constant<int, LIK_POISSON>{}); return BN_OK;
    case LIK_STUDENTS_T: f(std::integral_constant<int, LIK_STUDENTS_T>{}); return BN_OK;
    default: return set_error(BN_EINVAL, "unknown likelihood kind");
  }
}

int fail(int code, const std::string& m) { return set_error(code, m); }

#define BN_HIP_CHECK2(expr)                                                                      \
  do {                                                                                           \
    hipError_t e_ = (expr);                                                                      \
    if (e_ != hipSuccess) return fail(BN_EHIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

struct SiteIO {
  const float* y;
  const float* pm;
  const float* pv;
  double* nat1;
  double* nat2;
  float* py;
  float* pr;
  float* o_mean;
  float* o_jac;
  float* o_hess;
  double* part1;  // per chunk: sum |nat1_new - nat1| (update) or t1 (energy)
  double* part2;
  const float* batch;  // optional per slot: nonzero = in the mini-batch (NULL: every site)
};

template <int METHOD, int KIND>
__global__ __launch_bounds__(kBlock) void k_site_update(PlanView pl, LikParams L, GHTable gh, SiteIO io, double lr,
                                                        double power, int ensure_psd) {
  __shared__ double red[kBlock];
  L.kind = KIND;
  const int32_t j = blockIdx.x * kLanes + (int32_t)(threadIdx.x % kLanes);
  const int32_t w = (int32_t)(threadIdx.x / kLanes);
  const int32_t len = j < pl.n_chunks ? pl.chunk_len_[j] : 0;
  double d1 = 0.0, d2 = 0.0;
  for (int32_t k = w; k < len; k += kPhases) {
    const int64_t s = pl.slot(j, k);
    // group_natural_params (basemodels.py:195-199): sites outside the batch keep their values
    if (io.batch != nullptr && io.batch[s] == 0.0f) continue;
    const double n1 = io.nat1[s], n2 = io.nat2[s];
    SiteResult r = site_step<METHOD>(L, gh, (double)io.y[s], (double)io.pm[s], (double)io.pv[s], n1, n2, power,
                                     ensure_psd != 0);
    d1 += fabs(r.nat1 - n1);
    d2 += fabs(r.nat2 - n2);
    // damping (inference.py:84-87) and reparametrisation (basemodels.py:85-100)
    const double a1 = (1.0 - lr) * n1 + lr * r.nat1;
    const double a2 = (1.0 - lr) * n2 + lr * r.nat2;
    io.nat1[s] = a1;
    io.nat2[s] = a2;
    io.pr[s] = (float)(1.0 / a2);
    io.py[s] = (float)(a1 / a2);
    if (io.o_mean != nullptr) {
      io.o_mean[s] = (float)r.mean;
      io.o_jac[s] = (float)r.jac;
      io.o_hess[s] = (float)r.hess;
    }
  }
  d1 = phase_sum(d1, red);
  d2 = phase_sum(d2, red);
  if (w == 0 && j < pl.n_chunks) {
    io.part1[j] = d1;
    io.part2[j] = d2;
  }
}

template <int METHOD, int KIND>
__global__ __launch_bounds__(kBlock) void k_energy_terms(PlanView pl, LikParams L, GHTable gh, SiteIO io,
                                                         double power) {
  __shared__ double red[kBlock];
  L.kind = KIND;
  const int32_t j = blockIdx.x * kLanes + (int32_t)(threadIdx.x % kLanes);
  const int32_t w = (int32_t)(threadIdx.x / kLanes);
  const int32_t len = j < pl.n_chunks ? pl.chunk_len_[j] : 0;
  double s1 = 0.0, s2 = 0.0;
  for (int32_t k = w; k < len; k += kPhases) {
    const int64_t s = pl.slot(j, k);
    double t1, t2;
    energy_terms<METHOD>(L, gh, (double)io.y[s], (double)io.pm[s], (double)io.pv[s], io.nat1[s], io.nat2[s], power,
                         t1, t2);
    // the data term runs over the mini-batch only (inference.py:130-325); the pseudo term over every site
    if (io.batch == nullptr || io.batch[s] != 0.0f) s1 += t1;
    s2 += t2;
  }
  s1 = phase_sum(s1, red);
  s2 = phase_sum(s2, red);
  if (w == 0 && j < pl.n_chunks) {
    io.part1[j] = s1;
    io.part2[j] = s2;
  }
}

template <int METHOD, int KIND>
__global__ __launch_bounds__(kBlock) void k_lik_grad(PlanView pl, LikParams L, GHTable gh, SiteIO io, double power) {
  __shared__ double red[kBlock];
  L.kind = KIND;
  const int32_t j = blockIdx.x * kLanes + (int32_t)(threadIdx.x % kLanes);
  const int32_t w = (int32_t)(threadIdx.x / kLanes);
  const int32_t len = j < pl.n_chunks ? pl.chunk_len_[j] : 0;
  double s1 = 0.0;
  for (int32_t k = w; k < len; k += kPhases) {
    const int64_t s = pl.slot(j, k);
    const double n1 = METHOD == M_EP ? io.nat1[s] : 0.0, n2 = METHOD == M_EP ? io.nat2[s] : 0.0;
    s1 += dterm1_dparam<METHOD>(L, gh, (double)io.y[s], (double)io.pm[s], (double)io.pv[s], n1, n2, power);
  }
  s1 = phase_sum(s1, red);
  if (w == 0 && j < pl.n_chunks) io.part1[j] = s1;
}

__global__ void k_reparametrise(PlanView pl, const double* __restrict__ nat1, const double* __restrict__ nat2,
                                float* __restrict__ py, float* __restrict__ pr) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= pl.n_slots) return;
  pr[s] = (float)(1.0 / nat2[s]);
  py[s] = (float)(nat1[s] / nat2[s]);
}

__global__ __launch_bounds__(kBlock) void k_series_sum1(PlanView pl, const double* __restrict__ a,
                                                        double* __restrict__ oa) {
  __shared__ double ra[kBlock];
  const int32_t s = blockIdx.x;
  const int32_t c0 = pl.series_first_chunk[s];
  const int32_t nc = pl.series_nchunks[s];
  ra[threadIdx.x] = strided_sum(a + c0, nc, threadIdx.x, kBlock);
  __syncthreads();
  for (int32_t w = kBlock / 2; w > 0; w >>= 1) {
    if ((int32_t)threadIdx.x < w) ra[threadIdx.x] += ra[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) oa[s] = ra[0];
}

__global__ __launch_bounds__(kBlock) void k_series_sum2(PlanView pl, const double* __restrict__ a,
                                                        const double* __restrict__ b, double* __restrict__ oa,
                                                        double* __restrict__ ob) {
  __shared__ double ra[kBlock], rb[kBlock];
  const int32_t s = blockIdx.x;
  const int32_t c0 = pl.series_first_chunk[s];
  const int32_t nc = pl.series_nchunks[s];
  ra[threadIdx.x] = strided_sum(a + c0, nc, threadIdx.x, kBlock);
  rb[threadIdx.x] = strided_sum(b + c0, nc, threadIdx.x, kBlock);
  __syncthreads();
  for (int32_t w = kBlock / 2; w > 0; w >>= 1) {
    if ((int32_t)threadIdx.x < w) {
      ra[threadIdx.x] += ra[threadIdx.x + w];
      rb[threadIdx.x] += rb[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    oa[s] = ra[0];
    ob[s] = rb[0];
  }
}

// Orthonormal Hermite polynomials for the weight exp(-t^2) at t:
// p_0 = pi^-1/4, p_1 = sqrt(2) t p_0, p_{k+1} = sqrt(2/(k+1)) t p_k - sqrt(k/(k+1)) p_{k-1}.
// Returns p_{n-1}(t), p_n(t) and sum_{k<n} p_k(t)^2.
void hermite_orthonormal(int n, double t, double& pnm1, double& pn, double& sumsq) {
  double pm = 0.0, p = 0.7511255444649425;  // pi^-1/4
  sumsq = 0.0;
  for (int k = 0; k < n; ++k) {
    sumsq += p * p;
    const double next = std::sqrt(2.0 / (k + 1)) * t * p - std::sqrt((double)k / (k + 1)) * pm;
    pm = p;
    p = next;
  }
  pnm1 = pm;
  pn = p;
}

// Gauss-Hermite rule by Golub-Welsch: the nodes are the eigenvalues of the
// Jacobi matrix of the weight exp(-t^2) (zero diagonal, off-diagonal sqrt(k/2)),
// found with cyclic Jacobi rotations and polished by Newton steps on p_n; the
// weights are the Christoffel numbers 1 / sum_{k<n} p_k(t_i)^2, which keep full
// relative accuracy in the tails (eigenvector components would not).  Rescaled
// for expectations under N(0, 1): x = sqrt(2) t, w = weight / sqrt(pi).
bool golub_welsch(int n, double* x, double* w) {
  if (n < 1 || n > kMaxGH) return false;
  std::vector<double> A(n * n, 0.0);
  for (int k = 1; k < n; ++k) A[(k - 1) * n + k] = A[k * n + k - 1] = std::sqrt(0.5 * k);
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) off += A[p * n + q] * A[p * n + q];
    if (off < 1e-36) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = A[p * n + q];
        if (std::fabs(apq) < 1e-300) continue;
        const double theta = (A[q * n + q] - A[p * n + p]) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {  // A <- J' A J
          const double akp = A[k * n + p], akq = A[k * n + q];
          A[k * n + p] = c * akp - s * akq;
          A[k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          const double apk = A[p * n + k], aqk = A[q * n + k];
          A[p * n + k] = c * apk - s * aqk;
          A[q * n + k] = s * apk + c * aqk;
        }
      }
  }
  std::vector<double> t(n);
  for (int i = 0; i < n; ++i) t[i] = A[i * n + i];
  std::sort(t.begin(), t.end());
  for (int i = 0; i < n; ++i) {
    double ti = t[i], pnm1, pn, ss;
    for (int it = 0; it < 3; ++it) {  // Newton on p_n, p_n' = sqrt(2n) p_{n-1}
      hermite_orthonormal(n, ti, pnm1, pn, ss);
      const double d = std::sqrt(2.0 * n) * pnm1;
      if (d != 0.0) ti -= pn / d;
    }
    hermite_orthonormal(n, ti, pnm1, pn, ss);
    x[i] = std::sqrt(2.0) * ti;
    w[i] = 1.0 / (ss * std::sqrt(M_PI));
  }
  return true;
}

// The table for n nodes is computed once per thread and reused.
int make_gh(int n, GHTable& gh) {
  thread_local GHTable cache[kMaxGH + 1];
  thread_local bool have[kMaxGH + 1] = {};
  if (n < 1 || n > kMaxGH) return fail(BN_EINVAL, "gh_points must be in [1, 64]");
  if (!have[n]) {
    cache[n].n = n;
    if (!golub_welsch(n, cache[n].x, cache[n].w)) return fail(BN_EINVAL, "Golub-Welsch failed");
    have[n] = true;
  }
  gh = cache[n];
  return BN_OK;
}

int make_lik_checked(const bn_lik_desc* d, LikParams& L) {
  if (d == nullptr) return fail(BN_EINVAL, "likelihood descriptor is NULL");
  switch (d->kind) {
    case BN_LIK_GAUSSIAN:
      if (!(d->param0 > 0)) return fail(BN_EINVAL, "Gaussian variance must be > 0");
      break;
    case BN_LIK_BERNOULLI_PROBIT:
      break;
    case BN_LIK_POISSON:
      if (!(d->param0 > 0)) return fail(BN_EINVAL, "Poisson binsize must be > 0");
      break;
    case BN_LIK_STUDENTS_T:
      if (!(d->param0 > 0) || !(d->param1 > 0)) return fail(BN_EINVAL, "Student-t scale and df must be > 0");
      break;
    default:
      return fail(BN_EUNSUPPORTED, "unknown likelihood kind");
  }
  L = make_lik(d->kind, d->param0, d->param1);
  return BN_OK;
}

}  // namespace

namespace bn {
int make_gh_table(int n, GHTable& gh) { return make_gh(n, gh); }
int check_lik(const bn_lik_desc* d, LikParams& L) { return make_lik_checked(d, L); }
}  // namespace bn

// Predictive moments in data space and the predictive log density at one point
// (Likelihood.predict / predict_cubature, cubature.py:438-465; Gaussian.predict and
// Gaussian.log_density, likelihoods.py:784-803; log_density_cubature, cubature.py:249-276
// with Z clipped at 1e-8).  Conditional moments (likelihoods.py): Bernoulli (p, p - p^2)
// with the jittered probit link, Poisson (binsize e^f, binsize e^f), Student-t
// (f, scale^2 df / (df - 2)).
template <int KIND>
__global__ void k_lik_predict(LikParams L, GHTable gh, int64_t n, const double* __restrict__ mean_f,
                              const double* __restrict__ var_f, const double* __restrict__ y,
                              double* __restrict__ mean_y, double* __restrict__ var_y,
                              double* __restrict__ log_density) {
  L.kind = KIND;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double m = mean_f[i], v = var_f[i];
  if (L.kind == LIK_GAUSSIAN) {
    if (mean_y) mean_y[i] = m;
    if (var_y) var_y[i] = v + L.p0;
    if (log_density) {
      const double s = L.p0 + v, r = y[i] - m;
      log_density[i] = -0.5 * (kLog2Pi + log(s) + r * r / s);
    }
    return;
  }
  const double sv = sqrt(v);
  double Ey = 0.0, Ey2 = 0.0, Z = 0.0;
  const double yi = log_density ? y[i] : 0.0;
  for (int k = 0; k < gh.n; ++k) {
    const double f = m + sv * gh.x[k];
    double mu, s2;
    if (L.kind == LIK_BERNOULLI_PROBIT) {
      mu = norm_cdf(f) * (1.0 - 2.0 * kProbitEps) + kProbitEps;
      s2 = mu - mu * mu;
    } else if (L.kind == LIK_POISSON) {
      mu = exp(f) * L.p0;
      s2 = mu;
    } else {
      mu = f;
      s2 = L.p0 * L.p0 * L.p1 / (L.p1 - 2.0);
    }
    Ey += gh.w[k] * mu;
    Ey2 += gh.w[k] * (s2 + mu * mu);
    if (log_density) Z += gh.w[k] * exp(log_lik(L, yi, f));
  }
  if (mean_y) mean_y[i] = Ey;
  if (var_y) var_y[i] = Ey2 - Ey * Ey;
  // np.maximum(Z, 1e-8) propagates NaN (a missing y), so nanmean skips the point
  if (log_density) log_density[i] = (Z != Z) ? Z : log(Z > 1e-8 ? Z : 1e-8);
}

extern "C" {

int bn_gauss_hermite(int32_t n, double* x, double* w) {
  if (x == nullptr || w == nullptr) return fail(BN_EINVAL, "NULL output");
  GHTable gh;
  if (int rc = make_gh(n, gh)) return rc;
  for (int i = 0; i < n; ++i) {
    x[i] = gh.x[i];
    w[i] = gh.w[i];
  }
  return BN_OK;
}

int bn_site_reparametrise(const bn_plan* plan, const double* nat1, const double* nat2, float* pseudo_y,
                          float* pseudo_var, void* stream) {
  if (plan == nullptr || !nat1 || !nat2 || !pseudo_y || !pseudo_var) return fail(BN_EINVAL, "NULL argument");
  const int64_t n = plan->host.n_slots;
  hipLaunchKernelGGL(k_reparametrise, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                     plan->view, nat1, nat2, pseudo_y, pseudo_var);
  BN_HIP_CHECK2(hipGetLastError());
  return BN_OK;
}

int bn_site_update(const bn_plan* plan, const bn_kernel_desc* kern, const bn_lik_desc* lik, int32_t method,
                   int32_t gh_points, double lr, double power, int32_t ensure_psd, const float* y,
                   const float* post_mean, const float* post_var, const float* batch, double* nat1, double* nat2,
                   float* pseudo_y,
                   float* pseudo_var, float* out_mean, float* out_jac, float* out_hess, double* diff1, double* diff2,
                   void* workspace, void* stream) {
  if (plan == nullptr) return fail(BN_EINVAL, "plan is NULL");
  if (!y || !post_mean || !post_var || !nat1 || !nat2 || !pseudo_y || !pseudo_var || !diff1 || !diff2 || !workspace)
    return fail(BN_EINVAL, "NULL array argument to bn_site_update");
  if ((out_mean == nullptr) != (out_jac == nullptr) || (out_mean == nullptr) != (out_hess == nullptr))
    return fail(BN_EINVAL, "out_mean / out_jac / out_hess must be all NULL or all set");
  if (method == BN_EP && !(power > 0)) return fail(BN_EINVAL, "EP power must be > 0");
  LikParams L;
  if (int rc = make_lik_checked(lik, L)) return rc;
  GHTable gh;
  if (int rc = make_gh(gh_points, gh)) return rc;
  double* p1 = site_partials(plan, kern, workspace, 0);
  double* p2 = site_partials(plan, kern, workspace, 1);
  if (!p1 || !p2) return fail(BN_EINVAL, "bad kernel descriptor for the workspace");
  SiteIO io{y, post_mean, post_var, nat1, nat2, pseudo_y, pseudo_var, out_mean, out_jac, out_hess, p1, p2, batch};
  hipStream_t st = (hipStream_t)stream;
  const dim3 grid((plan->host.n_chunks + kLanes - 1) / kLanes);
  if (method != BN_VI && method != BN_NEWTON && method != BN_EP)
    return fail(BN_EINVAL, "method must be BN_VI, BN_NEWTON or BN_EP");
  if (int rc = with_kind(L.kind, [&](auto kd) {
        constexpr int KD = decltype(kd)::value;
        if (method == BN_VI)
          hipLaunchKernelGGL((k_site_update<M_VI, KD>), grid, dim3(kBlock), 0, st, plan->view, L, gh, io, lr, power,
                             ensure_psd);
        else if (method == BN_NEWTON)
          hipLaunchKernelGGL((k_site_update<M_NEWTON, KD>), grid, dim3(kBlock), 0, st, plan->view, L, gh, io, lr,
                             power, ensure_psd);
        else
          hipLaunchKernelGGL((k_site_update<M_EP, KD>), grid, dim3(kBlock), 0, st, plan->view, L, gh, io, lr, power,
                             ensure_psd);
      }))
    return rc;
  hipLaunchKernelGGL(k_series_sum2, dim3(plan->host.n_series), dim3(kBlock), 0, st, plan->view, p1, p2, diff1, diff2);
  BN_HIP_CHECK2(hipGetLastError());
  return BN_OK;
}

int bn_energy_terms(const bn_plan* plan, const bn_kernel_desc* kern, const bn_lik_desc* lik, int32_t method,
                    int32_t gh_points, double power, const float* y, const float* post_mean, const float* post_var,
                    const float* batch, const double* nat1, const double* nat2, double* term1, double* term2,
                    void* workspace,
                    void* stream) {
  if (plan == nullptr) return fail(BN_EINVAL, "plan is NULL");
  if (!y || !post_mean || !post_var || !nat1 || !nat2 || !term1 || !term2 || !workspace)
    return fail(BN_EINVAL, "NULL array argument to bn_energy_terms");
  if (method == BN_EP && !(power > 0)) return fail(BN_EINVAL, "EP power must be > 0");
  LikParams L;
  if (int rc = make_lik_checked(lik, L)) return rc;
  GHTable gh;
  if (int rc = make_gh(gh_points, gh)) return rc;
  double* p1 = site_partials(plan, kern, workspace, 0);
  double* p2 = site_partials(plan, kern, workspace, 1);
  if (!p1 || !p2) return fail(BN_EINVAL, "bad kernel descriptor for the workspace");
  SiteIO io{y, post_mean, post_var, const_cast<double*>(nat1), const_cast<double*>(nat2), nullptr, nullptr,
            nullptr, nullptr, nullptr, p1, p2, batch};
  hipStream_t st = (hipStream_t)stream;
  const dim3 grid((plan->host.n_chunks + kLanes - 1) / kLanes);
  if (method != BN_VI && method != BN_NEWTON && method != BN_EP)
    return fail(BN_EINVAL, "method must be BN_VI, BN_NEWTON or BN_EP");
  if (int rc = with_kind(L.kind, [&](auto kd) {
        constexpr int KD = decltype(kd)::value;
        if (method == BN_VI)
          hipLaunchKernelGGL((k_energy_terms<M_VI, KD>), grid, dim3(kBlock), 0, st, plan->view, L, gh, io, power);
        else if (method == BN_NEWTON)
          hipLaunchKernelGGL((k_energy_terms<M_NEWTON, KD>), grid, dim3(kBlock), 0, st, plan->view, L, gh, io, power);
        else
          hipLaunchKernelGGL((k_energy_terms<M_EP, KD>), grid, dim3(kBlock), 0, st, plan->view, L, gh, io, power);
      }))
    return rc;
  hipLaunchKernelGGL(k_series_sum2, dim3(plan->host.n_series), dim3(kBlock), 0, st, plan->view, p1, p2, term1, term2);
  BN_HIP_CHECK2(hipGetLastError());
  return BN_OK;
}

int bn_lik_param_grad(const bn_plan* plan, const bn_kernel_desc* kern, const bn_lik_desc* lik, int32_t method,
                      int32_t gh_points, double power, const float* y, const float* post_mean, const float* post_var,
                      const double* nat1, const double* nat2, double* grad, void* workspace, void* stream) {
  if (plan == nullptr) return fail(BN_EINVAL, "plan is NULL");
  if (!y || !post_mean || !post_var || !grad || !workspace) return fail(BN_EINVAL, "NULL array argument");
  if (method == BN_EP && (!nat1 || !nat2)) return fail(BN_EINVAL, "EP needs the sites nat1 / nat2");
  if (method == BN_EP && !(power > 0.0 && power <= 1.0)) return fail(BN_EINVAL, "EP power must be in (0, 1]");
  LikParams L;
  if (int rc = make_lik_checked(lik, L)) return rc;
  GHTable gh;
  if (int rc = make_gh(gh_points, gh)) return rc;
  double* p1 = site_partials(plan, kern, workspace, 0);
  if (!p1) return fail(BN_EINVAL, "bad kernel descriptor for the workspace");
  SiteIO io{y, post_mean, post_var, const_cast<double*>(nat1), const_cast<double*>(nat2), nullptr, nullptr,
            nullptr, nullptr, nullptr, p1, nullptr};
  hipStream_t st = (hipStream_t)stream;
  const dim3 grid((plan->host.n_chunks + kLanes - 1) / kLanes);
  if (method != BN_VI && method != BN_NEWTON && method != BN_EP)
    return fail(BN_EINVAL, "method must be BN_VI, BN_NEWTON or BN_EP");
  if (int rc = with_kind(L.kind, [&](auto kd) {
        constexpr int KD = decltype(kd)::value;
        if (method == BN_VI)
          hipLaunchKernelGGL((k_lik_grad<M_VI, KD>), grid, dim3(kBlock), 0, st, plan->view, L, gh, io, power);
        else if (method == BN_NEWTON)
          hipLaunchKernelGGL((k_lik_grad<M_NEWTON, KD>), grid, dim3(kBlock), 0, st, plan->view, L, gh, io, power);
        else
          hipLaunchKernelGGL((k_lik_grad<M_EP, KD>), grid, dim3(kBlock), 0, st, plan->view, L, gh, io, power);
      }))
    return rc;
  hipLaunchKernelGGL(k_series_sum1, dim3(plan->host.n_series), dim3(kBlock), 0, st, plan->view, p1, grad);
  BN_HIP_CHECK2(hipGetLastError());
  return BN_OK;
}

int bn_lik_predict(const bn_lik_desc* lik, int32_t gh_points, int64_t n, const double* mean_f, const double* var_f,
                   const double* y, double* mean_y, double* var_y, double* log_density, void* stream) {
  LikParams L;
  if (int rc = check_lik(lik, L)) return rc;
  GHTable gh;
  if (int rc = make_gh_table(gh_points, gh)) return rc;
  if (n < 0) return fail(BN_EINVAL, "n must be >= 0");
  if (n > 0 && (!mean_f || !var_f)) return fail(BN_EINVAL, "NULL mean_f / var_f");
  if (n > 0 && log_density && !y) return fail(BN_EINVAL, "log_density needs y");
  if (n > 0) {
    if (int rc = with_kind(L.kind, [&](auto kd) {
          hipLaunchKernelGGL((k_lik_predict<decltype(kd)::value>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                             (hipStream_t)stream, L, gh, n, mean_f, var_f, y, mean_y, var_y, log_density);
        }))
      return rc;
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(BN_EHIP, std::string("bn_lik_predict: ") + hipGetErrorString(e));
  return BN_OK;
}

}  // extern "C"
