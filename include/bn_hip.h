/*
 * bn_hip.h - C ABI of the MI355X-native Markov-GP hot path (libbn_hip.so).
 *
 * The reference (AaltoML/BayesNewton v1.3.4) is pure Python/JAX and has no FFI;
 * its boundary for this path is the Python API
 *   ops.kalman_filter(dt, kernel, y, noise_cov, mask, parallel, return_predict)
 *                                                   bayesnewton/ops.py:256-285
 *   ops.rauch_tung_striebel_smoother(dt, kernel, filter_mean, filter_cov,
 *                                    return_full, parallel) bayesnewton/ops.py:357-380
 *   MarkovGaussianProcess.update_posterior / compute_log_lik
 *                                     bayesnewton/basemodels.py:689-706, 726-741
 *   InferenceMixin.inference / <Newton|VariationalInference|
 *   ExpectationPropagation>.update_variational_params / energy
 *                                     bayesnewton/inference.py:65-325
 * Each entry point below names the reference function it replaces.  The Python
 * package bayesnewton_amd binds them with ctypes and re-exposes the reference
 * signatures.
 *
 * Conventions
 *  - All array arguments are DEVICE pointers, caller-allocated (e.g. torch
 *    tensors); `stream` is a hipStream_t passed as void* (NULL = null stream).
 *  - Per-step arrays live in the plan's TILED layout: the time steps of every
 *    series are cut into chunks of at most `chunk_len` steps; chunk j's step k
 *    sits at slot ((j / G) * chunk_len + k) * G + (j % G) with G = 64 (or the
 *    next power of two >= the chunk count when there are fewer than 64 chunks),
 *    so the 64 lanes of a wavefront that each walk one chunk read consecutive
 *    words.  Use
 *    bn_plan_gather_* / bn_plan_scatter_* to convert from/to the reference's
 *    series-contiguous [N] layout.  Per-chunk and per-tile scratch lives in a
 *    caller-provided workspace of bn_workspace_bytes() bytes.
 *  - Kernels: a sum of at most two Matern components (kernels.py:123-300,
 *    Sum 1624-1663).  Hyperparameters are natural-scale FP64 arrays
 *    lengthscale[b * hyper_stride + c], variance[b * hyper_stride + c]
 *    (hyper_stride = 0 shares one set across all series).
 *  - Numerics: per-step inputs (dt, y, r) and the fused path's posterior
 *    outputs are FP32; the state recurrences, the discretisation A(dt),
 *    Q(dt) = Pinf - A Pinf A', the log-likelihood accumulation, the scan
 *    elements and the filtered-moment scratch are FP64 (FP32 recurrences leave
 *    the posterior variance of a Matern32 + Matern52 sum ~1e-3 relative from
 *    the FP64 reference).  The ops.* mirrors return FP64 moments, as the
 *    reference does.
 *  - Return value: 0 on success, a negative BN_E* code otherwise; the message is
 *    available from bn_last_error() (thread-local).  Numerical failure (a
 *    non-PD matrix) is not an error: as in the reference, NaN propagates into the
 *    affected series' outputs, and the fused pass and the filter also report it per
 *    series in an optional status[n_series] array of BN_STATUS_* bits.
 */
#ifndef BN_HIP_H
#define BN_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BN_OK 0
#define BN_EINVAL -1
#define BN_EHIP -2
#define BN_ENOMEM -3
#define BN_EUNSUPPORTED -4

#define BN_MATERN12 1
#define BN_MATERN32 2
#define BN_MATERN52 3

/* per-series status bits (status[b], OR over the series' steps; 0 = none) */
#define BN_STATUS_INNOVATION_NONPD 1 /* some step's innovation variance H P- H' + r was not > 0 (or NaN) */
#define BN_STATUS_NONFINITE 2        /* the series' log-likelihood or final filtered mean is NaN / infinite */
#define BN_STATUS_SMOOTHER_NONPD 4   /* some smoothed variance was not > 0 or a smoothed mean not finite */

/* flags */
#define BN_RETURN_PREDICT 0x1u /* kalman_filter(return_predict=True), ops.py:172-175 */
#define BN_RETURN_FULL 0x2u    /* rauch_tung_striebel_smoother(return_full=True), ops.py:302-305 */

typedef struct bn_kernel_desc {
  int32_t n_comp;  /* 1 or 2 */
  int32_t kind[2]; /* BN_MATERN12 / BN_MATERN32 / BN_MATERN52, in Sum order */
} bn_kernel_desc;

typedef struct bn_plan bn_plan; /* opaque: chunk/tile tables on the device */

/* likelihoods (likelihoods.py: Gaussian 684-803, Bernoulli/probit 806-871,
 * Poisson 891-1008, StudentsT 1011-1044) */
#define BN_LIK_GAUSSIAN 1         /* param0 = noise variance */
#define BN_LIK_BERNOULLI_PROBIT 2 /* y in {0, 1}; link Phi(f)(1 - 2e-3) + 1e-3 */
#define BN_LIK_POISSON 3          /* param0 = binsize; exp link */
#define BN_LIK_STUDENTS_T 4       /* param0 = scale, param1 = degrees of freedom */

/* inference schemes (inference.py: Newton 99-157, VariationalInference
 * 160-222, ExpectationPropagation 225-325) */
#define BN_VI 0
#define BN_NEWTON 1
#define BN_EP 2

typedef struct bn_lik_desc {
  int32_t kind;
  double param0, param1;
} bn_lik_desc;

const char *bn_last_error(void);
int bn_version(void);

/* --- plan: the series -> chunk -> tile decomposition ------------------------
 * offsets_host[n_series + 1]: CSR offsets (host memory) of the series in the
 * [N] layout, as the reference's per-model arrays concatenated.  chunk_len:
 * steps walked sequentially by one lane (1..2^24).  Each series may span at
 * most 128 * 128 chunks (level-1 tiles of 128 chunks, at most 128 tiles).  Replaces input_admin's bookkeeping (utils.py:234-265) for
 * many series at once; the reference handles one series per model. */
int bn_plan_create(const int64_t *offsets_host, int32_t n_series, int32_t chunk_len, bn_plan **out);
int bn_plan_destroy(bn_plan *plan);
int64_t bn_plan_num_slots(const bn_plan *plan);
int32_t bn_plan_num_chunks(const bn_plan *plan);
int32_t bn_plan_num_tiles(const bn_plan *plan);
int64_t bn_workspace_bytes(const bn_plan *plan, const bn_kernel_desc *kern);

/* [N] -> tiled (padding slots get `pad`) and tiled -> [N].  `width` fields are
 * interleaved in the [N] array ([N, width] row-major) and stored as `width`
 * tiled planes of num_slots words each. */
int bn_plan_gather_f32(const bn_plan *plan, const float *src, float *dst, int32_t width, float pad, void *stream);
int bn_plan_scatter_f32(const bn_plan *plan, const float *src, float *dst, int32_t width, void *stream);
int bn_plan_gather_f64(const bn_plan *plan, const double *src, double *dst, int32_t width, double pad, void *stream);
int bn_plan_scatter_f64(const bn_plan *plan, const double *src, double *dst, int32_t width, void *stream);
int bn_plan_gather_u8(const bn_plan *plan, const uint8_t *src, uint8_t *dst, uint8_t pad, void *stream);

/* --- discretisation: vmap(kernel.state_transition)(dt) and
 * process_noise_covariance, ops.py:274-279 / kernels.py:216-224, 273-286.
 * dt[n] -> A[n, d, d], Q[n, d, d] (row-major, FP32, computed in FP64). */
int bn_discretise(const bn_kernel_desc *kern, const double *lengthscale, const double *variance, int64_t n,
                  const float *dt, float *A, float *Q, void *stream);

/* --- filter + smoother, fused: MarkovGaussianProcess.update_posterior,
 * basemodels.py:689-706 (filter ops.py:154-253, shifted dt basemodels.py:700,
 * smoother ops.py:288-354).  Inputs dt, y (pseudo mean), r (pseudo variance),
 * mask (1 = missing, may be NULL) are tiled.  Outputs: post_mean / post_var
 * (tiled, H m and H P H') and loglik[n_series] (FP64, the filter's log p(y),
 * i.e. compute_log_lik, basemodels.py:726-741).  With chunk_len >= the longest
 * series this is the sequential scan; otherwise the associative scan.  status[n_series]
 * (may be NULL) receives the BN_STATUS_* bits of every series. */
int bn_filter_smoother(const bn_plan *plan, const bn_kernel_desc *kern, const double *lengthscale,
                       const double *variance, int32_t hyper_stride, const float *dt, const float *y, const float *r,
                       const uint8_t *mask, float *post_mean, float *post_var, double *loglik, int32_t *status,
                       void *workspace, void *stream);

/* --- bn_filter_smoother without allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory during the call (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, kernels and results as bn_filter_smoother minus `workspace`:
 * the scratch lives in a device array of BN_NOALLOC_SCRATCH_BYTES inside the
 * library (one per GPU, held until the process ends; its address is looked up on
 * the first call on each device).  A plan whose bn_workspace_bytes() exceeds
 * that size is rejected with BN_EINVAL.  The call allocates nothing (the plan is
 * created beforehand with bn_plan_create) and records no timing events.  Calls
 * on one GPU share the array: calls that overlap on one GPU (other streams,
 * other host threads) race and must be serialised by the caller. */
#define BN_NOALLOC_SCRATCH_BYTES (64u * 1024u * 1024u)
int bn_filter_smoother_noalloc(const bn_plan *plan, const bn_kernel_desc *kern, const double *lengthscale,
                               const double *variance, int32_t hyper_stride, const float *dt, const float *y,
                               const float *r, const uint8_t *mask, float *post_mean, float *post_var,
                               double *loglik, int32_t *status, void *stream);

/* --- ops.kalman_filter, ops.py:256-285.  Writes the filtered (or, with
 * BN_RETURN_PREDICT, predicted) means [d tiled planes] and covariances [d*d
 * tiled planes, row-major], FP64, loglik[n_series] and (if not NULL) status[n_series]
 * (BN_STATUS_INNOVATION_NONPD / BN_STATUS_NONFINITE). */
int bn_kalman_filter(const bn_plan *plan, const bn_kernel_desc *kern, const double *lengthscale,
                     const double *variance, int32_t hyper_stride, const float *dt, const float *y, const float *r,
                     const uint8_t *mask, double *means, double *covs, double *loglik, uint32_t flags, int32_t *status,
                     void *workspace, void *stream);

/* --- log-likelihood only: MarkovGaussianProcess.compute_log_lik
 * (basemodels.py:726-741), the filter's log p(pseudo_y) per series without
 * writing any state. */
int bn_filter_loglik(const bn_plan *plan, const bn_kernel_desc *kern, const double *lengthscale,
                     const double *variance, int32_t hyper_stride, const float *dt, const float *y, const float *r,
                     const uint8_t *mask, double *loglik, void *workspace, void *stream);

/* --- gradient of the filter log-likelihood log Z (compute_log_lik) with
 * respect to every component's (lengthscale, variance), natural scale:
 * grad[b * 2 n_comp + 2 c + {0, 1}] = d log Z_b / d{lengthscale_c, variance_c}.
 * This is the kernel-hyperparameter part of objax.GradValues(model.energy)
 * (README.md:59, tests/test_gp_vs_markovgp_reg.py:87-96) with the sites and
 * posterior held fixed: d energy / d theta = -d log Z / d theta.  Exact
 * forward-mode tangents through the chunked scan (no per-step storage);
 * grad_workspace holds bn_grad_workspace_bytes() bytes.  loglik[n_series] is
 * the primal log Z. */
int64_t bn_grad_workspace_bytes(const bn_plan *plan, const bn_kernel_desc *kern);
int bn_filter_loglik_grad(const bn_plan *plan, const bn_kernel_desc *kern, const double *lengthscale,
                          const double *variance, int32_t hyper_stride, const float *dt, const float *y,
                          const float *r, const uint8_t *mask, double *loglik, double *grad, void *workspace,
                          void *grad_workspace, void *stream);

/* --- ops.rauch_tung_striebel_smoother, ops.py:357-380.  dt is the smoother's
 * (already shifted) step vector, tiled.  filter_mean [d planes], filter_cov
 * [d*d planes], FP64.  Outputs (FP64): with BN_RETURN_FULL, means [d planes] /
 * covs [d*d planes]; otherwise H m [1 plane] / H P H' [1 plane].  gains [d*d
 * planes] or NULL.  The last step of each series returns the filtered moments, as the
 * reference's associative form does (ops.py:314-315). */
int bn_rts_smoother(const bn_plan *plan, const bn_kernel_desc *kern, const double *lengthscale,
                    const double *variance, int32_t hyper_stride, const float *dt, const double *filter_mean,
                    const double *filter_cov, double *means, double *covs, double *gains, uint32_t flags,
                    void *workspace, void *stream);

/* --- prediction at test times: MarkovGaussianProcess.predict (basemodels.py:766-816)
 * with temporal_conditional / predict_from_state_ / compute_conditional_statistics
 * (utils.py:99-215).  For every test time the smoothed state is conditioned on the two
 * neighbouring training states (sentinels (0, Pinf) at -/+1e10 outside the data) and
 * projected by H.  train_offsets / test_offsets: device int64 [n_series + 1] CSR
 * offsets; x_train: the training times the plan was built on (sorted within each
 * series); x_test: test times of each series (any order); means / covs / gains: the
 * smoother's full output (bn_rts_smoother with BN_RETURN_FULL, tiled planes, FP64).
 * Outputs out_mean / out_var [n_test], FP64.  The call takes stream-ordered scratch
 * (hipMallocAsync / hipFreeAsync on `stream`, about 24 bytes per test time) for the
 * test times' neighbour indices and their sort by the state slot they read. */
int bn_predict(const bn_plan *plan, const bn_kernel_desc *kern, const double *lengthscale, const double *variance,
               int32_t hyper_stride, const int64_t *train_offsets, const double *x_train, const double *means,
               const double *covs, const double *gains, const int64_t *test_offsets, const double *x_test,
               int64_t n_test, double *out_mean, double *out_var, void *stream);

/* --- predictive moments in data space and log predictive density per point:
 * Likelihood.predict (predict_cubature, cubature.py:438-465; Gaussian.predict,
 * likelihoods.py:802-803) and Likelihood.log_density (log_density_cubature,
 * cubature.py:249-276; Gaussian.log_density, likelihoods.py:784-800), the per-point
 * maps of predict_y and negative_log_predictive_density (basemodels.py:165-193).
 * mean_f / var_f / y [n] FP64 (y only with log_density; NaN = missing gives NaN);
 * mean_y / var_y / log_density [n] FP64, each may be NULL. */
int bn_lik_predict(const bn_lik_desc *lik, int32_t gh_points, int64_t n, const double *mean_f, const double *var_f,
                   const double *y, double *mean_y, double *var_y, double *log_density, void *stream);

/* --- site update: one inference step's site map for every time step,
 * InferenceMixin.inference (inference.py:65-90) between its two
 * update_posterior calls: update_variational_params (VI 170-195, Newton
 * 105-128, EP 238-284 with the cavity of basemodels.py:230-245),
 * ensure_diagonal_positive_precision (utils.py:89-96) when ensure_psd,
 * newton_update (21-39), damping with lr and reparametrisation
 * (basemodels.py:85-100).  Tiled arrays: y (observations, NaN = missing),
 * post_mean / post_var (the current posterior), nat1 / nat2 (FP64 site
 * natural parameters, updated in place; nat2 = 1 / site variance as in
 * basemodels.py:56-57), pseudo_y / pseudo_var (written: the filter's inputs).
 * out_mean / out_jac / out_hess: the (mean, jacobian, hessian) inference()
 * returns, or all NULL.  diff1 / diff2 [n_series]: sums of |nat_new - nat_old|
 * before damping (inference.py:79-81 divides by N).  gh_points: Gauss-Hermite
 * nodes for the cubature likelihoods (reference default 20).  batch: optional tiled
 * mask (nonzero = in the mini-batch of inference(batch_ind=...)): sites outside it are
 * left as they are and add nothing to diff1 / diff2 (group_natural_params,
 * basemodels.py:195-199; inference.py:65-90); NULL updates every site. */
int bn_site_update(const bn_plan *plan, const bn_kernel_desc *kern, const bn_lik_desc *lik, int32_t method,
                   int32_t gh_points, double lr, double power, int32_t ensure_psd, const float *y,
                   const float *post_mean, const float *post_var, const float *batch, double *nat1, double *nat2,
                   float *pseudo_y, float *pseudo_var, float *out_mean, float *out_jac, float *out_hess,
                   double *diff1, double *diff2, void *workspace, void *stream);

/* --- energy terms per series (inference.py:130-154, 197-222, 286-325):
 * term1 = VI: sum E_q[log p(y|f)] | Newton: sum log p(y|m) | EP: sum log Z
 * term2 = VI/Newton: sum E_q[log N(pseudo_y | f, pseudo_var)] (compute_kl,
 *         basemodels.py:708-724) | EP: sum of lel_pseudo with pep_constant
 *         (basemodels.py:247-262).  Masked sites contribute 0.
 * With logZ the filter log-likelihood of the same sites: energy =
 * -(term1 - (term2 - logZ)) for VI / Newton, -(logZ + (term1 - term2) / power)
 * for EP.  batch: optional tiled mask of energy(batch_ind=...): term1 sums the
 * mini-batch only (the caller scales it by N / batch size), term2 every site;
 * NULL = every site. */
int bn_energy_terms(const bn_plan *plan, const bn_kernel_desc *kern, const bn_lik_desc *lik, int32_t method,
                    int32_t gh_points, double power, const float *y, const float *post_mean, const float *post_var,
                    const float *batch, const double *nat1, const double *nat2, double *term1, double *term2,
                    void *workspace, void *stream);

/* d term1 / d(likelihood hyperparameter) per series with the posterior and sites fixed
 * (the likelihood part of objax.GradValues(model.energy); d energy = -this):
 * Gaussian variance or Student-t scale, natural scale; 0 for likelihoods
 * without one.  VI, Newton and EP (term1 of EP is (1/power) sum log Z_n, the
 * moment-matched log partition at each site's cavity, which needs the sites
 * nat1 / nat2 and power; VI and Newton ignore them and nat1 / nat2 may be NULL). */
int bn_lik_param_grad(const bn_plan *plan, const bn_kernel_desc *kern, const bn_lik_desc *lik, int32_t method,
                      int32_t gh_points, double power, const float *y, const float *post_mean, const float *post_var,
                      const double *nat1, const double *nat2, double *grad, void *workspace, void *stream);

/* pseudo_y = nat1 / nat2, pseudo_var = 1 / nat2 on every slot (basemodels.py:85-90). */
int bn_site_reparametrise(const bn_plan *plan, const double *nat1, const double *nat2, float *pseudo_y,
                          float *pseudo_var, void *stream);

/* Gauss-Hermite nodes / weights for E_{z~N(0,1)} (cubature.py:56-84): x = sqrt(2)
 * times the Hermite roots, w = weights / sqrt(pi), by Golub-Welsch on the host. */
int bn_gauss_hermite(int32_t n, double *x, double *w);

/* Kalman filter over pairs of neighbouring states, kalman_filter_pairs (ops.py:383-426), the
 * filter of SparseMarkovGaussianProcess (basemodels.py:928-1011): the pair (x_{k-1}, x_k)
 * with A_pair = [[0, I], [0, A]], Q_pair = blkdiag(1e-32 I, Q), observed whole (H = I)
 * with a full 2d x 2d noise per step.  n_series series of steps offsets[s] ..
 * offsets[s+1] (int64, device); dt [N], y [N,2d], noise [N,2d,2d] (FP64).  Outputs the pair
 * means [N,2d] and covariances [N,2d,2d] after every step (the reference returns the first
 * state of the pairs after each series' first step) and loglik [n_series].  No mask. */
int bn_pairs_filter(const bn_kernel_desc *kern, const double *lengthscale, const double *variance,
                    int32_t hyper_stride, int32_t n_series, const int64_t *offsets, const double *dt, const double *y,
                    const double *noise, double *pair_means, double *pair_covs, double *loglik, void *stream);

/* --- the sparse Markov GP's per-transition work (SparseMarkovGaussianProcess,
 * basemodels.py:928-1152; VI / Newton mixins inference.py:65-222).  The n_z inducing
 * states sit at z_aug[1..n_z] (sorted) with sentinels z_aug[0] = -1e10, z_aug[n_z+1] =
 * 1e10; transition i in [0, n_z] joins u_i, u_{i+1} and carries a 2d-dim site nat1
 * [n_z+1, 2d], nat2 [n_z+1, 2d, 2d] with its mean / covariance site_mean, site_cov.
 * sm [n_z, d], sP [n_z, d, d], gain [n_z, d, d]: the RTS smoother's state, covariance and
 * gain over the inducing states (contiguous).  tr_off [n_z+2]: CSR offsets of the data
 * points per transition (t, y [N] sorted by time; y NaN = missing).  One lane per
 * transition builds the pair posterior (written to pair_mean / pair_cov when not NULL),
 * the conditional of every data point given its pair (utils.py:185-215) and then
 *   BN_SPARSE_UPDATE: the site terms, their pair-space map, newton_update, the grouping
 *     of group_natural_params (basemodels.py:1115-1140) and the lr damping, in place in
 *     nat1 / nat2 / site_mean / site_cov; part1 / part2 [n_z+1] = |d nat1|, |d nat2| sums;
 *   BN_SPARSE_ENERGY: part1 = the data terms (VI E_q log p, Newton log p at the mean),
 *     part2 = E_q log N(site_mean | pair, site_cov) (basemodels.py:214-222), per transition.
 * Temporal kernels (1 or 2 Matern components), VI and Newton, FP64. */
#define BN_SPARSE_UPDATE 0
#define BN_SPARSE_ENERGY 1
int bn_sparse_sites(const bn_kernel_desc *kern, const double *lengthscale, const double *variance,
                    const bn_lik_desc *lik, int32_t method, int32_t mode, int32_t gh_points, double lr,
                    int32_t ensure_psd, int32_t n_z, const double *z_aug, const double *sm, const double *sP,
                    const double *gain, const int64_t *tr_off, const double *t, const double *y, double *nat1,
                    double *nat2, double *site_mean, double *site_cov, double *pair_mean, double *pair_cov,
                    double *part1, double *part2, void *stream);

/* --- space-time models: SpatioTemporalKernel (kernels.py:385-586, sparse, 'Full'
 * conditional) driven by MarkovGaussianProcess (basemodels.py:625-764) --------
 * K independent runs ("runs"), each T steps with R observation locations in 2-D
 * and M <= 64 spatial inducing points z; temporal kernel Matern12 or Matern32
 * (state dim KT), state dim D = KT*M ordered [f_0..f_{M-1}, f'_0..f'_{M-1}].
 * All arrays are dense, run-major: [K,T,...].  FP64 throughout except the
 * FP32 observations y and locations r. */
typedef struct bn_st_dims {
  int32_t n_runs, n_steps, n_loc, n_inducing;
  int32_t temporal_kind; /* BN_MATERN12 or BN_MATERN32 */
} bn_st_dims;

typedef struct bn_st_hypers {
  const double *ls_t, *var_t; /* temporal lengthscale / variance (device, per run) */
  const double *var_s;        /* spatial kernel variance (device, per run) */
  int32_t stride;             /* 1: one value per run; 0: one shared value */
} bn_st_hypers;

#define BN_ST_UPDATE 0 /* bn_st_sites: site update + projected pseudo observation */
#define BN_ST_ENERGY 1 /* bn_st_sites: energy terms at the stored sites */
#define BN_ST_PROJECT 2 /* bn_st_sites: projected pseudo observation of the current sites only */

/* B_t = Krz Kzz^-1 Lzz (rows b_i = Lzz^-1 k_s(z, r_i)) and diag C_t =
 * var_t (var_s - |b_i|^2) of SpatioTemporalKernel.spatial_conditional
 * (kernels.py:486-506) for every step of every run.  z [M,2], Lzz [M,M] =
 * chol(k_s(z, z)) (inducing_precision, kernels.py:508-515), r [K,T,R,2];
 * B [K,T,R,M], cdiag [K,T,R].  Spatial kernel: spatial_kind (BN_MATERN12/32/52)
 * with lengthscale ls_s and variance hyp->var_s. */
int bn_st_conditional(const bn_st_dims *dims, const bn_st_hypers *hyp, int32_t spatial_kind, double ls_s,
                      const double *z,
                      const double *Lzz, const float *r, double *B, double *cdiag, void *stream);

/* Site map of one inference step or energy evaluation on every (run, step):
 * mean_f = B m, var_f = diag(B P B') + diag C from the posterior (post_mean
 * [K,T,M], post_cov [K,T,M,M]; conditional_posterior_to_data, basemodels.py:743-764).
 * mode BN_ST_UPDATE (inference.py:65-90 with 105-128 / 170-195): new sites
 * nat1/nat2 [K,T,R] (diagonal) damped by lr; part0/part1 [K,T] = per-step sums
 * of |dnat1|, |dnat2|; the projected pseudo observation of compute_full_pseudo_lik
 * (basemodels.py:676-687) in whitened form: Lw [K,T,M,M] = chol(B' diag(nat2) B +
 * 1e-12 I), zw [K,T,M] = Lw^-1 B' nat1, ldl [K,T] = sum log diag Lw.
 * mode BN_ST_ENERGY (inference.py:130-154 / 197-222): part0 = per-step sum of
 * E_q[log p(y|f)] (VI) or log p(y|m) (Newton); part1 = per-step
 * gaussian_expected_log_lik of the pseudo observation (utils.py:510-531) from
 * Lw, zw, ldl.  mode BN_ST_PROJECT: Lw, zw, ldl of the current sites only
 * (post_* unused).  Methods BN_VI and BN_NEWTON.  out_* optional [K,T,R].
 * batch: optional [K,T] mask of the mini-batch of time steps (inference(batch_ind=...),
 * energy(batch_ind=...)): steps outside it keep their sites (UPDATE) and add nothing to
 * part0 (ENERGY; the caller scales by T / batch size); NULL = every step. */
int bn_st_sites(const bn_st_dims *dims, const bn_lik_desc *lik, int32_t method, int32_t mode, int32_t gh_points,
                double lr, int32_t ensure_psd, const float *y, const double *B, const double *cdiag,
                const double *post_mean, const double *post_cov, const float *batch, double *nat1, double *nat2,
                double *Lw, double *zw, double *ldl, double *part0, double *part1, float *out_mean, float *out_jac,
                float *out_hess, void *stream);

/* Power EP on the space-time models (ExpectationPropagation, inference.py:238-325, with
 * the cavity of the projected sites, basemodels.py:230-262 and 667-674).  Per (run,
 * step): the cavity of the inducing states, cav_cov = ((post_cov + 1e-8 I)^-1 - power
 * B' diag(nat2) B)^-1 and cav_mean = cav_cov ((post_cov + 1e-8 I)^-1 post_mean - power
 * B' nat1), written to cav_mean [K,T,M] / cav_cov [K,T,M,M] (B' diag(nat2) B and B'
 * nat1 are taken from the current projection Lw, zw, which must match nat1 / nat2);
 * then, at every location, the cavity at the data (B cav_mean, diag(B cav_cov B') +
 * diag C) and the moment match of the likelihood.
 * mode BN_ST_UPDATE: new sites damped by lr, part0/part1 = per-step sums of |dnat1|,
 * |dnat2|, and the new projection Lw, zw, ldl (as bn_st_sites).  The reference
 * moment-matches all R locations of a step jointly; for the Gaussian likelihood that
 * closed form separates per location (the joint scale factor times the joint
 * derivatives is -I / sigma^2), so R > 1 takes the Gaussian likelihood only, R = 1
 * any likelihood.  Missing outputs: masked pseudo observations when M == R
 * (basemodels.py:138-144), otherwise the moment match at y = the cavity mean.
 * mode BN_ST_ENERGY (R = 1 only: the reference's EP energy refuses several locations
 * per step, inference.py:313-317): part0 = per-step log Z of the moment match
 * (0 where y is missing), part1 = per-step lel_pseudo = log N(pseudo_y | cav_mean,
 * pseudo_var / power + cav_cov) + pep_constant(pseudo_var, power) in whitened form.
 * The energy is -(log Z_filter + (sum part0 - sum part1) / power).
 * power in (0, 1].  out_* optional [K,T,R]: cavity mean at the data, jacobian, hessian. */
int bn_st_ep_sites(const bn_st_dims *dims, const bn_lik_desc *lik, int32_t mode, int32_t gh_points, double lr,
                   int32_t ensure_psd, double power, const float *y, const double *B, const double *cdiag,
                   const double *post_mean, const double *post_cov, double *nat1, double *nat2, double *Lw,
                   double *zw, double *ldl, double *cav_mean, double *cav_cov, double *part0, double *part1,
                   float *out_mean, float *out_jac, float *out_hess, void *stream);

/* Device scratch of bn_st_filter / bn_st_smoother (bytes). */
int64_t bn_st_workspace_bytes(const bn_st_dims *dims);

/* Kalman filter of the projected pseudo observations (kalman_filter,
 * ops.py:256-285, with y = (Lw Lw')^-1 Lw zw and noise (Lw Lw')^-1), one
 * workgroup per run.  dt [K,T] with dt[k,0] = 0.  Outputs: filtered means
 * filter_mean [K,T,D], filtered covariances filter_cov [K,T,D(D+1)/2] (upper
 * triangle by rows), per-step log N terms ell_steps [K,T] and their sums loglik [K]. */
int bn_st_filter(const bn_st_dims *dims, const bn_st_hypers *hyp, const double *dt, const double *Lw,
                 const double *zw, const double *ldl, double *filter_mean, double *filter_cov, double *ell_steps,
                 double *loglik, void *workspace, void *stream);

/* Smoothed posterior H m, H P H' at every step (rauch_tung_striebel_smoother,
 * ops.py:288-311, 357-380, as used by update_posterior, basemodels.py:689-706,
 * with the next step's dt) from the filter's outputs.  post_mean [K,T,M],
 * post_cov [K,T,M,M]. */
int bn_st_smoother(const bn_st_dims *dims, const bn_st_hypers *hyp, const double *dt, const double *filter_mean,
                   const double *filter_cov, double *post_mean, double *post_cov, void *workspace, void *stream);

/* bn_st_smoother with the full smoothed state as well (return_full, ops.py:294-305 as used by
 * MarkovGaussianProcess.predict, basemodels.py:787-793): smooth_mean [K,T,D], smooth_cov
 * [K,T,D,D] and smooth_cross [K,T,D,D] = G_t P_s,t+1 = Cov(x_t, x_t+1 | y) (gain times the
 * next smoothed covariance, as temporal_conditional joins neighbours, utils.py:132-137);
 * at the last step G = I (the reference's gain for a zero step) and smooth_cross = Pinf.
 * The three are given together or not at all. */
int bn_st_smoother_full(const bn_st_dims *dims, const bn_st_hypers *hyp, const double *dt, const double *filter_mean,
                        const double *filter_cov, double *post_mean, double *post_cov, double *smooth_mean,
                        double *smooth_cov, double *smooth_cross, void *workspace, void *stream);

/* Space-time prediction (MarkovGaussianProcess.predict, basemodels.py:766-816): for every
 * test time t_test [K,n_test] the state conditional on its two neighbouring smoothed states
 * (temporal_conditional, utils.py:125-141 with compute_conditional_statistics 185-215;
 * sentinels (0, Pinf) at -+1e10), then f(t*, r) = B H x(t*) + N(0, C) at n_loc_test
 * locations: B_test [K,n_test,n_loc_test,M] and cdiag_test from bn_st_conditional.
 * t_train [K,T].  Outputs mean, var [K,n_test,n_loc_test] (the diagonal of W P W' + C).
 * steady_cov [K,D,D] (may be NULL): the infinite-horizon form
 * (temporal_conditional_infinite_horizon, utils.py:163-177), whose state covariance is
 * the smoother's steady one at every test time; smooth_cov / smooth_cross are then unused. */
int bn_st_predict(const bn_st_dims *dims, const bn_st_hypers *hyp, const double *t_train, int32_t n_test,
                  int32_t n_loc_test, const double *t_test, const double *B_test, const double *cdiag_test,
                  const double *smooth_mean, const double *smooth_cov, const double *smooth_cross,
                  const double *steady_cov, double *mean, double *var, void *stream);

/* per_run[k] = sum_t per_step[k, t] (FP64, fixed order). */
int bn_st_sum(int32_t n_runs, int32_t n_steps, const double *per_step, double *per_run, void *stream);

/* q(f) at every data location: mean_f = B m, var_f = diag(B P B') + diag C, [K,T,R]
 * (conditional_posterior_to_data, basemodels.py:743-764, diagonal). */
int bn_st_marginals(const bn_st_dims *dims, const double *B, const double *cdiag, const double *post_mean,
                    const double *post_cov, double *mean_f, double *var_f, void *stream);

/* --- mean-field space-time filter and smoother: kalman_filter_meanfield /
 * rauch_tung_striebel_smoother_meanfield (ops.py:581-608, 681-706) as used by
 * MarkovMeanFieldGaussianProcess (basemodels.py:1158-1175).  The state is M independent
 * temporal blocks of size KT (stationary_covariance_meanfield / state_transition_meanfield,
 * kernels.py:526-561); each update keeps only the diagonal blocks of the covariance.
 * Layouts: y [K,T,M], noise [K,T,M,M] (symmetric), block means [K,T,M,KT], block
 * covariances [K,T,M,KT,KT].  dims->n_loc is not used (any value >= 1). */

/* Device scratch of bn_mf_filter's associative form (bytes). */
int64_t bn_mf_workspace_bytes(const bn_st_dims *dims);

/* Pseudo observation of compute_full_pseudo_lik (basemodels.py:676-687) from bn_st_sites'
 * whitened form: noise = (Lw Lw')^-1, y = noise Lw zw. */
int bn_mf_pseudo(const bn_st_dims *dims, const double *Lw, const double *zw, double *noise, double *y,
                 void *stream);

/* Mean-field Kalman filter.  parallel = 0: the sequential form (_sequential_kf_mf,
 * ops.py:429-465); parallel = 1: the associative-scan form (_parallel_kf_mf, ops.py:468-578),
 * a different approximation when the noise couples the blocks.  dt [K,T] (first step of a
 * run: its prior is the stationary state whatever dt[k,0]).  Outputs: filter_mean,
 * filter_cov, per-step log N terms ell_steps [K,T] and their sums loglik [K].  No mask:
 * the space-time models filter unmasked pseudo observations. */
int bn_mf_filter(const bn_st_dims *dims, const bn_st_hypers *hyp, int32_t parallel, const double *dt,
                 const double *y, const double *noise, double *filter_mean, double *filter_cov, double *ell_steps,
                 double *loglik, void *workspace, void *stream);

/* Mean-field RTS smoother (_sequential_rts_mf ops.py:611-641; _parallel_rts_mf 644-678 gives
 * the same result).  dt [K,T] is the smoother's step vector (update_posterior passes the next
 * step's dt and 0 at the end, basemodels.py:700).  Optional outputs: post_mean [K,T,M] = h m_s,
 * h P_s h' of block i at step kt written to post_var[kt*var_step_stride + i*var_elem_stride]
 * (stride M*M and M+1 fill the diagonal of a [K,T,M,M] array), the full smoothed blocks
 * smooth_mean / smooth_cov and the gains [K,T,M,KT,KT] (return_full). */
int bn_mf_smoother(const bn_st_dims *dims, const bn_st_hypers *hyp, const double *dt, const double *filter_mean,
                   const double *filter_cov, double *post_mean, double *post_var, int64_t var_step_stride,
                   int64_t var_elem_stride, double *smooth_mean, double *smooth_cov, double *gains, void *stream);

/* --- infinite-horizon space-time filter and smoother: kalman_filter_infinite_horizon /
 * rauch_tung_striebel_smoother_infinite_horizon (ops.py:796-1068) as used by
 * InfiniteHorizonGaussianProcess (basemodels.py:1257-1301).  Equispaced steps: one step
 * size per run (step [K]).  Dense state of dimension D = KT*M in the layout of bn_st_filter
 * (index c*M + m, H picks the first M entries); covariances and gains [K,D,D]. */

/* Device scratch of bn_ih_filter (and bn_ih_smoother, which needs less), bytes. */
int64_t bn_ih_workspace_bytes(const bn_st_dims *dims, int32_t heteroscedastic);

/* Tied pseudo noise of InfiniteHorizonGaussianProcess.filter (basemodels.py:1276-1285):
 * noise_tied [K,M,M] = (mean_t B' diag(nat2_t) B)^-1 from bn_st_sites' whitened Lw.  workspace:
 * bn_ih_workspace_bytes (the same buffer bn_ih_filter uses; its contents are not kept). */
int bn_ih_tied(const bn_st_dims *dims, const double *Lw, double *noise_tied, void *workspace, void *stream);

/* Infinite-horizon Kalman filter (ops.py:877-953).  dare [K,D,D]: the Riccati solution's
 * start (dare_init) on entry, after `iters` iterations of dare() (ops.py:796-823) on exit
 * (Pdare).  filter_cov [K,D,D] = Pdare - K H Pdare.  heteroscedastic = 1: per-step gains
 * from S_t = H Pdare H' + noise_t (noise [K,T,M,M]); 0: the tied gain for every step (noise
 * unused).  y [K,T,M]; filter_mean [K,T,D]; ell_steps [K,T] log N(y_t | H A m_{t-1}, S_t);
 * loglik [K].  The associative form composes the same maps (ops.py:849-874). */
int bn_ih_filter(const bn_st_dims *dims, const bn_st_hypers *hyp, const double *step, int32_t iters,
                 int32_t heteroscedastic, const double *y, const double *noise, const double *noise_tied, double *dare,
                 double *filter_cov, double *filter_mean, double *ell_steps, double *loglik, void *workspace,
                 void *stream);

/* Infinite-horizon RTS smoother (ops.py:1018-1068): gain G = filter_cov A' Pdare^-1 (optional
 * output [K,D,D]); dare_smoother [K,D,D]: start on entry, after `iters` iterations of
 * rts_dare (ops.py:956-973) on exit; means sm_t = fm_t + G (sm_{t+1} - A fm_t) from sm =
 * fm_{T-1}.  Optional outputs: post_mean [K,T,M] = H sm_t, post_cov [K,T,M,M] = H X_s H' at
 * every step, smooth_mean [K,T,D]. */
int bn_ih_smoother(const bn_st_dims *dims, const bn_st_hypers *hyp, const double *step, int32_t iters,
                   const double *filter_mean, const double *dare, const double *filter_cov, double *dare_smoother,
                   double *gain, double *post_mean, double *post_cov, double *smooth_mean, void *workspace,
                   void *stream);

/* --- learned warm start of the space-time sites.  No counterpart in the reference:
 * it replaces the fixed site initialisation N(0, 1e2 I) (basemodels.py:130-133) via
 * update_nat_params (basemodels.py:97-100); the plain iteration then runs to the same
 * convergence test.  A 3 x 128 SiLU MLP (BF16 weights, FP32 accumulation) maps
 * BN_WS_FEATURES features per site to the converged marginal (m, log v) of f there.
 * Parameters: flat FP32 array of BN_WS_PARAMS (W1[F][128], b1, W2[128][128], b2,
 * W3[128][128], b3, W4[128][2], b4) and its BF16 copy (same layout).  norm: 2F + 4
 * floats = feature means, feature std devs, output means, output std devs. */
#define BN_WS_FEATURES 8
#define BN_WS_PARAMS 34434

/* Features of every site [K,T,R,F]: log(1 + y), three kernel-smoothed log rates over a
 * +-6 step window (bandwidths (0.07, 1), (0.15, 2.5), (0.3, 6) in space and time),
 * and the logs of ls_t, ls_s, var_t, binsize.  t [K,T] (FP64 times). */
int bn_ws_features(const bn_st_dims *dims, const float *r, const float *y, const double *t, double ls_t,
                   double ls_s, double var_t, double binsize, float *X, void *stream);

/* out[n, 2] = MLP(X[n, F]) on the original output scale. */
int bn_mlp_forward(int64_t n_rows, const float *X, const uint16_t *w_bf16, const float *params, const float *norm,
                   float *out, void *stream);

/* One mini-batch of L = sum (out - target)^2 / (2 n) on standardised outputs: gradient
 * (optional, grad[BN_WS_PARAMS]), loss (optional, one float) and, for step >= 1, one
 * Adam step (lr, beta1 0.9, beta2 0.999, eps 1e-8) on params with the BF16 copy
 * refreshed.  step = 0: gradient only. */
int64_t bn_mlp_grad_workspace_bytes(int64_t n_rows);
int bn_mlp_grad_step(int64_t n_rows, const float *X, const float *target, float *params, uint16_t *w_bf16,
                     const float *norm, float *adam_m, float *adam_v, int32_t step, float lr, float *grad,
                     float *loss, void *workspace, void *stream);
int bn_mlp_to_bf16(const float *params, uint16_t *w_bf16, void *stream);

/* Initial sites from predicted marginals pred[n] = (m, log v): the VI site step at
 * q(f) = N(m, v) (inference.py:160-195, 21-39).  nat1 / nat2 [n] FP64. */
int bn_ws_sites(int64_t n_sites, const bn_lik_desc *lik, const float *y, const float *pred, double *nat1,
                double *nat2, void *stream);

/* --- measurement hooks (no counterpart in the reference).  When enabled, each
 * launch of the chunked filter/smoother kernels is bracketed by hipEvents on
 * its own stream; bn_timing_read sums the elapsed time of one kernel id.
 * Not thread-safe; for benchmarks only.  bn_timing_enable clears records. */
int bn_timing_enable(int32_t enable);
int bn_timing_read(int32_t kernel_id, double *total_ms, int32_t *count);
const char *bn_kernel_name(int32_t kernel_id);
int32_t bn_num_kernel_ids(void);

#ifdef __cplusplus
}
#endif

#endif /* BN_HIP_H */
