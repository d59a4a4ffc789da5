// Per-site maps of the Markov-GP inference step (FP64, __host__ __device__):
// likelihood expectations and derivatives for VI, Newton (Laplace) and EP,
// the Newton natural-parameter update with damping, and the energy terms.
//
//   VI      inference.py:170-195, likelihoods.py:363-383, cubature.py:198-246
//   Newton  inference.py:105-128, likelihoods.py:326-355 (analytic derivatives
//           replace jacrev)
//   EP      inference.py:238-284, basemodels.py:230-245, utils.py:534-541,
//           likelihoods.py:401-412, cubature.py:310-371
//   update  inference.py:21-39 (newton_update), 79-86 (damping, diffs),
//           utils.py:89-96 (ensure_diagonal_positive_precision)
//   energy  inference.py:130-154, 197-222, 286-325; basemodels.py:247-262;
//           utils.py:431-445 (pep_constant), 510-531 (gaussian_expected_log_lik)
// Sites are scalar (one per time step).  A NaN observation is a missing one.
#pragma once
#include "bn_core.h"
#include "../../include/bn_hip.h"

namespace bn {

enum : int32_t { LIK_GAUSSIAN = 1, LIK_BERNOULLI_PROBIT = 2, LIK_POISSON = 3, LIK_STUDENTS_T = 4 };
enum : int32_t { M_VI = 0, M_NEWTON = 1, M_EP = 2 };

static constexpr int kMaxGH = 64;
struct GHTable {
  int32_t n;
  double x[kMaxGH];  // nodes for E_{z~N(0,1)}: sqrt(2) * Hermite roots
  double w[kMaxGH];  // weights / sqrt(pi)
};

struct LikParams {
  int32_t kind;
  double p0, p1;  // Gaussian: variance | Student-t: scale, df | Poisson: binsize
  double c0;      // Student-t log-density constant
  double t0;      // Student-t: 1 / (scale^2 df), so that a node costs no division
};

inline LikParams make_lik(int32_t kind, double p0, double p1) {
  LikParams L{kind, p0, p1, 0.0, 0.0};
  if (kind == LIK_STUDENTS_T) {
    const double df = p1, s = p0;
    L.c0 = lgamma((df + 1.0) * 0.5) - lgamma(df * 0.5) - 0.5 * (log(s * s) + log(df) + log(M_PI));
    L.t0 = 1.0 / (s * s * df);
  }
  return L;
}

static constexpr double kProbitEps = 1e-3;  // likelihoods.py:828

BN_HD double norm_cdf(double x) { return 0.5 * erfc(-x * 0.70710678118654752440); }

// log p(y | f) (likelihoods.py:708-720, 833-852, 941-946, 1031-1041)
BN_HD double log_lik(const LikParams& L, double y, double f) {
  switch (L.kind) {
    case LIK_GAUSSIAN:
      return -0.5 * log(2.0 * M_PI * L.p0) - 0.5 * (y - f) * (y - f) / L.p0;
    case LIK_BERNOULLI_PROBIT: {
      const double p = norm_cdf(f) * (1.0 - 2.0 * kProbitEps) + kProbitEps;
      return log_fast(y == 1.0 ? p : 1.0 - p);
    }
    case LIK_POISSON: {
      const double mu = exp(f) * L.p0;
      return y * log(mu) - mu - lgamma(y + 1.0);
    }
    default: {
      const double r = y - f;
      return L.c0 - 0.5 * (L.p1 + 1.0) * log_fast(1.0 + r * r * L.t0);
    }
  }
}

// log p, d/df, d2/df2 in closed form (replaces jacrev, likelihoods.py:326-335)
BN_HD void log_lik_derivs(const LikParams& L, double y, double f, double& ll, double& J, double& H) {
  switch (L.kind) {
    case LIK_GAUSSIAN:
      ll = log_lik(L, y, f);
      J = (y - f) / L.p0;
      H = -1.0 / L.p0;
      return;
    case LIK_BERNOULLI_PROBIT: {
      const double s = (y == 1.0) ? 1.0 : -1.0;
      const double phi = exp(-0.5 * f * f) * 0.39894228040143267794;
      const double p = kProbitEps + (1.0 - 2.0 * kProbitEps) * norm_cdf(s * f);
      ll = log(p);
      J = s * (1.0 - 2.0 * kProbitEps) * phi / p;
      H = -(1.0 - 2.0 * kProbitEps) * s * f * phi / p - J * J;
      return;
    }
    case LIK_POISSON: {
      const double mu = exp(f) * L.p0;
      ll = y * log(mu) - mu - lgamma(y + 1.0);
      J = y - mu;
      H = -mu;
      return;
    }
    default: {
      const double r = y - f, s2 = L.p0 * L.p0, df = L.p1;
      const double den = df * s2 + r * r;
      ll = log_lik(L, y, f);
      J = (df + 1.0) * r / den;
      H = (df + 1.0) * (r * r - df * s2) / (den * den);
      return;
    }
  }
}

// Variational expectation E_q[log p], dE/dm, d2E/dm2 (closed forms for
// Gaussian likelihoods.py:744-753 and Poisson 996-1007, else GH cubature)
BN_HD void var_exp(const LikParams& L, const GHTable& gh, double y, double m, double v, double& E, double& dE,
                   double& d2E) {
  if (L.kind == LIK_GAUSSIAN) {
    E = -0.5 * log(2.0 * M_PI) - 0.5 * log(L.p0) - 0.5 * ((y - m) * (y - m) + v) / L.p0;
    dE = (y - m) / L.p0;
    d2E = -1.0 / L.p0;
    return;
  }
  if (L.kind == LIK_POISSON) {
    const double emc = L.p0 * exp(m + 0.5 * v);
    E = y * log(L.p0) + y * m - emc - lgamma(y + 1.0);
    dE = y - emc;
    d2E = -emc;
    return;
  }
  const double sv = sqrt(v), iv = 1.0 / v;
  double e = 0.0, d1 = 0.0, dv = 0.0;
  for (int i = 0; i < gh.n; ++i) {
    const double d = sv * gh.x[i];
    const double wl = gh.w[i] * log_lik(L, y, m + d);
    e += wl;
    d1 += iv * d * wl;
    dv += (0.5 * iv * iv * d * d - 0.5 * iv) * wl;
  }
  E = e;
  dE = d1;
  d2E = 2.0 * dv;
}

// Moment matching log Z, dlogZ/dm, d2logZ/dm2 of the tilted distribution
// (Gaussian closed form likelihoods.py:771-782 with pep_constant; else GH
// cubature with Z clipped at 1e-8, cubature.py:340-371)
BN_HD void moment_match(const LikParams& L, const GHTable& gh, double y, double m, double v, double power,
                        double& lZ, double& dlZ, double& d2lZ) {
  if (L.kind == LIK_GAUSSIAN) {
    const double var = L.p0 / power + v;
    const double r = y - m;
    lZ = -0.5 * (kLog2Pi + log(var) + r * r / var) +
         0.5 * ((1.0 - power) * kLog2Pi - log(power)) + 0.5 * (1.0 - power) * log(L.p0);
    dlZ = r / var;
    d2lZ = -1.0 / var;
    return;
  }
  const double sv = sqrt(v), iv = 1.0 / v;
  double Z = 0.0, dZ = 0.0, d2Z = 0.0;
  for (int i = 0; i < gh.n; ++i) {
    const double d = sv * gh.x[i];
    const double wl = gh.w[i] * exp(power * log_lik(L, y, m + d));
    Z += wl;
    dZ += iv * d * wl;
    d2Z += (iv * iv * d * d - iv) * wl;
  }
  const double Zc = Z > 1e-8 ? Z : 1e-8;
  lZ = log(Zc);
  const double iZ = 1.0 / Zc;
  dlZ = iZ * dZ;
  d2lZ = -dlZ * dlZ + iZ * d2Z;
}

// d log p(y | f) / d(likelihood hyperparameter): Gaussian variance, Student-t
// scale; 0 for the likelihoods without one (Bernoulli, Poisson with fixed binsize).
BN_HD double dlog_lik_dparam(const LikParams& L, double y, double f) {
  if (L.kind == LIK_GAUSSIAN) return -0.5 / L.p0 + 0.5 * (y - f) * (y - f) / (L.p0 * L.p0);
  if (L.kind == LIK_STUDENTS_T) {
    const double r = y - f, s = L.p0, df = L.p1;
    return -1.0 / s + (df + 1.0) * r * r / (s * (df * s * s + r * r));
  }
  return 0.0;
}

// d/d(likelihood hyperparameter) of the moment-matched log Z = log E_cav[p(y | f)^power]
// (moment_match above) at the cavity (m, v): Gaussian closed form, else by the same
// cubature, d log Z = sum w_i p_i^a a d log p_i / Z (0 where Z is clipped at 1e-8).
BN_HD double dmoment_match_dparam(const LikParams& L, const GHTable& gh, double y, double m, double v, double power) {
  if (L.kind == LIK_GAUSSIAN) {
    const double var = L.p0 / power + v;
    const double r = y - m;
    return -0.5 * (1.0 / var - r * r / (var * var)) / power + 0.5 * (1.0 - power) / L.p0;
  }
  const double sv = sqrt(v);
  double Z = 0.0, dZ = 0.0;
  for (int i = 0; i < gh.n; ++i) {
    const double f = m + sv * gh.x[i];
    const double wl = gh.w[i] * exp(power * log_lik(L, y, f));
    Z += wl;
    dZ += wl * power * dlog_lik_dparam(L, y, f);
  }
  return Z > 1e-8 ? dZ / Z : 0.0;
}

// d term1 / d(likelihood hyperparameter) of one site (see energy_terms), with the
// posterior and the sites held fixed: VI E_q[d log p / d phi] (closed form for the
// Gaussian), Newton d log p(y | m) / d phi, EP (1 / power) d log Z / d phi at the
// site's cavity (the lel term of inference.py:286-325; the pseudo terms carry no
// likelihood hyperparameter).  Masked sites contribute 0.
template <int METHOD>
BN_HD double dterm1_dparam(const LikParams& L, const GHTable& gh, double y, double pm, double pv, double n1,
                           double n2, double power) {
  if (y != y) return 0.0;
  if (METHOD == M_NEWTON) return dlog_lik_dparam(L, y, pm);
  if (METHOD == M_EP) {
    // cavity of utils.py:534-541, as site_step forms it
    const double pn2 = 1.0 / (pv + 1e-8);
    const double cv = 1.0 / (pn2 - power * n2);
    const double cm = cv * (pn2 * pm - power * n1);
    return dmoment_match_dparam(L, gh, y, cm, cv, power) / power;
  }
  if (L.kind == LIK_GAUSSIAN) return -0.5 / L.p0 + 0.5 * ((y - pm) * (y - pm) + pv) / (L.p0 * L.p0);
  const double sv = sqrt(pv);
  double e = 0.0;
  for (int i = 0; i < gh.n; ++i) e += gh.w[i] * dlog_lik_dparam(L, y, pm + sv * gh.x[i]);
  return e;
}

struct SiteResult {
  double mean, jac, hess;  // terms returned by inference() (hess after ensure_psd)
  double nat1, nat2;       // newton_update of those terms
};

// The site's new natural parameters for one inference step.
template <int METHOD>
BN_HD SiteResult site_step(const LikParams& L, const GHTable& gh, double y, double pm, double pv, double n1,
                           double n2, double power, bool ensure_psd) {
  const bool miss = y != y;
  double mean, jac, hess;
  if (METHOD == M_VI) {
    double E, dE, d2E;
    var_exp(L, gh, miss ? pm : y, pm, pv, E, dE, d2E);
    mean = pm;
    jac = miss ? NAN : dE;
    hess = miss ? NAN : d2E;
  } else if (METHOD == M_NEWTON) {
    double ll, J, H;
    log_lik_derivs(L, miss ? pm : y, pm, ll, J, H);
    mean = pm;
    jac = miss ? NAN : J;
    hess = miss ? NAN : H;
  } else {
    // cavity (utils.py:534-541): post_nat2 = (v + 1e-8)^-1
    const double pn2 = 1.0 / (pv + 1e-8);
    const double cv = 1.0 / (pn2 - power * n2);
    const double cm = cv * (pn2 * pm - power * n1);
    double lZ, dlZ, d2lZ;
    moment_match(L, gh, miss ? cm : y, cm, cv, power, lZ, dlZ, d2lZ);
    // scale_factor = cav_prec (d2lZ + cav_prec)^-1 / power (inference.py:263-265)
    const double cp = 1.0 / cv;
    const double sc = cp / (d2lZ + cp) / power;
    mean = cm;
    jac = miss ? NAN : sc * dlZ;
    hess = miss ? NAN : sc * d2lZ;
  }
  // ensure_diagonal_positive_precision(-hess): -hess < 0 -> 1e-2
  if (ensure_psd && -hess < 0.0) hess = -1e-2;
  // newton_update (inference.py:21-39)
  const double H = (hess != hess) ? -1e-6 : hess;
  const double J = (jac != jac) ? H * mean : jac;
  return SiteResult{mean, jac, hess, J - H * mean, -H};
}

// EP site step from a cavity already formed (cm, cv): the space-time models remove
// the projected sites from the posterior of the inducing states and carry the cavity
// to the data (basemodels.py:230-245, 743-764) before moment matching
// (inference.py:238-284).  mask_pseudo: missing outputs give NaN terms (the model
// masks its pseudo observations); otherwise the moment match runs at y = cm, as
// likelihoods.py:401-412 substitutes, and its terms are kept.
BN_HD SiteResult ep_site_from_cavity(const LikParams& L, const GHTable& gh, double y, double cm, double cv,
                                     double power, bool mask_pseudo, bool ensure_psd) {
  const bool miss = y != y;
  double lZ, dlZ, d2lZ;
  moment_match(L, gh, miss ? cm : y, cm, cv, power, lZ, dlZ, d2lZ);
  const double cp = 1.0 / cv;
  const double sc = cp / (d2lZ + cp) / power;
  double jac = sc * dlZ, hess = sc * d2lZ;
  if (miss && mask_pseudo) { jac = NAN; hess = NAN; }
  if (ensure_psd && -hess < 0.0) hess = -1e-2;
  const double H = (hess != hess) ? -1e-6 : hess;
  const double J = (jac != jac) ? H * cm : jac;
  return SiteResult{cm, jac, hess, J - H * cm, -H};
}

// log Z of the moment match at a formed cavity (the lel term of inference.py:286-325);
// a missing output contributes 0.
BN_HD double ep_lel_from_cavity(const LikParams& L, const GHTable& gh, double y, double cm, double cv,
                                double power) {
  if (y != y) return 0.0;
  double lZ, dlZ, d2lZ;
  moment_match(L, gh, y, cm, cv, power, lZ, dlZ, d2lZ);
  return lZ;
}

// Energy terms of one site (summed per series by the caller):
//   VI      t1 = E_q[log p(y|f)]             t2 = E_q[log N(py | f, pr)]
//   Newton  t1 = log p(y | m)                t2 = E_q[log N(py | f, pr)]
//   EP      t1 = log Z (cavity, masked 0)    t2 = log N(py | cm, pr/power + cv) + pep_constant(pr, power)
// with the pseudo site (py, pr) = (n1/n2, 1/n2).  Masked sites contribute 0.
template <int METHOD>
BN_HD void energy_terms(const LikParams& L, const GHTable& gh, double y, double pm, double pv, double n1, double n2,
                        double power, double& t1, double& t2) {
  const bool miss = y != y;
  const double pr = 1.0 / n2, py = n1 / n2;
  if (METHOD == M_EP) {
    const double pn2 = 1.0 / (pv + 1e-8);
    const double cv = 1.0 / (pn2 - power * n2);
    const double cm = cv * (pn2 * pm - power * n1);
    double lZ, dlZ, d2lZ;
    moment_match(L, gh, miss ? cm : y, cm, cv, power, lZ, dlZ, d2lZ);
    t1 = miss ? 0.0 : lZ;
    const double var = pr / power + cv;
    const double r = py - cm;
    const double lp = -0.5 * (r * r / var + kLog2Pi + log(var));
    const double pep = 0.5 * ((1.0 - power) * kLog2Pi - log(power)) + 0.5 * (1.0 - power) * log(pr);
    t2 = miss ? 0.0 : lp + pep;
    return;
  }
  if (METHOD == M_VI) {
    double E, dE, d2E;
    var_exp(L, gh, miss ? pm : y, pm, pv, E, dE, d2E);
    t1 = miss ? 0.0 : E;
  } else {
    double ll, J, H;
    log_lik_derivs(L, miss ? pm : y, pm, ll, J, H);
    t1 = miss ? 0.0 : ll;
  }
  // gaussian_expected_log_lik (utils.py:510-531); masked: noise 1/(2 pi), q_mu = y, q_cov 1e-20
  if (miss) {
    t2 = -0.5 * 1e-20 * 2.0 * M_PI;
  } else {
    const double r = py - pm;
    t2 = -0.5 * (kLog2Pi + log(pr) + r * r / pr) - 0.5 * pv / pr;
  }
}

// Host helpers shared by the translation units that launch site kernels:
// Gauss-Hermite table for n nodes, and a validated likelihood descriptor.
int make_gh_table(int n, GHTable& gh);
int check_lik(const bn_lik_desc* d, LikParams& L);

}  // namespace bn
