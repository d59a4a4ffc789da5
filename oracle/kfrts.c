/*
 * FP64 sequential Kalman filter + RTS smoother, C restatement of the oracle.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/ as a fast checker for long series and
 * by bench.py's cpu_baseline leg.  Never linked into bayesnewton_amd.
 *
 * Follows the reference's dense recurrences exactly:
 *   filter   bayesnewton/ops.py:154-180 (_sequential_kf) with kalman_filter 256-285
 *   smoother bayesnewton/ops.py:288-311 (_sequential_rts) with the shifted step
 *            vector of basemodels.py:700 (smoother step n uses dt[n+1], last 0)
 *   log N    bayesnewton/utils.py:376-396 (mask -> cov 1/(2 pi), contributes 0)
 *   kernels  Matern12/32/52 closed-form A(dt), Pinf (kernels.py:123-300),
 *            Sum = block-diagonal state, concatenated H (kernels.py:1624-1663)
 *   Q = Pinf - A Pinf A'  (ops.py:149-151)
 * Observations are scalar per step (p = 1), the only case of the 1-D models.
 * Series are independent (CSR offsets); OpenMP runs them in parallel.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXD 12
#define LOG2PI 1.8378770664093453

typedef struct {
    int n_comp;
    int kind[4];      /* 1 = Matern12, 2 = Matern32, 3 = Matern52 (= state dim) */
    double ls[4];
    double var[4];
    int d;
} model_t;

static void comp_transition(int kind, double ls, double dt, double *A /*kind x kind*/) {
    if (kind == 1) {
        A[0] = exp(-dt / ls);
    } else if (kind == 2) {
        double lam = sqrt(3.0) / ls, e = exp(-dt * lam);
        A[0] = e * (dt * lam + 1.0);  A[1] = e * dt;
        A[2] = e * (-dt * lam * lam); A[3] = e * (-dt * lam + 1.0);
    } else {
        double lam = sqrt(5.0) / ls, dl = dt * lam, e = exp(-dl);
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

static void comp_pinf(int kind, double ls, double var, double *P) {
    if (kind == 1) {
        P[0] = var;
    } else if (kind == 2) {
        P[0] = var; P[1] = 0; P[2] = 0; P[3] = 3.0 * var / (ls * ls);
    } else {
        double kap = 5.0 / 3.0 * var / (ls * ls);
        P[0] = var; P[1] = 0; P[2] = -kap;
        P[3] = 0; P[4] = kap; P[5] = 0;
        P[6] = -kap; P[7] = 0; P[8] = 25.0 * var / (ls * ls * ls * ls);
    }
}

static void model_pinf(const model_t *m, double *P) {
    int d = m->d, off = 0;
    memset(P, 0, sizeof(double) * d * d);
    for (int c = 0; c < m->n_comp; ++c) {
        double blk[9];
        int k = m->kind[c];
        comp_pinf(k, m->ls[c], m->var[c], blk);
        for (int i = 0; i < k; ++i)
            for (int j = 0; j < k; ++j) P[(off + i) * d + off + j] = blk[i * k + j];
        off += k;
    }
}

static void model_transition(const model_t *m, double dt, double *A) {
    int d = m->d, off = 0;
    memset(A, 0, sizeof(double) * d * d);
    for (int c = 0; c < m->n_comp; ++c) {
        double blk[9];
        int k = m->kind[c];
        comp_transition(k, m->ls[c], dt, blk);
        for (int i = 0; i < k; ++i)
            for (int j = 0; j < k; ++j) A[(off + i) * d + off + j] = blk[i * k + j];
        off += k;
    }
}

/* C = A B (d x d) */
static void mm(int d, const double *A, const double *B, double *C) {
    for (int i = 0; i < d; ++i)
        for (int j = 0; j < d; ++j) {
            double s = 0;
            for (int k = 0; k < d; ++k) s += A[i * d + k] * B[k * d + j];
            C[i * d + j] = s;
        }
}

/* C = A B' */
static void mmt(int d, const double *A, const double *B, double *C) {
    for (int i = 0; i < d; ++i)
        for (int j = 0; j < d; ++j) {
            double s = 0;
            for (int k = 0; k < d; ++k) s += A[i * d + k] * B[j * d + k];
            C[i * d + j] = s;
        }
}

/* A <- chol(A) lower, in place; returns 0 on success */
static int chol(int d, double *A) {
    for (int j = 0; j < d; ++j) {
        double s = A[j * d + j];
        for (int k = 0; k < j; ++k) s -= A[j * d + k] * A[j * d + k];
        if (!(s > 0)) return -1;
        double l = sqrt(s);
        A[j * d + j] = l;
        for (int i = j + 1; i < d; ++i) {
            double t = A[i * d + j];
            for (int k = 0; k < j; ++k) t -= A[i * d + k] * A[j * d + k];
            A[i * d + j] = t / l;
        }
        for (int i = 0; i < j; ++i) A[i * d + j] = 0.0;
    }
    return 0;
}

/* X <- L^-T L^-1 X for nrhs columns stored row-major d x nrhs */
static void chol_solve(int d, const double *L, double *X, int nrhs) {
    for (int c = 0; c < nrhs; ++c) {
        for (int i = 0; i < d; ++i) {
            double s = X[i * nrhs + c];
            for (int k = 0; k < i; ++k) s -= L[i * d + k] * X[k * nrhs + c];
            X[i * nrhs + c] = s / L[i * d + i];
        }
        for (int i = d - 1; i >= 0; --i) {
            double s = X[i * nrhs + c];
            for (int k = i + 1; k < d; ++k) s -= L[k * d + i] * X[k * nrhs + c];
            X[i * nrhs + c] = s / L[i * d + i];
        }
    }
}

static void run_series(const model_t *mdl, int64_t n, const double *dt, const double *y, const double *r,
                       const uint8_t *mask, double *post_mean, double *post_var, double *loglik,
                       double *fm, double *fP) {
    int d = mdl->d;
    double Pinf[MAXD * MAXD], A[MAXD * MAXD], Q[MAXD * MAXD], T1[MAXD * MAXD], T2[MAXD * MAXD];
    double m[MAXD], P[MAXD * MAXD], mp[MAXD], Pp[MAXD * MAXD];
    int H[MAXD];
    int off = 0;
    for (int i = 0; i < d; ++i) H[i] = 0;
    for (int c = 0; c < mdl->n_comp; ++c) { H[off] = 1; off += mdl->kind[c]; }
    model_pinf(mdl, Pinf);
    for (int i = 0; i < d; ++i) m[i] = 0.0;
    memcpy(P, Pinf, sizeof(double) * d * d);
    double ell = 0.0;
    /* ---- filter: ops.py:156-175 ---- */
    for (int64_t t = 0; t < n; ++t) {
        model_transition(mdl, dt[t], A);
        mm(d, A, Pinf, T1); mmt(d, T1, A, T2);
        for (int i = 0; i < d * d; ++i) Q[i] = Pinf[i] - T2[i];
        for (int i = 0; i < d; ++i) {
            double s = 0;
            for (int k = 0; k < d; ++k) s += A[i * d + k] * m[k];
            mp[i] = s;
        }
        mm(d, A, P, T1); mmt(d, T1, A, Pp);
        for (int i = 0; i < d * d; ++i) Pp[i] += Q[i];
        double hm = 0, S = 0, HP[MAXD];
        for (int k = 0; k < d; ++k) if (H[k]) hm += mp[k];
        for (int j = 0; j < d; ++j) {
            double s = 0;
            for (int k = 0; k < d; ++k) if (H[k]) s += Pp[k * d + j];
            HP[j] = s;
        }
        for (int j = 0; j < d; ++j) if (H[j]) S += HP[j];
        S += r[t];
        double v = y[t] - hm;
        if (!(mask && mask[t])) ell += -0.5 * (v * v / S + LOG2PI + log(S));
        for (int i = 0; i < d; ++i) m[i] = mp[i] + HP[i] / S * v;
        for (int i = 0; i < d; ++i)
            for (int j = 0; j < d; ++j) P[i * d + j] = Pp[i * d + j] - HP[i] / S * HP[j];
        memcpy(fm + t * d, m, sizeof(double) * d);
        memcpy(fP + t * d * d, P, sizeof(double) * d * d);
    }
    *loglik = ell;
    /* ---- smoother: ops.py:290-305, smoother step t uses dt[t+1] (last: 0) ---- */
    double sm[MAXD], sP[MAXD * MAXD], G[MAXD * MAXD];
    memcpy(sm, fm + (n - 1) * d, sizeof(double) * d);
    memcpy(sP, fP + (n - 1) * d * d, sizeof(double) * d * d);
    for (int64_t t = n - 1; t >= 0; --t) {
        const double *fmt = fm + t * d, *fPt = fP + t * d * d;
        double dts = (t + 1 < n) ? dt[t + 1] : 0.0;
        model_transition(mdl, dts, A);
        mm(d, A, Pinf, T1); mmt(d, T1, A, T2);
        for (int i = 0; i < d * d; ++i) Q[i] = Pinf[i] - T2[i];
        double pm[MAXD];
        for (int i = 0; i < d; ++i) {
            double s = 0;
            for (int k = 0; k < d; ++k) s += A[i * d + k] * fmt[k];
            pm[i] = s;
        }
        double AfP[MAXD * MAXD], pP[MAXD * MAXD];
        mm(d, A, fPt, AfP);
        mmt(d, AfP, A, pP);
        for (int i = 0; i < d * d; ++i) pP[i] += Q[i];
        /* C = solve(pP, AfP)' */
        if (chol(d, pP) != 0) {
            for (int i = 0; i < d * d; ++i) G[i] = NAN;
        } else {
            memcpy(T1, AfP, sizeof(double) * d * d);
            chol_solve(d, pP, T1, d);
            for (int i = 0; i < d; ++i)
                for (int j = 0; j < d; ++j) G[i * d + j] = T1[j * d + i];
        }
        /* recompute pP (overwritten by the factor) */
        mmt(d, AfP, A, pP);
        for (int i = 0; i < d * d; ++i) pP[i] += Q[i];
        double nm[MAXD], dm[MAXD];
        for (int i = 0; i < d; ++i) dm[i] = sm[i] - pm[i];
        for (int i = 0; i < d; ++i) {
            double s = 0;
            for (int k = 0; k < d; ++k) s += G[i * d + k] * dm[k];
            nm[i] = fmt[i] + s;
        }
        double D[MAXD * MAXD];
        for (int i = 0; i < d * d; ++i) D[i] = sP[i] - pP[i];
        mm(d, G, D, T1); mmt(d, T1, G, T2);
        for (int i = 0; i < d * d; ++i) sP[i] = fPt[i] + T2[i];
        memcpy(sm, nm, sizeof(double) * d);
        double hm = 0, hv = 0;
        for (int i = 0; i < d; ++i) if (H[i]) {
            hm += sm[i];
            for (int j = 0; j < d; ++j) if (H[j]) hv += sP[i * d + j];
        }
        post_mean[t] = hm;
        post_var[t] = hv;
    }
}

/*
 * kinds[n_comp], ls/var: per-series hypers [B * n_comp] (hyper_stride = n_comp) or shared
 * (hyper_stride = 0).  offsets[B+1] CSR into dt/y/r/mask/post_*.  mask may be NULL.
 * Returns 0, or -1 on bad arguments.
 */
int oracle_filter_smoother(int n_comp, const int *kinds, const double *ls, const double *var, int hyper_stride,
                           const int64_t *offsets, int B, const double *dt, const double *y, const double *r,
                           const uint8_t *mask, double *post_mean, double *post_var, double *loglik,
                           int nthreads) {
    if (n_comp < 1 || n_comp > 4) return -1;
    int d = 0;
    for (int c = 0; c < n_comp; ++c) {
        if (kinds[c] < 1 || kinds[c] > 3) return -1;
        d += kinds[c];
    }
    if (d > MAXD) return -1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    int err = 0;
#pragma omp parallel
    {
        int64_t cap = 0;
        double *fm = NULL, *fP = NULL;
#pragma omp for schedule(dynamic, 1)
        for (int b = 0; b < B; ++b) {
            int64_t n = offsets[b + 1] - offsets[b];
            if (n <= 0) { loglik[b] = 0.0; continue; }
            if (n > cap) {
                free(fm); free(fP);
                cap = n;
                fm = (double *)malloc(sizeof(double) * d * n);
                fP = (double *)malloc(sizeof(double) * d * d * n);
                if (!fm || !fP) { err = -2; cap = 0; continue; }
            }
            model_t mdl;
            mdl.n_comp = n_comp; mdl.d = d;
            for (int c = 0; c < n_comp; ++c) {
                mdl.kind[c] = kinds[c];
                mdl.ls[c] = ls[(int64_t)b * hyper_stride + c];
                mdl.var[c] = var[(int64_t)b * hyper_stride + c];
            }
            int64_t o = offsets[b];
            run_series(&mdl, n, dt + o, y + o, r + o, mask ? mask + o : NULL, post_mean + o, post_var + o,
                       loglik + b, fm, fP);
        }
        free(fm); free(fP);
    }
    return err;
}
