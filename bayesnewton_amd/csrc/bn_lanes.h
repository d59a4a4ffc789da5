// Per-lane bodies of the chunked filter / smoother kernels.
//
// Decomposition (one series = one reference model, bayesnewton/basemodels.py:625-706):
//   series -> chunks of <= chunk_len consecutive steps -> tiles of <= 128 chunks.
// One lane walks one chunk sequentially; chunk summaries are combined with the
// associative operators of ops.py:203-219 (filter) and 328-335 (smoother) by a
// two-level scan (within a tile, then across the tiles of a series).
//
//   filter_reduce   chunk -> filtering element (conditional filter from x_{s-1})
//   [tile scan, series scan of the filtering elements]
//   filter_pass     incoming filtered state -> filtered moments of every step,
//                   log-likelihood partial, and the chunk's smoothing element
//   [tile scan, series scan of the smoothing elements, in reverse]
//   smoother_pass   incoming smoothed state -> smoothed moments of every step
//
// Everything here is __host__ __device__: the HIP kernels call these bodies with
// their lane index, and a host build runs them in plain loops.
#pragma once
#include <type_traits>

#include "bn_core.h"

namespace bn {

enum : uint8_t { CH_FIRST = 1, CH_LAST = 2 };
enum : uint8_t { TL_FIRST = 1, TL_LAST = 2 };
// Per-chunk (then per-series, OR-reduced) numerical status bits.
enum : int32_t { ST_INNOV_NONPD = 1, ST_NONFINITE = 2, ST_SMOOTHER_NONPD = 4 };

// false for NaN and +-Inf (no reliance on the math library's classification)
BN_HD bool finite_d(double x) { return x - x == 0.0; }

struct PlanView {
  int32_t n_series, n_chunks, n_chunks_pad, n_tiles, chunk_len;
  int32_t gshift;  // log2 of the lane-group width (64, or fewer for < 64 chunks)
  int64_t n_slots;
  const int32_t* chunk_series;
  const int32_t* chunk_len_;
  const uint8_t* chunk_flags;
  const int32_t* chunk_tile;
  const int64_t* chunk_start;
  const int32_t* tile_first;
  const int32_t* tile_count;
  const int32_t* tile_series;
  const uint8_t* tile_flags;
  const int32_t* series_first_tile;
  const int32_t* series_ntiles;
  const int32_t* series_first_chunk;
  const int32_t* series_nchunks;
  int32_t n_groups;            // scan groups of the tile scans
  const int32_t* group_tiles;  // kGroupMax tiles per group, -1 for an empty place
  const uint8_t* group_w;      // width of each group: its tiles have at most kTile / w chunks
  const int32_t* wave_maxlen;  // longest chunk among chunks 64w .. 64w + 63

  BN_HD int64_t slot(int32_t j, int32_t k) const {
    return (((int64_t)(j >> gshift) * chunk_len + k) << gshift) + (j & ((1 << gshift) - 1));
  }
};

struct HyperView {
  const double* ls;
  const double* var;
  int32_t stride;
};

template <int K1, int K2>
BN_HD Model<K1, K2> load_model(const HyperView& hv, int32_t b) {
  Model<K1, K2> m;
  int64_t o = (int64_t)b * hv.stride;
  m.init(hv.ls[o], hv.var[o], K2 > 0 ? hv.ls[o + 1] : 1.0, K2 > 0 ? hv.var[o + 1] : 1.0);
  return m;
}

// Storage precision of the chunk/tile elements and of the filtered-moment
// scratch.  FP64 for both: with FP32 storage the posterior variance of a
// Matern32 + Matern52 sum (whose observed variance is a small difference of
// large component entries) drifts to ~1e-4 relative from the FP64 oracle;
// with FP64 storage it stays at ~1e-7.  The macros exist for that experiment.
#ifndef BN_AGG_STORE
#define BN_AGG_STORE double
#endif
#ifndef BN_STATE_STORE
#define BN_STATE_STORE double
#endif
// Storage of the per-step rows the Bryson-Frazier smoother reads (H P-, H m-): FP32.
// They enter the smoothed moments linearly and the chunk-local recursion runs at most
// one chunk from its FP64 anchor, so their rounding stays at the 1e-7 level of the
// FP32 outputs; FP64 rows cost 24 B more per step in each direction.
#ifndef BN_HROW_STORE
#define BN_HROW_STORE float
#endif
using agg_t = BN_AGG_STORE;
using state_t = BN_STATE_STORE;
using hrow_t = BN_HROW_STORE;

// Field-major (structure-of-arrays) element storage: field f of item i at base[f * stride + i].
struct SoA {
  agg_t* base;
  int64_t stride;
};
struct CSoA {
  const agg_t* base;
  int64_t stride;
};

template <class E>
BN_HD void soa_store(const E& e, SoA s, int64_t i) {
  e.store([&](int f, auto v) { s.base[(int64_t)f * s.stride + i] = (agg_t)v; });
}
template <class E>
BN_HD void soa_load(E& e, CSoA s, int64_t i) {
  e.load([&](int f) { return s.base[(int64_t)f * s.stride + i]; });
}

// The earlier operand of fcombine_from read from SoA storage (FElem's packed order:
// Phi, beta, Sig upper, J upper, eta) field group by field group, when the combine's
// phase needs it, instead of whole before the combine.
template <int D, class T>
struct SoaOperand {
  CSoA s;
  int64_t i;
  static constexpr int NS = D * (D + 1) / 2;
  BN_HD T at(int f) const { return (T)s.base[(int64_t)f * s.stride + i]; }
  BN_HD void phi(T P[D][D]) const { BN_UNROLL for (int r = 0; r < D; ++r) BN_UNROLL for (int c = 0; c < D; ++c) P[r][c] = at(r * D + c); }
  BN_HD void beta(T b[D]) const { BN_UNROLL for (int r = 0; r < D; ++r) b[r] = at(D * D + r); }
  BN_HD void sym(int f0, T S[D][D]) const {
    int f = f0;
    BN_UNROLL for (int r = 0; r < D; ++r) BN_UNROLL for (int c = r; c < D; ++c) { const T v = at(f++); S[r][c] = v; S[c][r] = v; }
  }
  BN_HD void sig(T S[D][D]) const { sym(D * D + D, S); }
  BN_HD void jm(T J[D][D]) const { sym(D * D + D + NS, J); }
  BN_HD void eta(T v[D]) const { BN_UNROLL for (int r = 0; r < D; ++r) v[r] = at(D * D + D + 2 * NS + r); }
};

// Where the filtered moments go / come from.  packed: D mean planes then NS
// packed-upper covariance planes; full: D mean planes + D*D covariance planes.
struct StateIO {
  state_t* mean;
  state_t* cov;
  bool packed;
};
struct CStateIO {
  const state_t* mean;
  const state_t* cov;
  bool packed;
};

template <int K1, int K2, class T>
BN_HD void write_state(const StateIO& io, int64_t n_slots, int64_t s, const T m[Spec<K1, K2>::D],
                       const T P[Spec<K1, K2>::D][Spec<K1, K2>::D]) {
  using S = Spec<K1, K2>;
  constexpr int D = S::D;
  BN_UNROLL for (int i = 0; i < D; ++i) io.mean[i * n_slots + s] = (state_t)m[i];
  if (io.packed) {
    BN_UNROLL for (int i = 0; i < D; ++i)
      BN_UNROLL for (int j = i; j < D; ++j) io.cov[S::sidx(i, j) * n_slots + s] = (state_t)P[i][j];
  } else {
    BN_UNROLL for (int i = 0; i < D; ++i)
      BN_UNROLL for (int j = 0; j < D; ++j) io.cov[(i * D + j) * n_slots + s] = (state_t)P[i][j];
  }
}

// write_state from a packed covariance (NS entries, upper triangle by rows).
template <int K1, int K2, class T>
BN_HD void write_state_packed(const StateIO& io, int64_t n_slots, int64_t s, const T m[Spec<K1, K2>::D],
                              const T P[Spec<K1, K2>::NS]) {
  using S = Spec<K1, K2>;
  constexpr int D = S::D;
  BN_UNROLL for (int i = 0; i < D; ++i) io.mean[i * n_slots + s] = (state_t)m[i];
  if (io.packed) {
    BN_UNROLL for (int q = 0; q < S::NS; ++q) io.cov[q * n_slots + s] = (state_t)P[q];
  } else {
    BN_UNROLL for (int i = 0; i < D; ++i)
      BN_UNROLL for (int j = 0; j < D; ++j)
        io.cov[(i * D + j) * n_slots + s] = (state_t)P[i <= j ? S::sidx(i, j) : S::sidx(j, i)];
  }
}

template <int K1, int K2, class T>
BN_HD void read_state(const CStateIO& io, int64_t n_slots, int64_t s, T m[Spec<K1, K2>::D],
                      T P[Spec<K1, K2>::D][Spec<K1, K2>::D]) {
  using S = Spec<K1, K2>;
  constexpr int D = S::D;
  BN_UNROLL for (int i = 0; i < D; ++i) m[i] = (T)io.mean[i * n_slots + s];
  if (io.packed) {
    BN_UNROLL for (int i = 0; i < D; ++i)
      BN_UNROLL for (int j = i; j < D; ++j) {
        T v = (T)io.cov[S::sidx(i, j) * n_slots + s];
        P[i][j] = v;
        P[j][i] = v;
      }
  } else {
    // symmetrise the caller's full covariance
    BN_UNROLL for (int i = 0; i < D; ++i)
      BN_UNROLL for (int j = i; j < D; ++j) {
        T v = T(0.5) * ((T)io.cov[(i * D + j) * n_slots + s] + (T)io.cov[(j * D + i) * n_slots + s]);
        P[i][j] = v;
        P[j][i] = v;
      }
  }
}

// ---------------------------------------------------------------------------
// filter_reduce: the chunk's filtering element, built by running the filter
// conditionally on the (unknown) state x before the chunk: m = Phi x + beta,
// P conditional covariance, and the information (J, eta) that y_chunk carries
// about x.  The first chunk of a series starts from the prior (m0 = 0,
// P0 = Pinf, ops.py:272-273) and yields a state element (Phi = 0).
// ---------------------------------------------------------------------------
template <int K1, int K2, class T>
BN_HD void lane_filter_reduce(const PlanView& pl, const HyperView& hv, const float* __restrict__ dt,
                              const float* __restrict__ y, const float* __restrict__ r, SoA agg, int32_t j) {
  // The element is kept with packed symmetric Sig / J (NS entries each) and updated
  // in place, which keeps the live FP64 state at D*D + 2*NS + 2*D values.
  using S = Spec<K1, K2>;
  constexpr int D = S::D;
  constexpr int NS = S::NS;
  const int32_t b = pl.chunk_series[j];
  const int32_t L = pl.chunk_len_[j];
  const uint8_t fl = pl.chunk_flags[j];
  const int64_t o = (int64_t)b * hv.stride;
  const double ls1 = hv.ls[o], v1 = hv.var[o];
  const double ls2 = K2 > 0 ? hv.ls[o + 1] : 1.0, v2 = K2 > 0 ? hv.var[o + 1] : 1.0;
  T Phi[D][D], beta[D], Sig[NS], J[NS], eta[D];
  BN_UNROLL for (int i = 0; i < D; ++i) {
    beta[i] = T(0);
    eta[i] = T(0);
    BN_UNROLL for (int c = 0; c < D; ++c) Phi[i][c] = T(i == c ? 1 : 0);
  }
  BN_UNROLL for (int q = 0; q < NS; ++q) { Sig[q] = T(0); J[q] = T(0); }
  int32_t k0 = 0;
  if (fl & CH_FIRST) {
    // the series' first step: predicted state (0, Pinf) and no dependence on the
    // incoming state (Phi- = 0)
    Model<K1, K2> mdl;
    mdl.init(ls1, v1, ls2, v2);
    BN_UNROLL for (int i = 0; i < D; ++i) {
      BN_UNROLL for (int c = i; c < D; ++c) Sig[S::sidx(i, c)] = (T)mdl.Pinf[i][c];
      BN_UNROLL for (int c = 0; c < D; ++c) Phi[i][c] = T(0);
    }
    k0 = 0;
  }
  Model<K1, K2> mdl;
  mdl.lam1 = comp_lambda<K1>(ls1);
  mdl.lam2 = K2 > 0 ? comp_lambda<(K2 > 0 ? K2 : 1)>(ls2) : 0.0;
  mdl.var1 = v1;
  mdl.var2 = v2;
  // the next step's inputs are loaded one step ahead, off the recurrence's critical path
  float nx_dt = 0.f, nx_y = 0.f, nx_r = 0.f;
  if (k0 < L) {
    const int64_t s0 = pl.slot(j, k0);
    nx_dt = dt[s0];
    nx_y = y[s0];
    nx_r = r[s0];
  }
  for (int32_t k = k0; k < L; ++k) {
    const float dtk = nx_dt, yk = nx_y, rk = nx_r;
    if (k + 1 < L) {
      const int64_t s1 = pl.slot(j, k + 1);
      nx_dt = dt[s1];
      nx_y = y[s1];
      nx_r = r[s1];
    }
    {
      // The series' first step has no predict: a zero step there (A = I and Q = 0
      // exactly) keeps the body free of a divergent branch, whose merge costs a register
      // copy of every updated value per step.
      const float dte = (k == 0 && (fl & CH_FIRST)) ? 0.f : dtk;
      T A[D][D], Q[D][D];
      mdl.discretise(dte, A, Q);
      // beta = A beta, Phi = A Phi (block-diagonal A; new rows of a block from old ones)
      T nb[D];
      BN_UNROLL for (int i = 0; i < D; ++i) {
        T t = T(0);
        BN_UNROLL for (int q = 0; q < D; ++q) if (S::nz(i, q)) t += A[i][q] * beta[q];
        nb[i] = t;
      }
      BN_UNROLL for (int i = 0; i < D; ++i) beta[i] = nb[i];
      BN_UNROLL for (int c = 0; c < D; ++c) {
        T col[D];
        BN_UNROLL for (int i = 0; i < D; ++i) {
          T t = T(0);
          BN_UNROLL for (int q = 0; q < D; ++q) if (S::nz(i, q)) t += A[i][q] * Phi[q][c];
          col[i] = t;
        }
        BN_UNROLL for (int i = 0; i < D; ++i) Phi[i][c] = col[i];
      }
      // Sig = A Sig A' + Q on the packed upper triangle
      T ns[NS];
      BN_UNROLL for (int i = 0; i < D; ++i)
        BN_UNROLL for (int c = i; c < D; ++c) {
          T t = S::nz(i, c) ? Q[i][c] : T(0);
          BN_UNROLL for (int p = 0; p < D; ++p) {
            if (!S::nz(i, p)) continue;
            T u = T(0);
            BN_UNROLL for (int q = 0; q < D; ++q)
              if (S::nz(c, q)) u += Sig[p <= q ? S::sidx(p, q) : S::sidx(q, p)] * A[c][q];
            t += A[i][p] * u;
          }
          ns[S::sidx(i, c)] = t;
        }
      BN_UNROLL for (int q = 0; q < NS; ++q) Sig[q] = ns[q];
    }
    // u = H Phi-, hP = H P-, S = H P- H' + r, v = y - H beta-
    T u[D], hP[D];
    T hb = T(0);
    BN_UNROLL for (int c = 0; c < D; ++c) {
      T su = T(0), sp = T(0);
      BN_UNROLL for (int i = 0; i < D; ++i)
        if (S::h(i)) {
          su += Phi[i][c];
          sp += Sig[i <= c ? S::sidx(i, c) : S::sidx(c, i)];
        }
      u[c] = su;
      hP[c] = sp;
    }
    double Sd = (double)rk;
    BN_UNROLL for (int i = 0; i < D; ++i) if (S::h(i)) { Sd += (double)hP[i]; hb += beta[i]; }
    const double iS = 1.0 / Sd;
    const T v = (T)((double)yk - (double)hb);
    const T iSf = (T)iS;
    const T vS = v * iSf;
    // rank-one updates with the 1 / S folded into one factor first (one FMA per entry)
    BN_UNROLL for (int i = 0; i < D; ++i) {
      eta[i] += u[i] * vS;
      const T ki = hP[i] * iSf, ui = u[i] * iSf;
      beta[i] += ki * v;
      BN_UNROLL for (int c = i; c < D; ++c) {
        J[S::sidx(i, c)] += ui * u[c];
        Sig[S::sidx(i, c)] -= ki * hP[c];
      }
      BN_UNROLL for (int c = 0; c < D; ++c) Phi[i][c] -= ki * u[c];
    }
  }
  // storage order of FElem: Phi, beta, Sig (upper), J (upper), eta
  int f = 0;
  BN_UNROLL for (int i = 0; i < D; ++i) BN_UNROLL for (int c = 0; c < D; ++c) agg.base[(int64_t)(f++) * agg.stride + j] = (agg_t)Phi[i][c];
  BN_UNROLL for (int i = 0; i < D; ++i) agg.base[(int64_t)(f++) * agg.stride + j] = (agg_t)beta[i];
  BN_UNROLL for (int q = 0; q < NS; ++q) agg.base[(int64_t)(f++) * agg.stride + j] = (agg_t)Sig[q];
  BN_UNROLL for (int q = 0; q < NS; ++q) agg.base[(int64_t)(f++) * agg.stride + j] = (agg_t)J[q];
  BN_UNROLL for (int i = 0; i < D; ++i) agg.base[(int64_t)(f++) * agg.stride + j] = (agg_t)eta[i];
}

// Incoming filtered state of chunk j (not the first of its series): the
// inclusive prefix of all earlier chunks.  `agg` holds tile-local inclusive
// prefixes, `carry[t]` the prefix of all tiles of the series before tile t.
template <int D, class T>
BN_HD void filter_incoming(const PlanView& pl, CSoA agg, CSoA carry, int32_t j, T m[D], T P[D][D]) {
  const int32_t t = pl.chunk_tile[j];
  FElem<D, T> e;
  if (pl.tile_first[t] == j) {
    soa_load(e, carry, t);
    BN_UNROLL for (int i = 0; i < D; ++i) {
      m[i] = e.beta[i];
      BN_UNROLL for (int c = 0; c < D; ++c) P[i][c] = e.Sig[i][c];
    }
    return;
  }
  soa_load(e, agg, j - 1);
  if (pl.tile_flags[t] & TL_FIRST) {
    BN_UNROLL for (int i = 0; i < D; ++i) {
      m[i] = e.beta[i];
      BN_UNROLL for (int c = 0; c < D; ++c) P[i][c] = e.Sig[i][c];
    }
    return;
  }
  FElem<D, T> cst;
  soa_load(cst, carry, t);
  BN_UNROLL for (int i = 0; i < D; ++i) {
    m[i] = cst.beta[i];
    BN_UNROLL for (int c = 0; c < D; ++c) P[i][c] = cst.Sig[i][c];
  }
  fapply<D, T>(m, P, e);
}

// Accumulate the smoothing element of one step into the chunk's running
// composition acc = e_s o ... o e_{k-1} (forward in time, see SElem):
//   G = fP A' pP^-1, g = fm - G pm, L = fP - G A fP   (ops.py:318-325)
template <int D, class T>
BN_HD void smooth_accumulate(SElem<D, T>& acc, const T fm[D], const T fP[D][D], const T AfP[D][D],
                             const T pm[D], const T pP[D][D]) {
  T G[D][D];
  rts_gain<D, T>(AfP, pP, G);
  T g[D], Lk[D][D];
  BN_UNROLL for (int i = 0; i < D; ++i) {
    T s = fm[i];
    BN_UNROLL for (int k = 0; k < D; ++k) s -= G[i][k] * pm[k];
    g[i] = s;
  }
  BN_UNROLL for (int i = 0; i < D; ++i)
    BN_UNROLL for (int c = i; c < D; ++c) {
      T s1 = T(0), s2 = T(0);
      BN_UNROLL for (int k = 0; k < D; ++k) { s1 += G[i][k] * AfP[k][c]; s2 += G[c][k] * AfP[k][i]; }
      T s = fP[i][c] - T(0.5) * (s1 + s2);
      Lk[i][c] = s;
      Lk[c][i] = s;
    }
  // acc.g += acc.E g ; acc.L += acc.E Lk acc.E' ; acc.E = acc.E G
  T EL[D][D];
  mat_mul<D, T>(acc.E, Lk, EL);
  BN_UNROLL for (int i = 0; i < D; ++i) {
    T s = acc.g[i];
    BN_UNROLL for (int k = 0; k < D; ++k) s += acc.E[i][k] * g[k];
    acc.g[i] = s;
    BN_UNROLL for (int c = i; c < D; ++c) {
      T t = acc.L[i][c];
      BN_UNROLL for (int k = 0; k < D; ++k) t += EL[i][k] * acc.E[c][k];
      acc.L[i][c] = t;
      acc.L[c][i] = t;
    }
  }
  T EG[D][D];
  mat_mul<D, T>(acc.E, G, EG);
  BN_UNROLL for (int i = 0; i < D; ++i) BN_UNROLL for (int c = 0; c < D; ++c) acc.E[i][c] = EG[i][c];
}

// Terminal element of a series: the smoothed state at the last step is the
// filtered one (ops.py:314-315): acc = acc o (0, fm, fP).
template <int D, class T>
BN_HD void smooth_terminal(SElem<D, T>& acc, const T fm[D], const T fP[D][D]) {
  T EL[D][D];
  mat_mul<D, T>(acc.E, fP, EL);
  BN_UNROLL for (int i = 0; i < D; ++i) {
    T s = acc.g[i];
    BN_UNROLL for (int k = 0; k < D; ++k) s += acc.E[i][k] * fm[k];
    acc.g[i] = s;
    BN_UNROLL for (int c = i; c < D; ++c) {
      T t = acc.L[i][c];
      BN_UNROLL for (int k = 0; k < D; ++k) t += EL[i][k] * acc.E[c][k];
      acc.L[i][c] = t;
      acc.L[c][i] = t;
    }
  }
  set_zero<D>(acc.E);
}

// Prediction without Q: A P A' + Q = A (P - Pinf) A' + Pinf exactly, since
// Q = Pinf - A Pinf A' (ops.py:149-151).  Both forms carry rounding of order
// eps |Pinf|, as the reference's own Q does; the filter and smoother passes
// start from a filtered state, so only their absolute accuracy matters.  Also
// returns A P (needed for the RTS gain).
template <int K1, int K2, class T>
BN_HD void predict_ap(const T A[Spec<K1, K2>::D][Spec<K1, K2>::D], const double Pinf[Spec<K1, K2>::D][Spec<K1, K2>::D],
                      const T m[Spec<K1, K2>::D], const T P[Spec<K1, K2>::D][Spec<K1, K2>::D],
                      T mp[Spec<K1, K2>::D], T Pp[Spec<K1, K2>::D][Spec<K1, K2>::D],
                      T AP[Spec<K1, K2>::D][Spec<K1, K2>::D]) {
  using S = Spec<K1, K2>;
  constexpr int D = S::D;
  T AD[D][D];
  BN_UNROLL for (int i = 0; i < D; ++i) {
    T s = T(0);
    BN_UNROLL for (int k = 0; k < D; ++k) if (S::nz(i, k)) s += A[i][k] * m[k];
    mp[i] = s;
  }
  BN_UNROLL for (int i = 0; i < D; ++i)
    BN_UNROLL for (int c = 0; c < D; ++c) {
      T s = T(0), sp = T(0);
      BN_UNROLL for (int k = 0; k < D; ++k) if (S::nz(i, k)) {
        s += A[i][k] * P[k][c];
        if (S::nz(k, c)) sp += A[i][k] * (T)Pinf[k][c];
      }
      AP[i][c] = s;
      AD[i][c] = s - sp;
    }
  BN_UNROLL for (int i = 0; i < D; ++i)
    BN_UNROLL for (int c = i; c < D; ++c) {
      T s = S::nz(i, c) ? (T)Pinf[i][c] : T(0);
      BN_UNROLL for (int k = 0; k < D; ++k) if (S::nz(c, k)) s += AD[i][k] * A[c][k];
      Pp[i][c] = s;
      Pp[c][i] = s;
    }
}

// A only (the passes never need Q)
template <int K1, int K2, class T>
BN_HD void transition_only(const Model<K1, K2>& mdl, double dt, T A[Spec<K1, K2>::D][Spec<K1, K2>::D]) {
  constexpr int D = Spec<K1, K2>::D;
  double a1[9], a2[9];
  comp_transition<K1>(mdl.lam1, dt, a1);
  if (K2 > 0) comp_transition<(K2 > 0 ? K2 : 1)>(mdl.lam2, dt, a2);
  BN_UNROLL for (int i = 0; i < D; ++i)
    BN_UNROLL for (int j = 0; j < D; ++j) {
      double v = 0.0;
      if (i < K1 && j < K1) v = a1[i * K1 + j];
      else if (i >= K1 && j >= K1) v = a2[(i - K1) * (K2 > 0 ? K2 : 1) + (j - K1)];
      A[i][j] = (T)v;
    }
}

struct FilterPassOut {
  StateIO state;       // filtered (or predicted) moments; state.mean == nullptr: not written
  bool predict;        // write predicted instead of filtered moments
  SoA sagg;            // smoothing element per chunk; sagg.base == nullptr: not built
  double* ell;         // per-chunk log-likelihood partial
  // (D + 1) planes: H P- (D) and H m- at every step, all the Bryson-Frazier smoother
  // reads per step; when set, `state` is written only at each chunk's first step (the
  // smoother's anchor).  nullptr: `state` at every step.
  hrow_t* hrow = nullptr;
  // per-chunk status bits (ST_*), or nullptr: ST_INNOV_NONPD when some step's
  // innovation variance H P- H' + r is not > 0 (or NaN), ST_NONFINITE when the
  // chunk's log-likelihood or final filtered mean is NaN / infinite
  int32_t* status = nullptr;
};

// filter_pass: the filter over one chunk from its true incoming state
// (ops.py:156-175), writing moments, the log-likelihood partial and the
// chunk's smoothing element.  The smoothing element x_s | x_{e+1}, y is built
// by a fixed-point smoother run alongside the filter: with the chunk's first
// step s, it tracks mu = E[x_s | y_1..k], V = Var(x_s | y_1..k) and
// C = Cov(x_s, x_k | y_1..k) (rank-1 updates since p = 1); conditioning the
// joint of (x_s, x_{e+1}) given y_1..e then gives E = C- Pp^-1, g = mu - E mp,
// L = V - E C-', the element of ops.py:318-335 composed over the chunk, with
// one solve per chunk.  The smoother uses dt of the NEXT step (basemodels.py:700).
// WANT_S: build the smoother element (out.sagg is set); without it the fixed-point
// smoother is not run.
template <int K1, int K2, class T, bool WANT_S>
BN_HD void lane_filter_pass(const PlanView& pl, const HyperView& hv, const float* __restrict__ dt,
                            const float* __restrict__ y, const float* __restrict__ r, const uint8_t* __restrict__ mask,
                            CSoA fagg, CSoA fcarry, const FilterPassOut& out, int32_t j) {
  // Covariances P (filter) and V (fixed-point smoother) are kept packed (NS
  // entries) and updated in place; C = Cov(x_s, x_k | y_1..k) is a full D x D.
  using S = Spec<K1, K2>;
  constexpr int D = S::D;
  constexpr int NS = S::NS;
  const int32_t b = pl.chunk_series[j];
  const int32_t L = pl.chunk_len_[j];
  const uint8_t fl = pl.chunk_flags[j];
  Model<K1, K2> mdl = load_model<K1, K2>(hv, b);
  T m[D], P[NS];
  if (!(fl & CH_FIRST)) {
    T Pf[D][D];
    filter_incoming<D, T>(pl, fagg, fcarry, j, m, Pf);
    BN_UNROLL for (int i = 0; i < D; ++i) BN_UNROLL for (int c = i; c < D; ++c) P[S::sidx(i, c)] = Pf[i][c];
  } else {
    // the series' first step: predicted state (0, Pinf), reached below by a zero step
    BN_UNROLL for (int i = 0; i < D; ++i) {
      m[i] = T(0);
      BN_UNROLL for (int c = i; c < D; ++c) P[S::sidx(i, c)] = (T)mdl.Pinf[i][c];
    }
  }
  constexpr bool want_s = WANT_S;
  T mu[D], V[NS], C[D][D];  // fixed-point smoother of x_s
  // log S is summed as the log of a running product kept as mantissa * 2^exponent
  // (two frexp per step in place of a log; the mantissa is renormalised every step)
  double ell = 0.0, s_mant = 1.0;
  int32_t s_exp = 0;
  int32_t st = 0;
  auto pk = [](int i, int c) { return i <= c ? S::sidx(i, c) : S::sidx(c, i); };
  // the next step's inputs are loaded one step ahead, off the recurrence's critical path
  float nx_dt = 0.f, nx_y = 0.f, nx_r = 0.f;
  bool nx_masked = false;
  if (L > 0) {
    const int64_t s0 = pl.slot(j, 0);
    nx_dt = dt[s0];
    nx_y = y[s0];
    nx_r = r[s0];
    nx_masked = mask != nullptr && mask[s0];
  }
  // One step of the chunk.  The chunk's first step is a separate instance (FIRST), so
  // that no step carries a branch on the step index: the merge after such a branch
  // costs a register copy of every value it updates.
  auto step = [&](int32_t k, auto first) {
    constexpr bool FIRST = decltype(first)::value;
    const int64_t s = pl.slot(j, k);
    const float dtk = nx_dt, yk = nx_y, rk = nx_r;
    const bool masked = nx_masked;
    if (k + 1 < L) {
      const int64_t s1 = pl.slot(j, k + 1);
      nx_dt = dt[s1];
      nx_y = y[s1];
      nx_r = r[s1];
      nx_masked = mask != nullptr && mask[s1];
    }
    {
      // predict in place: m = A m, P = A (P - Pinf) A' + Pinf (Q-free form), C = C A'.
      // The series' first step has no predict: a zero step there (A = I exactly) keeps
      // the body free of a divergent branch, whose merge costs a register copy of every
      // updated value per step.
      const float dte = (FIRST && (fl & CH_FIRST)) ? 0.f : dtk;
      T A[D][D];
      transition_only<K1, K2, T>(mdl, dte, A);
      T nm[D];
      BN_UNROLL for (int i = 0; i < D; ++i) {
        T t = T(0);
        BN_UNROLL for (int q = 0; q < D; ++q) if (S::nz(i, q)) t += A[i][q] * m[q];
        nm[i] = t;
      }
      BN_UNROLL for (int i = 0; i < D; ++i) m[i] = nm[i];
      T np_[NS];
      BN_UNROLL for (int i = 0; i < D; ++i)
        BN_UNROLL for (int c = i; c < D; ++c) {
          T t = S::nz(i, c) ? (T)mdl.Pinf[i][c] : T(0);
          BN_UNROLL for (int p = 0; p < D; ++p) {
            if (!S::nz(i, p)) continue;
            T u = T(0);
            BN_UNROLL for (int q = 0; q < D; ++q)
              if (S::nz(c, q)) u += (P[pk(p, q)] - (S::nz(p, q) ? (T)mdl.Pinf[p][q] : T(0))) * A[c][q];
            t += A[i][p] * u;
          }
          np_[S::sidx(i, c)] = t;
        }
      BN_UNROLL for (int q = 0; q < NS; ++q) P[q] = np_[q];
      if constexpr (want_s && !FIRST) {
        BN_UNROLL for (int i = 0; i < D; ++i) {
          T row[D];
          BN_UNROLL for (int c = 0; c < D; ++c) {
            T t = T(0);
            BN_UNROLL for (int q = 0; q < D; ++q) if (S::nz(c, q)) t += C[i][q] * A[c][q];
            row[c] = t;
          }
          BN_UNROLL for (int c = 0; c < D; ++c) C[i][c] = row[c];
        }
      }
    }
    if (out.state.mean != nullptr && out.predict && (out.hrow == nullptr || FIRST))
      write_state_packed<K1, K2, T>(out.state, pl.n_slots, s, m, P);
    // measurement update (ops.py:163-171): hP = H P-, S = H P- H' + r, v = y - H m-
    T hP[D];
    BN_UNROLL for (int c = 0; c < D; ++c) {
      T t = T(0);
      BN_UNROLL for (int i = 0; i < D; ++i) if (S::h(i)) t += P[pk(i, c)];
      hP[c] = t;
    }
    double Sd = (double)rk;
    T hm = T(0);
    BN_UNROLL for (int q = 0; q < D; ++q) if (S::h(q)) { Sd += (double)hP[q]; hm += m[q]; }
    if (!(Sd > 0.0)) st |= ST_INNOV_NONPD;
    if (out.hrow != nullptr) {
      BN_UNROLL for (int q = 0; q < D; ++q) out.hrow[q * pl.n_slots + s] = (hrow_t)hP[q];
      out.hrow[D * pl.n_slots + s] = (hrow_t)hm;
    }
    const double vd = (double)yk - (double)hm;
    const double iSd = 1.0 / Sd;
    const T vS = (T)(vd * iSd);
    const T iS = (T)iSd;
    if (!masked) {
      ell += -0.5 * (vd * vd * iSd + kLog2Pi);
      int e1, e2;
      s_mant = frexp(s_mant * frexp(Sd, &e1), &e2);
      s_exp += e1 + e2;
    }
    if constexpr (want_s && !FIRST) {
      // fixed-point smoother: ch = C- H'; mu += ch v / S; V -= ch ch' / S; C = C- - ch hP' / S
      T ch[D];
      BN_UNROLL for (int i = 0; i < D; ++i) {
        T t = T(0);
        BN_UNROLL for (int q = 0; q < D; ++q) if (S::h(q)) t += C[i][q];
        ch[i] = t;
      }
      BN_UNROLL for (int i = 0; i < D; ++i) {
        mu[i] += ch[i] * vS;
        const T ci = ch[i] * iS;
        BN_UNROLL for (int c = i; c < D; ++c) V[S::sidx(i, c)] -= ci * ch[c];
        BN_UNROLL for (int c = 0; c < D; ++c) C[i][c] -= ci * hP[c];
      }
    }
    BN_UNROLL for (int i = 0; i < D; ++i) {
      m[i] += hP[i] * vS;
      const T ki = hP[i] * iS;
      BN_UNROLL for (int c = i; c < D; ++c) P[S::sidx(i, c)] -= ki * hP[c];
    }
    if constexpr (want_s && FIRST) {
      BN_UNROLL for (int i = 0; i < D; ++i) {
        mu[i] = m[i];
        BN_UNROLL for (int c = i; c < D; ++c) V[S::sidx(i, c)] = P[S::sidx(i, c)];
        BN_UNROLL for (int c = 0; c < D; ++c) C[i][c] = P[pk(i, c)];
      }
    }
    if (out.state.mean != nullptr && !out.predict) write_state_packed<K1, K2, T>(out.state, pl.n_slots, s, m, P);  };
  if (L > 0) step(0, std::true_type{});
  for (int32_t k = 1; k < L; ++k) step(k, std::false_type{});

  if (want_s) {
    SElem<D, T> acc;
    if (fl & CH_LAST) {
      // terminal: x_s | y_1..N is a state (ops.py:314-315)
      set_zero<D, T>(acc.E);
      BN_UNROLL for (int i = 0; i < D; ++i) {
        acc.g[i] = mu[i];
        BN_UNROLL for (int c = 0; c < D; ++c) acc.L[i][c] = V[pk(i, c)];
      }
    } else {
      T A[D][D], mp[D], Pp[D][D], AP[D][D], Cn[D][D], Pf[D][D];
      BN_UNROLL for (int i = 0; i < D; ++i) BN_UNROLL for (int c = 0; c < D; ++c) Pf[i][c] = P[pk(i, c)];
      transition_only<K1, K2, T>(mdl, dt[pl.slot(j + 1, 0)], A);
      predict_ap<K1, K2, T>(A, mdl.Pinf, m, Pf, mp, Pp, AP);
      BN_UNROLL for (int i = 0; i < D; ++i)
        BN_UNROLL for (int c = 0; c < D; ++c) {
          T t = T(0);
          BN_UNROLL for (int q = 0; q < D; ++q) if (S::nz(c, q)) t += C[i][q] * A[c][q];
          Cn[i][c] = t;
        }
      // E = Cn Pp^-1  (E' = Pp^-1 Cn', Cholesky of Pp in FP64)
      double Lc[D][D], X[D][D];
      BN_UNROLL for (int i = 0; i < D; ++i)
        BN_UNROLL for (int c = 0; c < D; ++c) { Lc[i][c] = (double)Pp[i][c]; X[i][c] = (double)Cn[c][i]; }
      chol_spd<D>(Lc);
      lower_solve<D, D, double>(Lc, X);
      upper_t_solve<D, D, double>(Lc, X);
      BN_UNROLL for (int i = 0; i < D; ++i) BN_UNROLL for (int c = 0; c < D; ++c) acc.E[i][c] = (T)X[c][i];
      BN_UNROLL for (int i = 0; i < D; ++i) {
        T t = mu[i];
        BN_UNROLL for (int q = 0; q < D; ++q) t -= acc.E[i][q] * mp[q];
        acc.g[i] = t;
      }
      BN_UNROLL for (int i = 0; i < D; ++i)
        BN_UNROLL for (int c = i; c < D; ++c) {
          T t1 = T(0), t2 = T(0);
          BN_UNROLL for (int q = 0; q < D; ++q) { t1 += acc.E[i][q] * Cn[c][q]; t2 += acc.E[c][q] * Cn[i][q]; }
          const T t = V[S::sidx(i, c)] - T(0.5) * (t1 + t2);
          acc.L[i][c] = t;
          acc.L[c][i] = t;
        }
    }
    soa_store(acc, out.sagg, j);
  }
  ell -= 0.5 * (log(s_mant) + (double)s_exp * kLn2);
  out.ell[j] = ell;
  if (out.status != nullptr) {
    bool fin = finite_d(ell);
    BN_UNROLL for (int i = 0; i < D; ++i) fin = fin && finite_d((double)m[i]);
    if (!fin) st |= ST_NONFINITE;
    out.status[j] = st;
  }
}

// smoother_reduce (standalone smoother API): the chunk's smoothing element
// from given filtered moments and an explicit (already shifted) dt.
template <int K1, int K2, class T>
BN_HD void lane_smoother_reduce(const PlanView& pl, const HyperView& hv, const float* __restrict__ dts,
                                const CStateIO& fin, SoA sagg, int32_t j) {
  using S = Spec<K1, K2>;
  constexpr int D = S::D;
  const int32_t b = pl.chunk_series[j];
  const int32_t L = pl.chunk_len_[j];
  const uint8_t fl = pl.chunk_flags[j];
  Model<K1, K2> mdl = load_model<K1, K2>(hv, b);
  SElem<D, T> acc;
  acc.set_identity_elem();
  for (int32_t k = 0; k < L; ++k) {
    const int64_t s = pl.slot(j, k);
    T fm[D], fP[D][D];
    read_state<K1, K2, T>(fin, pl.n_slots, s, fm, fP);
    if (k == L - 1 && (fl & CH_LAST)) {
      smooth_terminal<D, T>(acc, fm, fP);
    } else {
      T A[D][D], mp[D], Pp[D][D], AP[D][D];
      transition_only<K1, K2, T>(mdl, dts[s], A);
      predict_ap<K1, K2, T>(A, mdl.Pinf, fm, fP, mp, Pp, AP);
      smooth_accumulate<D, T>(acc, fm, fP, AP, mp, Pp);
    }
  }
  soa_store(acc, sagg, j);
}

// Incoming smoothed state of chunk j (not the last of its series): the
// inclusive suffix of all later chunks.  `agg` holds tile-local inclusive
// suffixes, `carry[t]` the suffix of all tiles of the series after tile t.
template <int D, class T>
BN_HD void smoother_incoming(const PlanView& pl, CSoA agg, CSoA carry, int32_t j, T m[D], T P[D][D]) {
  const int32_t t = pl.chunk_tile[j];
  SElem<D, T> e;
  if (pl.tile_first[t] + pl.tile_count[t] - 1 == j) {
    soa_load(e, carry, t);
    BN_UNROLL for (int i = 0; i < D; ++i) {
      m[i] = e.g[i];
      BN_UNROLL for (int c = 0; c < D; ++c) P[i][c] = e.L[i][c];
    }
    return;
  }
  soa_load(e, agg, j + 1);
  if (pl.tile_flags[t] & TL_LAST) {
    BN_UNROLL for (int i = 0; i < D; ++i) {
      m[i] = e.g[i];
      BN_UNROLL for (int c = 0; c < D; ++c) P[i][c] = e.L[i][c];
    }
    return;
  }
  SElem<D, T> cst;
  soa_load(cst, carry, t);
  BN_UNROLL for (int i = 0; i < D; ++i) {
    m[i] = cst.g[i];
    BN_UNROLL for (int c = 0; c < D; ++c) P[i][c] = cst.L[i][c];
  }
  sapply<D, T>(e, m, P);
}

template <class OT>
struct SmootherPassIO {
  CStateIO fin;        // filtered moments
  const float* dt;     // step vector
  bool shifted;        // dt is the smoother's own (shifted) vector; else use dt[k+1], 0 at the end
  OT* out_mean;        // H m planes (1) or full (D) when full
  OT* out_cov;         // H P H' (1) or full (D*D)
  OT* gains;           // D*D planes or nullptr
  bool full;
};

// smoother_pass: the RTS recursion over one chunk, backwards, from its true
// incoming smoothed state (ops.py:290-305).
template <int K1, int K2, class T, class OT>
BN_HD void lane_smoother_pass(const PlanView& pl, const HyperView& hv, CSoA sagg, CSoA scarry,
                              const SmootherPassIO<OT>& io, int32_t j) {
  using S = Spec<K1, K2>;
  constexpr int D = S::D;
  const int32_t b = pl.chunk_series[j];
  const int32_t L = pl.chunk_len_[j];
  const uint8_t fl = pl.chunk_flags[j];
  Model<K1, K2> mdl = load_model<K1, K2>(hv, b);
  T sm[D], sP[D][D];
  if (!(fl & CH_LAST)) smoother_incoming<D, T>(pl, sagg, scarry, j, sm, sP);
  const int64_t ns = pl.n_slots;
  // One step, walking backwards.  The series' terminal step (the last step of its last
  // chunk) is a separate instance (LAST), so that no other step carries that branch,
  // whose merge would cost a register copy of the smoothed state per step.
  auto step = [&](int32_t k, auto last) {
    constexpr bool LAST = decltype(last)::value;
    const int64_t s = pl.slot(j, k);
    T fm[D], fP[D][D], G[D][D];
    read_state<K1, K2, T>(io.fin, ns, s, fm, fP);
    const bool terminal = LAST && (fl & CH_LAST);
    double dtn;
    if (io.shifted) dtn = io.dt[s];
    else dtn = terminal ? 0.0 : (double)io.dt[k + 1 < L ? pl.slot(j, k + 1) : pl.slot(j + 1, 0)];
    T A[D][D], pm[D], pP[D][D], AfP[D][D];
    transition_only<K1, K2, T>(mdl, dtn, A);
    predict_ap<K1, K2, T>(A, mdl.Pinf, fm, fP, pm, pP, AfP);
    rts_gain<D, T>(AfP, pP, G);
    if (terminal) {
      BN_UNROLL for (int i = 0; i < D; ++i) {
        sm[i] = fm[i];
        BN_UNROLL for (int c = 0; c < D; ++c) sP[i][c] = fP[i][c];
      }
    } else {
      T dm[D], nm[D], Dd[D][D], GD[D][D];
      BN_UNROLL for (int i = 0; i < D; ++i) dm[i] = sm[i] - pm[i];
      BN_UNROLL for (int i = 0; i < D; ++i) {
        T t = fm[i];
        BN_UNROLL for (int c = 0; c < D; ++c) t += G[i][c] * dm[c];
        nm[i] = t;
      }
      BN_UNROLL for (int i = 0; i < D; ++i) BN_UNROLL for (int c = 0; c < D; ++c) Dd[i][c] = sP[i][c] - pP[i][c];
      mat_mul<D, T>(G, Dd, GD);
      BN_UNROLL for (int i = 0; i < D; ++i)
        BN_UNROLL for (int c = i; c < D; ++c) {
          T t = fP[i][c];
          BN_UNROLL for (int q = 0; q < D; ++q) t += GD[i][q] * G[c][q];
          sP[i][c] = t;
          sP[c][i] = t;
        }
      BN_UNROLL for (int i = 0; i < D; ++i) sm[i] = nm[i];
    }
    if (io.full) {
      BN_UNROLL for (int i = 0; i < D; ++i) io.out_mean[i * ns + s] = (OT)sm[i];
      BN_UNROLL for (int i = 0; i < D; ++i)
        BN_UNROLL for (int c = 0; c < D; ++c) io.out_cov[(i * D + c) * ns + s] = (OT)sP[i][c];
    } else {
      T hm = T(0), hv2 = T(0);
      BN_UNROLL for (int i = 0; i < D; ++i) if (S::h(i)) {
        hm += sm[i];
        BN_UNROLL for (int c = 0; c < D; ++c) if (S::h(c)) hv2 += sP[i][c];
      }
      io.out_mean[s] = (OT)hm;
      io.out_cov[s] = (OT)hv2;
    }
    if (io.gains != nullptr) {
      BN_UNROLL for (int i = 0; i < D; ++i)
        BN_UNROLL for (int c = 0; c < D; ++c) io.gains[(i * D + c) * ns + s] = (OT)G[i][c];
    }
  };
  int32_t k0 = L - 1;  // the last step the walk below takes
  if (L > 0 && (fl & CH_LAST)) {
    step(L - 1, std::true_type{});
    k0 = L - 2;
  }
  // the remaining steps in step with the wavefront's longest chunk, as in
  // lane_smoother_mbf (every lane reads step k of its chunk at the same time)
  const int32_t kmax = pl.wave_maxlen[j >> 6];
  for (int32_t k = kmax - 1; k >= 0; --k)
    if (k <= k0) step(k, std::false_type{});

}

// smoother_mbf: the smoothed marginals H m, H P H' of one chunk by the modified
// Bryson-Frazier recursion (backwards over the chunk) from the filter's PREDICTED
// moments (m-, P-) and the observations.  It carries the adjoint (lh, Lh) with
//   m_s = m- - P- lt,  P_s = P- - P- Lt P-,
//   lt = lh - h (k' lh + v / S),  Lt = Lh - h g' - g h' + (k' g + 1/S) h h',
//   g = Lh k, k = P- h / S,  lh <- A' lt, Lh <- A' Lt A (A of the step),
// which gives the posterior of ops.py:288-311 without factorising P- per step.
// The chunk's incoming adjoint comes from the smoothed state (sm, sP) at the next
// chunk's first step: lt = -P-^-1 (sm - m-), Lt = P-^-1 (P- - sP) P-^-1.
template <int K1, int K2, class T>
BN_HD void lane_smoother_mbf(const PlanView& pl, const HyperView& hv, const float* __restrict__ dt,
                             const float* __restrict__ y, const float* __restrict__ r, CSoA sagg, CSoA scarry,
                             const CStateIO& fin, const hrow_t* __restrict__ hrow, float* __restrict__ out_mean,
                             float* __restrict__ out_var, int32_t* __restrict__ status, int32_t j) {
  using S = Spec<K1, K2>;
  constexpr int D = S::D;
  constexpr int NS = S::NS;
  const int32_t b = pl.chunk_series[j];
  const int32_t L = pl.chunk_len_[j];
  const uint8_t fl = pl.chunk_flags[j];
  const int64_t ns = pl.n_slots;
  const int64_t o = (int64_t)b * hv.stride;
  Model<K1, K2> mdl;
  mdl.lam1 = comp_lambda<K1>(hv.ls[o]);
  mdl.lam2 = K2 > 0 ? comp_lambda<(K2 > 0 ? K2 : 1)>(hv.ls[o + 1]) : 0.0;
  auto pk = [](int i, int c) { return i <= c ? S::sidx(i, c) : S::sidx(c, i); };
  T lh[D], Lh[NS];
  BN_UNROLL for (int i = 0; i < D; ++i) lh[i] = T(0);
  BN_UNROLL for (int q = 0; q < NS; ++q) Lh[q] = T(0);
  // propagate (lt, Lt) of a step back over its transition A: lh = A' lt, Lh = A' Lt A
  auto back = [&](double dts, const T lt[D], const T Lt[NS]) {
    T A[D][D];
    transition_only<K1, K2, T>(mdl, dts, A);
    BN_UNROLL for (int i = 0; i < D; ++i) {
      T t = T(0);
      BN_UNROLL for (int q = 0; q < D; ++q) if (S::nz(q, i)) t += A[q][i] * lt[q];
      lh[i] = t;
    }
    BN_UNROLL for (int i = 0; i < D; ++i)
      BN_UNROLL for (int c = i; c < D; ++c) {
        T t = T(0);
        BN_UNROLL for (int p = 0; p < D; ++p) {
          if (!S::nz(p, i)) continue;
          T u = T(0);
          BN_UNROLL for (int q = 0; q < D; ++q) if (S::nz(q, c)) u += Lt[pk(p, q)] * A[q][c];
          t += A[p][i] * u;
        }
        Lh[S::sidx(i, c)] = t;
      }
  };
  if (!(fl & CH_LAST)) {
    T sm[D], sP[D][D], pm[D], pP[D][D];
    smoother_incoming<D, T>(pl, sagg, scarry, j, sm, sP);
    const int64_t sn = pl.slot(j + 1, 0);
    read_state<K1, K2, T>(fin, ns, sn, pm, pP);
    double Lc[D][D], xv[D][1], X[D][D];
    BN_UNROLL for (int i = 0; i < D; ++i) {
      xv[i][0] = (double)pm[i] - (double)sm[i];
      BN_UNROLL for (int c = 0; c < D; ++c) {
        Lc[i][c] = (double)pP[i][c];
        X[i][c] = (double)pP[i][c] - (double)sP[i][c];
      }
    }
    chol_spd<D>(Lc);
    lower_solve<D, 1, double>(Lc, xv);
    upper_t_solve<D, 1, double>(Lc, xv);
    lower_solve<D, D, double>(Lc, X);
    upper_t_solve<D, D, double>(Lc, X);   // X = P-^-1 (P- - sP)
    double Xt[D][D];
    BN_UNROLL for (int i = 0; i < D; ++i) BN_UNROLL for (int c = 0; c < D; ++c) Xt[i][c] = X[c][i];
    lower_solve<D, D, double>(Lc, Xt);
    upper_t_solve<D, D, double>(Lc, Xt);  // Xt = P-^-1 (P- - sP) P-^-1 (transposed; symmetric)
    T lt[D], Lt[NS];
    BN_UNROLL for (int i = 0; i < D; ++i) {
      lt[i] = (T)xv[i][0];
      BN_UNROLL for (int c = i; c < D; ++c) Lt[S::sidx(i, c)] = (T)(0.5 * (Xt[i][c] + Xt[c][i]));
    }
    back((double)dt[sn], lt, Lt);
  }
  bool smooth_ok = true;
  // the next (earlier) step's inputs and filter rows, loaded one step ahead
  float nx_dt = 0.f, nx_y = 0.f, nx_r = 0.f;
  hrow_t nx_h[D + 1];
  BN_UNROLL for (int c = 0; c <= D; ++c) nx_h[c] = hrow_t(0);
  if (L > 0) {
    const int64_t s0 = pl.slot(j, L - 1);
    nx_dt = dt[s0];
    nx_y = y[s0];
    nx_r = r[s0];
    BN_UNROLL for (int c = 0; c <= D; ++c) nx_h[c] = hrow[c * ns + s0];
  }
  // The steps run in step with the wavefront's longest chunk (wave_maxlen): a shorter
  // chunk's lane waits until the others reach its last step, so that every lane reads
  // step k of its chunk at the same time.  Staggered, the reads of each short chunk's lane
  // fetched cache lines of their own (a third more read traffic for this batch).
  const int32_t kmax = pl.wave_maxlen[j >> 6];
  for (int32_t k = kmax - 1; k >= 0; --k) {
    if (k >= L) continue;
    const int64_t s = pl.slot(j, k);
    const float dtk = nx_dt, yk = nx_y, rk = nx_r;
    T hP[D], kv[D];
    BN_UNROLL for (int c = 0; c < D; ++c) hP[c] = (T)nx_h[c];
    const T hm = (T)nx_h[D];
    if (k > 0) {
      const int64_t s1 = pl.slot(j, k - 1);
      nx_dt = dt[s1];
      nx_y = y[s1];
      nx_r = r[s1];
      BN_UNROLL for (int c = 0; c <= D; ++c) nx_h[c] = hrow[c * ns + s1];
    }
    T hph = T(0);
    BN_UNROLL for (int i = 0; i < D; ++i) if (S::h(i)) hph += hP[i];
    const double Sd = (double)rk + (double)hph;
    const double iSd = 1.0 / Sd;
    const T iS = (T)iSd;
    const T vS = (T)(((double)yk - (double)hm) * iSd);
    BN_UNROLL for (int i = 0; i < D; ++i) kv[i] = hP[i] * iS;
    T kl = T(0);
    BN_UNROLL for (int i = 0; i < D; ++i) kl += kv[i] * lh[i];
    T lt[D], g[D], Lt[NS];
    BN_UNROLL for (int i = 0; i < D; ++i) lt[i] = S::h(i) ? lh[i] - (kl + vS) : lh[i];
    BN_UNROLL for (int i = 0; i < D; ++i) {
      T t = T(0);
      BN_UNROLL for (int q = 0; q < D; ++q) t += Lh[pk(i, q)] * kv[q];
      g[i] = t;
    }
    T al = iS;
    BN_UNROLL for (int i = 0; i < D; ++i) al += kv[i] * g[i];
    BN_UNROLL for (int i = 0; i < D; ++i)
      BN_UNROLL for (int c = i; c < D; ++c) {
        T t = Lh[S::sidx(i, c)];
        if (S::h(i)) t -= g[c];
        if (S::h(c)) t -= g[i];
        if (S::h(i) && S::h(c)) t += al;
        Lt[S::sidx(i, c)] = t;
      }
    // H m_s = H m- - hP' lt ;  H P_s H' = H P- H' - hP' Lt hP
    T om = hm, ov = hph;
    BN_UNROLL for (int i = 0; i < D; ++i) {
      om -= hP[i] * lt[i];
      T t = T(0);
      BN_UNROLL for (int c = 0; c < D; ++c) t += Lt[pk(i, c)] * hP[c];
      ov -= hP[i] * t;
    }
    out_mean[s] = (float)om;
    out_var[s] = (float)ov;
    smooth_ok = smooth_ok && (double)ov > 0.0 && finite_d((double)om);
    if (k > 0) back((double)dtk, lt, Lt);
  }
  // the filter pass wrote this chunk's entry before this kernel ran
  if (status != nullptr && !smooth_ok) status[j] |= ST_SMOOTHER_NONPD;
}

}  // namespace bn
