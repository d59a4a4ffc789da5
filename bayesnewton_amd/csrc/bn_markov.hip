// HIP kernels and C ABI of the chunked Kalman filter / RTS smoother
// (bayesnewton/ops.py:149-380, basemodels.py:689-741) for gfx950.
//
// Launch structure of one fused filter + smoother pass (bn_filter_smoother_f32):
//   k_filter_reduce   lane per chunk            -> filtering element per chunk
//   k_fscan_tile      64-thread block per tile  -> tile-local inclusive prefixes (pairs + LDS Hillis-Steele)
//   k_fscan_series    block per series          -> exclusive tile carries
//   k_filter_pass     lane per chunk            -> filtered moments (scratch), loglik partials,
//                                                  smoothing element per chunk
//   k_sscan_tile      block per tile            -> tile-local inclusive suffixes
//   k_sscan_series    block per series          -> exclusive tile carries (from the end)
//   k_smoother_pass   lane per chunk            -> posterior mean / variance
//   k_series_sum      block per series          -> FP64 log-likelihood per series
// A series that fits in one chunk skips both scans (the sequential filter).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <mutex>
#include <string>
#include <vector>

#include "bn_grad.h"
#include "bn_internal.h"

using namespace bn;

namespace {
thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define BN_HIP_CHECK(expr)                                                              \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) return fail(BN_EHIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)


constexpr int kBlock = 256;  // lane-per-chunk kernels
constexpr int kScan = kTile;  // scan kernels: one thread per element of a tile
using CT = double;  // compute type of the recurrences (see bn_lanes.h)
}  // namespace

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
template <int K1, int K2>
__global__ __launch_bounds__(kBlock, 2) void k_filter_reduce(PlanView pl, HyperView hv, const float* __restrict__ dt,
                                                          const float* __restrict__ y, const float* __restrict__ r,
                                                          SoA agg) {
  const int32_t j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= pl.n_chunks) return;
  if ((pl.chunk_flags[j] & (CH_FIRST | CH_LAST)) == (CH_FIRST | CH_LAST)) return;  // single-chunk series
  lane_filter_reduce<K1, K2, CT>(pl, hv, dt, y, r, agg, j);
}

// Tile scan with two elements per thread: each thread combines its pair, a
// Hillis-Steele scan runs over the pair totals in LDS, and the first element of each
// pair takes the previous pair's prefix.  About 3.5 combines per element instead of
// log2(128) = 7 for a scan over the elements themselves.  A workgroup scans one scan
// group of the plan: up to 8 short tiles side by side, each on its own slice of the
// lanes and of the LDS pair slots (most series fit in one short tile, and packing them
// cuts the workgroups, which run one wavefront per SIMD, from one per tile).
template <int D>
__global__ __launch_bounds__(kScan / 2) void k_fscan_tile(PlanView pl, SoA agg, SoA tagg) {
  extern __shared__ agg_t lds[];
  using E = FElem<D, CT>;
  constexpr int NP = kScan / 2;
  const int32_t g = blockIdx.x;
  const int32_t w = pl.group_w[g];
  const int32_t lp = NP / w;                // lanes (pair slots) per tile
  const int32_t h = threadIdx.x / lp;       // place of this lane's tile in the group
  const int32_t tid = threadIdx.x - h * lp;  // lane within the tile
  const int32_t base = h * lp;               // the tile's first LDS pair slot
  const int32_t t = pl.group_tiles[g * kGroupMax + h];
  const int32_t first = t >= 0 ? pl.tile_first[t] : 0;
  const int32_t count = t >= 0 ? pl.tile_count[t] : 0;
  const int32_t np = (count + 1) / 2;
  int32_t npmax = 0;  // every lane runs the group's longest scan (the barriers are shared)
  for (int32_t u = 0; u < w; ++u) {
    const int32_t tu = pl.group_tiles[g * kGroupMax + u];
    if (tu >= 0) npmax = max(npmax, (pl.tile_count[tu] + 1) / 2);
  }
  const int32_t i0 = 2 * tid, i1 = 2 * tid + 1;
  const CSoA ag{agg.base, agg.stride};
  E pr;
  if (i1 < count) {
    E e1;
    soa_load(e1, ag, first + i1);
    fcombine_from<D, CT>(SoaOperand<D, CT>{ag, first + i0}, e1, pr);
  } else if (i0 < count) {
    soa_load(pr, ag, first + i0);
  } else {
    pr.set_identity_elem();
  }
  const SoA ls{lds, NP};
  const CSoA lc{lds, NP};
  for (int32_t off = 1; off < npmax; off <<= 1) {
    soa_store(pr, ls, base + tid);
    __syncthreads();
    if (tid >= off && tid < np) {
      E o;
      fcombine_from<D, CT>(SoaOperand<D, CT>{lc, base + tid - off}, pr, o);
      pr = o;
    }
    __syncthreads();
  }
  soa_store(pr, ls, base + tid);
  __syncthreads();
  // the pair's second output is the inclusive prefix itself; store it first so that
  // the prefix is no longer live during the first output's combine
  if (i1 < count) {
    soa_store(pr, agg, first + i1);
    if (i1 == count - 1) soa_store(pr, tagg, t);
  }
  if (i0 < count) {
    // the pair's first element is read again here rather than held through the scan
    // (an element is 65 doubles); the opaque index keeps the load below the scan
    int32_t at = first + i0;
    asm volatile("" : "+v"(at));
    E e0, out0;
    soa_load(e0, ag, at);
    if (tid == 0) {
      out0 = e0;
    } else {
      fcombine_from<D, CT>(SoaOperand<D, CT>{lc, base + tid - 1}, e0, out0);
    }
    soa_store(out0, agg, first + i0);
    if (i0 == count - 1) soa_store(out0, tagg, t);
  }
}

template <int D>
__global__ __launch_bounds__(kScan) void k_fscan_series(PlanView pl, SoA tagg, SoA carry) {
  extern __shared__ agg_t lds[];
  using E = FElem<D, CT>;
  const int32_t b = blockIdx.x;
  const int32_t nt = pl.series_ntiles[b];
  if (nt <= 1) return;
  const int32_t ft = pl.series_first_tile[b];
  const int32_t tid = threadIdx.x;
  E e;
  if (tid < nt) soa_load(e, CSoA{tagg.base, tagg.stride}, ft + tid);
  else e.set_identity_elem();
  const SoA ls{lds, kScan};
  const CSoA lc{lds, kScan};
  for (int32_t off = 1; off < nt; off <<= 1) {
    soa_store(e, ls, tid);
    __syncthreads();
    if (tid >= off && tid < nt) {
      E o;
      fcombine_from<D, CT>(SoaOperand<D, CT>{lc, tid - off}, e, o);
      e = o;
    }
    __syncthreads();
  }
  if (tid < nt - 1) soa_store(e, carry, ft + tid + 1);
}

// Two wavefronts per SIMD up to state dimension 5 (the one-step-ahead input loads fit
// in 256 registers with a few spills); larger states keep the register file to one.
template <int K1, int K2, bool WANT_S>
__global__ __launch_bounds__(kBlock, (K1 + K2 <= 5 ? 2 : 1)) void k_filter_pass(
    PlanView pl, HyperView hv, const float* __restrict__ dt, const float* __restrict__ y, const float* __restrict__ r,
    const uint8_t* __restrict__ mask, CSoA fagg, CSoA fcarry, FilterPassOut out) {
  const int32_t j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= pl.n_chunks) return;
  lane_filter_pass<K1, K2, CT, WANT_S>(pl, hv, dt, y, r, mask, fagg, fcarry, out, j);
}

template <int K1, int K2>
__global__ __launch_bounds__(kBlock) void k_smoother_reduce(PlanView pl, HyperView hv, const float* __restrict__ dts,
                                                            CStateIO fin, SoA sagg) {
  const int32_t j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= pl.n_chunks) return;
  lane_smoother_reduce<K1, K2, CT>(pl, hv, dts, fin, sagg, j);
}

// Suffix scan of the smoother elements of each tile, one element per thread, over the
// scan groups of the plan as in k_fscan_tile (a tile of a group of width w on
// kScan / w lanes).
template <int D>
__global__ __launch_bounds__(kScan) void k_sscan_tile(PlanView pl, SoA agg, SoA tagg) {
  extern __shared__ agg_t lds[];
  using E = SElem<D, CT>;
  const int32_t g = blockIdx.x;
  const int32_t w = pl.group_w[g];
  const int32_t lp = kScan / w;
  const int32_t h = threadIdx.x / lp;
  const int32_t tid = threadIdx.x - h * lp;
  const int32_t base = h * lp;
  const int32_t t = pl.group_tiles[g * kGroupMax + h];
  const int32_t first = t >= 0 ? pl.tile_first[t] : 0;
  const int32_t count = t >= 0 ? pl.tile_count[t] : 0;
  int32_t cmax = 0;  // every lane runs the group's longest scan (the barriers are shared)
  for (int32_t u = 0; u < w; ++u) {
    const int32_t tu = pl.group_tiles[g * kGroupMax + u];
    if (tu >= 0) cmax = max(cmax, pl.tile_count[tu]);
  }
  E e;
  if (tid < count) soa_load(e, CSoA{agg.base, agg.stride}, first + tid);
  else e.set_identity_elem();
  const SoA ls{lds, kScan};
  const CSoA lc{lds, kScan};
  for (int32_t off = 1; off < cmax; off <<= 1) {
    soa_store(e, ls, base + tid);
    __syncthreads();
    if (tid + off < count) {
      E p, o;
      soa_load(p, lc, base + tid + off);
      scombine<D, CT>(e, p, o);
      e = o;
    }
    __syncthreads();
  }
  if (tid < count) soa_store(e, agg, first + tid);
  if (tid == 0 && count > 0) soa_store(e, tagg, t);
}

template <int D>
__global__ __launch_bounds__(kScan) void k_sscan_series(PlanView pl, SoA tagg, SoA carry) {
  extern __shared__ agg_t lds[];
  using E = SElem<D, CT>;
  const int32_t b = blockIdx.x;
  const int32_t nt = pl.series_ntiles[b];
  if (nt <= 1) return;
  const int32_t ft = pl.series_first_tile[b];
  const int32_t tid = threadIdx.x;
  E e;
  if (tid < nt) soa_load(e, CSoA{tagg.base, tagg.stride}, ft + tid);
  else e.set_identity_elem();
  const SoA ls{lds, kScan};
  const CSoA lc{lds, kScan};
  for (int32_t off = 1; off < nt; off <<= 1) {
    soa_store(e, ls, tid);
    __syncthreads();
    if (tid + off < nt) {
      E p, o;
      soa_load(p, lc, tid + off);
      scombine<D, CT>(e, p, o);
      e = o;
    }
    __syncthreads();
  }
  if (tid >= 1 && tid < nt) soa_store(e, carry, ft + tid - 1);
}

template <int K1, int K2, class OT>
__global__ __launch_bounds__(kBlock) void k_smoother_pass(PlanView pl, HyperView hv, CSoA sagg, CSoA scarry,
                                                          SmootherPassIO<OT> io) {
  const int32_t j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= pl.n_chunks) return;
  lane_smoother_pass<K1, K2, CT, OT>(pl, hv, sagg, scarry, io, j);
}

// Fused path: smoothed marginals by the Bryson-Frazier recursion from the filter's
// predicted moments (lane_smoother_mbf).
template <int K1, int K2>
__global__ __launch_bounds__(kBlock) void k_smoother_mbf(PlanView pl, HyperView hv, const float* __restrict__ dt,
                                                         const float* __restrict__ y, const float* __restrict__ r,
                                                         CSoA sagg, CSoA scarry, CStateIO fin,
                                                         const hrow_t* __restrict__ hrow, float* pmean,
                                                         float* pvar, int32_t* chst) {
  const int32_t j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= pl.n_chunks) return;
  lane_smoother_mbf<K1, K2, CT>(pl, hv, dt, y, r, sagg, scarry, fin, hrow, pmean, pvar, chst, j);
}

// Two launches over the column classes of lane_grad_col: the NZ unit tangents (HK = -1)
// and the NH hyperparameter tangents (HK = 2: each wave takes the lengthscale or the
// variance instance, so both run in one launch and neither carries the other's terms).
template <int K1, int K2, int HK>
__global__ __launch_bounds__(kBlock) void k_grad_chunk(PlanView pl, HyperView hv, const float* __restrict__ dt,
                                                       const float* __restrict__ y, const float* __restrict__ r,
                                                       const uint8_t* __restrict__ mask, CSoA fagg, CSoA fcarry,
                                                       GradOut out) {
  // lane l -> column u = l / stride of the class, chunk l % stride: a wave covers
  // consecutive chunks of one column
  using GD = GradDims<K1, K2>;
  constexpr int NU = HK < 0 ? GD::NZ : (HK == 2 ? GD::NH : GD::NH / 2);
  const int64_t l = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int u = (int)(l / out.stride);
  const int32_t j = (int32_t)(l - (int64_t)u * out.stride);
  if (u >= NU || j >= pl.n_chunks) return;
  if constexpr (HK < 0) {
    lane_grad_col<K1, K2, -1>(pl, hv, dt, y, r, mask, fagg, fcarry, out, j, u);
  } else if ((u & 1) == 0) {  // column NZ + u: lengthscale for even u, variance for odd
    lane_grad_col<K1, K2, 0>(pl, hv, dt, y, r, mask, fagg, fcarry, out, j, GD::NZ + u);
  } else {
    lane_grad_col<K1, K2, 1>(pl, hv, dt, y, r, mask, fagg, fcarry, out, j, GD::NZ + u);
  }
}

// The chunk elements e = (M, O, f, c) of a series compose associatively (chunk order):
//   (n o a) = (Mn Ma, Mn Oa + On, fa + Ma' fn, ca + cn + Oa' fn),
// and the gradient is c of the whole series' composition (its incoming tangent is 0).
// A tree of groups: each k_grad_group launch composes groups of kGradGroup consecutive
// elements of a level (elements `span` chunks apart; one workgroup per group, the result
// replaces the group's first element), the next level's span is kGradGroup times larger,
// and k_grad_total composes the at most kGradGroup elements of the last level.  The
// dependent chain of a long series is then kGradGroup compositions per level, not the
// series' chunk count / 64.  Layout of e: M column-major (NZ*NZ), O (NH columns of NZ),
// f (NZ), c (NH); field q of chunk j at agg[q * stride + j].
constexpr int kGradGroup = 16;

template <int NZ, int NH>
__device__ void compose_elem(const double* __restrict__ a, const double* __restrict__ n, double* __restrict__ out) {
  constexpr int MM = NZ * NZ, FO = MM, FF = MM + NH * NZ, FC = FF + NZ, NF = FC + NH;
  for (int e = threadIdx.x; e < NF; e += blockDim.x) {
    double v;
    if (e < MM) {
      const int q = e / NZ, p = e - q * NZ;
      v = 0.0;
      for (int k = 0; k < NZ; ++k) v = fma(n[k * NZ + p], a[q * NZ + k], v);
    } else if (e < FF) {
      const int h = (e - FO) / NZ, p = (e - FO) - h * NZ;
      v = n[e];
      for (int k = 0; k < NZ; ++k) v = fma(n[k * NZ + p], a[FO + h * NZ + k], v);
    } else if (e < FC) {
      const int q = e - FF;
      v = a[e];
      for (int k = 0; k < NZ; ++k) v = fma(a[q * NZ + k], n[FF + k], v);
    } else {
      const int h = e - FC;
      v = a[e] + n[e];
      for (int k = 0; k < NZ; ++k) v = fma(a[FO + h * NZ + k], n[FF + k], v);
    }
    out[e] = v;
  }
}

// Composes `cnt` elements at chunk indices j0, j0 + step, ... into acc (LDS, NF doubles);
// the next element is read from memory into registers while the current one is composed.
template <int NZ, int NH, int NT>
__device__ const double* compose_run(const double* __restrict__ agg, int64_t stride, int64_t j0, int64_t step,
                                     int32_t cnt, double* buf) {
  constexpr int NF = NZ * NZ + NH * NZ + NZ + NH;
  constexpr int NR = (NF + NT - 1) / NT;
  double* ac = buf;
  double* nx = buf + NF;
  double* out = buf + 2 * NF;
  double pre[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const int e = threadIdx.x + r * NT;
    if (e < NF) ac[e] = agg[(int64_t)e * stride + j0];
    pre[r] = (cnt > 1 && e < NF) ? agg[(int64_t)e * stride + j0 + step] : 0.0;
  }
  for (int32_t i = 1; i < cnt; ++i) {
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int e = threadIdx.x + r * NT;
      if (e < NF) nx[e] = pre[r];
      pre[r] = (i + 1 < cnt && e < NF) ? agg[(int64_t)e * stride + j0 + (int64_t)(i + 1) * step] : 0.0;
    }
    __syncthreads();
    compose_elem<NZ, NH>(ac, nx, out);
    __syncthreads();
    double* t = ac;
    ac = out;
    out = t;
  }
  __syncthreads();
  return ac;
}

template <int K1, int K2>
__global__ __launch_bounds__((GradDims<K1, K2>::NF <= 128 ? 128 : 256)) void k_grad_group(PlanView pl,
                                                                                      double* __restrict__ agg,
                                                                                      int64_t stride, int32_t span) {
  using GD = GradDims<K1, K2>;
  constexpr int NT = GD::NF <= 128 ? 128 : 256;
  __shared__ double buf[3 * GD::NF];
  const int32_t b = blockIdx.x;
  const int32_t nc = pl.series_nchunks[b];
  const int32_t ne = (nc + span - 1) / span;  // elements of this level in the series
  const int32_t first = blockIdx.y * kGradGroup;
  if (first >= ne) return;
  const int64_t j0 = (int64_t)pl.series_first_chunk[b] + (int64_t)first * span;
  const int32_t cnt = min(kGradGroup, ne - first);
  const double* ac = compose_run<GD::NZ, GD::NH, NT>(agg, stride, j0, span, cnt, buf);
  for (int e = threadIdx.x; e < GD::NF; e += NT) agg[(int64_t)e * stride + j0] = ac[e];
}

template <int K1, int K2>
__global__ __launch_bounds__((GradDims<K1, K2>::NF <= 128 ? 128 : 256)) void k_grad_total(PlanView pl,
                                                                                      const double* __restrict__ agg,
                                                                                      int64_t stride, int32_t span,
                                                                                      double* __restrict__ grad) {
  using GD = GradDims<K1, K2>;
  constexpr int NT = GD::NF <= 128 ? 128 : 256;
  constexpr int FC = GD::NZ * GD::NZ + GD::NH * GD::NZ + GD::NZ;
  __shared__ double buf[3 * GD::NF];
  const int32_t b = blockIdx.x;
  const int32_t nc = pl.series_nchunks[b];
  if (nc == 0) {
    if ((int)threadIdx.x < GD::NH) grad[(int64_t)b * GD::NH + threadIdx.x] = 0.0;
    return;
  }
  const int32_t ng = (nc + span - 1) / span;  // elements of the last level
  const double* ac = compose_run<GD::NZ, GD::NH, NT>(agg, stride, pl.series_first_chunk[b], span, ng, buf);
  if ((int)threadIdx.x < GD::NH) grad[(int64_t)b * GD::NH + threadIdx.x] = ac[FC + threadIdx.x];
}

// ---- prediction (MarkovGaussianProcess.predict, basemodels.py:766-816) --------------
// One thread per test time.  temporal_conditional / predict_from_state_ /
// compute_conditional_statistics (utils.py:99-215): the training times of the test
// point's series are augmented with sentinels at -/+1e10 holding (0, Pinf); the left
// neighbour is the last training time < x*, the state at x* given the two neighbours
// u-, u+ is N(P [u-; u+], T) with
//   Q-+ = Q_b + A_b Q_f A_b',  V = (Q-+ + 1e-8 I)^-1 A_b,  W = Q_f V',
//   P = [A_f - W A_b A_f, W],  T = Q_f - (A_b Q_f)' V Q_f,
// (A_f, Q_f) over x* - x-, (A_b, Q_b) over x+ - x*, and the joint smoothed moments of
// (u-, u+) with cross covariance G- P+ (G- the smoother gain at u-).  Outputs H-projected
// mean and variance.
struct PredictIO {
  const int64_t* train_off;  // [n_series + 1] CSR offsets of the training times
  const double* x_train;     // training times (sorted within each series)
  const double* means;       // smoothed state means, d tiled planes
  const double* covs;        // smoothed state covariances, d*d tiled planes (row-major)
  const double* gains;       // smoother gains, d*d tiled planes
  const int64_t* test_off;   // [n_series + 1] CSR offsets of the test times
  const double* x_test;
  int64_t n_test;
  double* out_mean;
  double* out_var;
};

constexpr double kPredictEdge = 1e10;   // utils.py / basemodels.py:795-796 sentinel times
constexpr double kPredictJitter = 1e-8;  // utils.py:207

// Prediction runs in three launches.  k_predict_locate finds each test time's series and
// neighbour count (two binary searches, at the occupancy of a small kernel) and a sort
// key: the slot of the training state it reads the most (the left neighbour; the right
// one before the series' first training time).  A radix sort orders the test times by
// that key, and k_predict works in that order, so that neighbouring lanes read
// neighbouring slots of the state planes (in input order each lane's reads of the 55
// state values each touched a cache line of their own).
__global__ __launch_bounds__(kBlock) void k_predict_locate(PlanView pl, PredictIO io, uint32_t* __restrict__ key,
                                                           uint32_t* __restrict__ idx, int64_t* __restrict__ loc) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= io.n_test) return;
  // series of test point i: the last b with test_off[b] <= i
  int32_t lo = 0, hi = pl.n_series - 1;
  while (lo < hi) {
    const int32_t mid = (lo + hi + 1) >> 1;
    if (io.test_off[mid] <= i) lo = mid; else hi = mid - 1;
  }
  const int32_t b = lo;
  const double xs = io.x_test[i];
  const int64_t t0 = io.train_off[b];
  const int64_t n = io.train_off[b + 1] - t0;
  // cnt = #{training times < x*}
  int64_t a = 0, c = n;
  while (a < c) {
    const int64_t mid = (a + c) >> 1;
    if (io.x_train[t0 + mid] < xs) a = mid + 1; else c = mid;
  }
  loc[i] = ((int64_t)b << 32) | (int64_t)(uint32_t)a;
  if (key != nullptr) {
    uint32_t k = (uint32_t)pl.n_slots;  // no training state: after every slot
    if (n > 0) {
      const int64_t r = a > 0 ? a - 1 : 0;
      k = (uint32_t)pl.slot(pl.series_first_chunk[b] + (int32_t)(r / pl.chunk_len), (int32_t)(r % pl.chunk_len));
    }
    key[i] = k;
    idx[i] = (uint32_t)i;
  }
}

// One thread per test time, in the order `order` gives (NULL: input order).
template <int K1, int K2>
__global__ __launch_bounds__(kBlock) void k_predict(PlanView pl, HyperView hv, PredictIO io,
                                                    const uint32_t* __restrict__ order,
                                                    const int64_t* __restrict__ loc) {
  using S = Spec<K1, K2>;
  constexpr int D = S::D;
  const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (p >= io.n_test) return;
  const int64_t i = order != nullptr ? (int64_t)order[p] : p;
  const int64_t li = loc[i];
  const int32_t b = (int32_t)(li >> 32);
  const int64_t cnt = (int64_t)(uint32_t)(li & 0xffffffff);
  const double xs = io.x_test[i];
  const int64_t t0 = io.train_off[b];
  const int64_t n = io.train_off[b + 1] - t0;
  const Model<K1, K2> mdl = load_model<K1, K2>(hv, b);
  const int32_t fc = pl.series_first_chunk[b];
  auto slot_of = [&](int64_t k) {
    return pl.slot(fc + (int32_t)(k / pl.chunk_len), (int32_t)(k % pl.chunk_len));
  };
  const int64_t ns = pl.n_slots;
  double ml[D], mr[D], Pl[D][D], Pr[D][D], G[D][D];
  double xl, xr;
  if (cnt == 0) {
    xl = -kPredictEdge;
    for (int p = 0; p < D; ++p) { ml[p] = 0.0; for (int q = 0; q < D; ++q) { Pl[p][q] = mdl.Pinf[p][q]; G[p][q] = 0.0; } }
  } else {
    const int64_t s = slot_of(cnt - 1);
    xl = io.x_train[t0 + cnt - 1];
    for (int p = 0; p < D; ++p) {
      ml[p] = io.means[p * ns + s];
      for (int q = 0; q < D; ++q) {
        Pl[p][q] = io.covs[(int64_t)(p * D + q) * ns + s];
        G[p][q] = io.gains[(int64_t)(p * D + q) * ns + s];
      }
    }
  }
  if (cnt == n) {
    xr = kPredictEdge;
    for (int p = 0; p < D; ++p) { mr[p] = 0.0; for (int q = 0; q < D; ++q) Pr[p][q] = mdl.Pinf[p][q]; }
  } else {
    const int64_t s = slot_of(cnt);
    xr = io.x_train[t0 + cnt];
    for (int p = 0; p < D; ++p) {
      mr[p] = io.means[p * ns + s];
      for (int q = 0; q < D; ++q) Pr[p][q] = io.covs[(int64_t)(p * D + q) * ns + s];
    }
  }
  double Af[D][D], Qf[D][D], Ab[D][D], Qb[D][D];
  mdl.discretise(xs - xl, Af, Qf);
  mdl.discretise(xr - xs, Ab, Qb);
  // ABQ = A_b Q_f ; Qmp = Q_b + ABQ A_b' + jitter
  double ABQ[D][D], L[D][D];
  for (int p = 0; p < D; ++p)
    for (int q = 0; q < D; ++q) {
      double t = 0.0;
      for (int k = 0; k < D; ++k) t = fma(Ab[p][k], Qf[k][q], t);
      ABQ[p][q] = t;
    }
  for (int p = 0; p < D; ++p)
    for (int q = 0; q < D; ++q) {
      double t = Qb[p][q] + (p == q ? kPredictJitter : 0.0);
      for (int k = 0; k < D; ++k) t = fma(ABQ[p][k], Ab[q][k], t);
      L[p][q] = t;
    }
  chol_spd<D>(L);
  // V = Qmp^-1 A_b
  double V[D][D];
  for (int p = 0; p < D; ++p)
    for (int q = 0; q < D; ++q) V[p][q] = Ab[p][q];
  lower_solve<D, D, double>(L, V);
  upper_t_solve<D, D, double>(L, V);
  // h W = h Q_f V' (row), h T h' = h Q_f h' - (A_b Q_f h')' (V Q_f h')
  double hQ[D], hW[D], QfH[D];
  for (int q = 0; q < D; ++q) {
    double t = 0.0;
    for (int p = 0; p < D; ++p) if (S::h(p)) t += Qf[p][q];
    hQ[q] = t;
    QfH[q] = t;  // Q_f symmetric: Q_f h' = (h Q_f)'
  }
  for (int q = 0; q < D; ++q) {
    double t = 0.0;
    for (int k = 0; k < D; ++k) t = fma(hQ[k], V[q][k], t);
    hW[q] = t;
  }
  double hTh = 0.0;
  for (int p = 0; p < D; ++p) if (S::h(p)) hTh += QfH[p];
  {
    double VQh[D], ABQh[D];  // V Q_f h' and A_b Q_f h'
    for (int p = 0; p < D; ++p) {
      double t = 0.0, u = 0.0;
      for (int k = 0; k < D; ++k) {
        t = fma(V[p][k], QfH[k], t);
        if (S::h(k)) u += ABQ[p][k];
      }
      VQh[p] = t;
      ABQh[p] = u;
    }
    for (int p = 0; p < D; ++p) hTh -= ABQh[p] * VQh[p];
  }
  // u1 = h (A_f - W A_b A_f), u2 = h W
  double hWAb[D], u1[D];
  for (int q = 0; q < D; ++q) {
    double t = 0.0;
    for (int k = 0; k < D; ++k) t = fma(hW[k], Ab[k][q], t);
    hWAb[q] = t;
  }
  for (int q = 0; q < D; ++q) {
    double t = 0.0;
    for (int k = 0; k < D; ++k) t += ((S::h(k) ? 1.0 : 0.0) - hWAb[k]) * Af[k][q];
    u1[q] = t;
  }
  // mean = u1 ml + u2 mr ; var = u1 Pl u1' + 2 u1 (G Pr) u2' + u2 Pr u2' + h T h'
  double mean = 0.0;
  for (int p = 0; p < D; ++p) mean += u1[p] * ml[p] + hW[p] * mr[p];
  double Pru2[D];
  for (int p = 0; p < D; ++p) {
    double t = 0.0;
    for (int q = 0; q < D; ++q) t = fma(Pr[p][q], hW[q], t);
    Pru2[p] = t;
  }
  double var = hTh;
  for (int p = 0; p < D; ++p) {
    double t = 0.0, g = 0.0;
    for (int q = 0; q < D; ++q) { t = fma(Pl[p][q], u1[q], t); g = fma(G[p][q], Pru2[q], g); }
    var += u1[p] * t + 2.0 * u1[p] * g + hW[p] * Pru2[p];
  }
  io.out_mean[i] = mean;
  io.out_var[i] = var;
}

// Deterministic per-series sum of a per-chunk FP64 array (fixed tree order).
// With chst / status, also the OR of the series' per-chunk status bits.
__global__ __launch_bounds__(kBlock) void k_series_sum(PlanView pl, const double* __restrict__ per_chunk,
                                                       double* __restrict__ out, const int32_t* __restrict__ chst,
                                                       int32_t* __restrict__ status) {
  __shared__ double red[kBlock];
  __shared__ int32_t ored[kBlock];
  const int32_t b = blockIdx.x;
  const int32_t c0 = pl.series_first_chunk[b];
  const int32_t nc = pl.series_nchunks[b];
  const double s = strided_sum(per_chunk + c0, nc, threadIdx.x, kBlock);
  const int32_t f = chst != nullptr ? strided_or(chst + c0, nc, threadIdx.x, kBlock) : 0;
  red[threadIdx.x] = s;
  ored[threadIdx.x] = f;
  __syncthreads();
  for (int32_t w = kBlock / 2; w > 0; w >>= 1) {
    if ((int32_t)threadIdx.x < w) {
      red[threadIdx.x] += red[threadIdx.x + w];
      ored[threadIdx.x] |= ored[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[b] = red[0];
    if (status != nullptr) status[b] = ored[0];
  }
}

__device__ __forceinline__ bool slot_to_step(const PlanView& pl, int64_t slot, int64_t& csr) {
  const int32_t lane = (int32_t)(slot & ((1 << pl.gshift) - 1));
  const int64_t q = slot >> pl.gshift;
  const int32_t k = (int32_t)(q % pl.chunk_len);
  const int32_t j = (int32_t)(((q / pl.chunk_len) << pl.gshift) + lane);
  if (j >= pl.n_chunks || k >= pl.chunk_len_[j]) return false;
  csr = pl.chunk_start[j] + k;
  return true;
}

__global__ void k_gather_f32(PlanView pl, const float* __restrict__ src, float* __restrict__ dst, int32_t width,
                             float pad) {
  const int64_t slot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= pl.n_slots) return;
  int64_t csr;
  const bool ok = slot_to_step(pl, slot, csr);
  for (int32_t f = 0; f < width; ++f) dst[f * pl.n_slots + slot] = ok ? src[csr * width + f] : pad;
}

__global__ void k_gather_f64(PlanView pl, const double* __restrict__ src, double* __restrict__ dst, int32_t width,
                             double pad) {
  const int64_t slot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= pl.n_slots) return;
  int64_t csr;
  const bool ok = slot_to_step(pl, slot, csr);
  for (int32_t f = 0; f < width; ++f) dst[f * pl.n_slots + slot] = ok ? src[csr * width + f] : pad;
}

__global__ void k_scatter_f64(PlanView pl, const double* __restrict__ src, double* __restrict__ dst, int32_t width) {
  const int64_t slot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= pl.n_slots) return;
  int64_t csr;
  if (!slot_to_step(pl, slot, csr)) return;
  for (int32_t f = 0; f < width; ++f) dst[csr * width + f] = src[f * pl.n_slots + slot];
}

__global__ void k_gather_u8(PlanView pl, const uint8_t* __restrict__ src, uint8_t* __restrict__ dst, uint8_t pad) {
  const int64_t slot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= pl.n_slots) return;
  int64_t csr;
  const bool ok = slot_to_step(pl, slot, csr);
  dst[slot] = ok ? src[csr] : pad;
}

__global__ void k_scatter_f32(PlanView pl, const float* __restrict__ src, float* __restrict__ dst, int32_t width) {
  const int64_t slot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= pl.n_slots) return;
  int64_t csr;
  if (!slot_to_step(pl, slot, csr)) return;
  for (int32_t f = 0; f < width; ++f) dst[csr * width + f] = src[f * pl.n_slots + slot];
}

template <int K1, int K2>
__global__ void k_discretise(HyperView hv, int64_t n, const float* __restrict__ dt, float* __restrict__ A,
                             float* __restrict__ Q) {
  constexpr int D = K1 + K2;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Model<K1, K2> mdl = load_model<K1, K2>(hv, 0);
  float a[D][D], q[D][D];
  mdl.discretise((double)dt[i], a, q);
  for (int r = 0; r < D; ++r)
    for (int c = 0; c < D; ++c) {
      A[(i * D + r) * D + c] = a[r][c];
      Q[(i * D + r) * D + c] = q[r][c];
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
namespace {

int state_dim(const bn_kernel_desc* k) {
  int d = 0;
  for (int c = 0; c < k->n_comp; ++c) d += k->kind[c];
  return d;
}

int check_desc(const bn_kernel_desc* k) {
  if (k == nullptr) return fail(BN_EINVAL, "kernel descriptor is NULL");
  if (k->n_comp < 1 || k->n_comp > 2) return fail(BN_EUNSUPPORTED, "kernel must have 1 or 2 Matern components");
  for (int c = 0; c < k->n_comp; ++c)
    if (k->kind[c] < 1 || k->kind[c] > 3) return fail(BN_EUNSUPPORTED, "component kind must be Matern12/32/52");
  return BN_OK;
}

struct WsLayout {
  int64_t fagg, ftagg, ftcarry, sagg, stagg, stcarry, fbuf, ell, sp1, sp2, chst, total;
};

int64_t align256(int64_t x) { return (x + 255) & ~int64_t(255); }

WsLayout ws_layout(const PlanHost& p, int D) {
  const int64_t FA = D * D + 2 * D + D * (D + 1);
  const int64_t SA = D * D + D + D * (D + 1) / 2;
  const int64_t NS = D * (D + 1) / 2;
  const int64_t nt = p.n_tiles > 0 ? p.n_tiles : 1;
  WsLayout w;
  int64_t o = 0;
  const int64_t ea = sizeof(agg_t), es = sizeof(state_t);
  w.fagg = o;    o = align256(o + FA * p.n_chunks_pad * ea);
  w.ftagg = o;   o = align256(o + FA * nt * ea);
  w.ftcarry = o; o = align256(o + FA * nt * ea);
  w.sagg = o;    o = align256(o + SA * p.n_chunks_pad * ea);
  w.stagg = o;   o = align256(o + SA * nt * ea);
  w.stcarry = o; o = align256(o + SA * nt * ea);
  w.fbuf = o;    o = align256(o + (D + NS + D + 1) * p.n_slots * es);  // state planes, then H P-, H m-
  w.ell = o;     o = align256(o + p.n_chunks_pad * 8);
  w.sp1 = o;     o = align256(o + p.n_chunks_pad * 8);
  w.sp2 = o;     o = align256(o + p.n_chunks_pad * 8);
  w.chst = o;    o = align256(o + p.n_chunks_pad * 4);
  w.total = o;
  return w;
}

struct Bufs {
  SoA fagg, ftagg, ftcarry, sagg, stagg, stcarry;
  state_t* fbuf;
  double* ell;
  int32_t* chst;
};

Bufs ws_bufs(const PlanHost& p, int D, void* ws) {
  WsLayout w = ws_layout(p, D);
  char* base = (char*)ws;
  const int64_t nt = p.n_tiles > 0 ? p.n_tiles : 1;
  Bufs b;
  b.fagg = SoA{(agg_t*)(base + w.fagg), p.n_chunks_pad};
  b.ftagg = SoA{(agg_t*)(base + w.ftagg), nt};
  b.ftcarry = SoA{(agg_t*)(base + w.ftcarry), nt};
  b.sagg = SoA{(agg_t*)(base + w.sagg), p.n_chunks_pad};
  b.stagg = SoA{(agg_t*)(base + w.stagg), nt};
  b.stcarry = SoA{(agg_t*)(base + w.stcarry), nt};
  b.fbuf = (state_t*)(base + w.fbuf);
  b.ell = (double*)(base + w.ell);
  b.chst = (int32_t*)(base + w.chst);
  return b;
}

inline CSoA cs(SoA s) { return CSoA{s.base, s.stride}; }

inline dim3 lane_grid(const PlanHost& p) { return dim3((p.n_chunks + kBlock - 1) / kBlock); }

bool needs_scan(const PlanHost& p) { return p.needs_scan; }

// filter side: reduce + both scans
template <int K1, int K2>
void launch_filter_scan(const bn_plan* plan, HyperView hv, const float* dt, const float* y, const float* r,
                        const Bufs& b, hipStream_t st) {
  constexpr int D = K1 + K2;
  const PlanHost& p = plan->host;
  if (!needs_scan(p)) return;
  const size_t ldsF = sizeof(agg_t) * FElem<D, CT>::NF * kScan;
  { BN_TIMED(KID_FILTER_REDUCE, st);
  hipLaunchKernelGGL((k_filter_reduce<K1, K2>), lane_grid(p), dim3(kBlock), 0, st, plan->view, hv, dt, y, r, b.fagg); }
  { BN_TIMED(KID_FSCAN_TILE, st);
  if (plan->view.n_groups > 0)
    hipLaunchKernelGGL((k_fscan_tile<D>), dim3(plan->view.n_groups), dim3(kScan / 2), ldsF / 2, st, plan->view, b.fagg,
                     b.ftagg); }
  { BN_TIMED(KID_FSCAN_SERIES, st);
  hipLaunchKernelGGL((k_fscan_series<D>), dim3(p.n_series), dim3(kScan), ldsF, st, plan->view, b.ftagg, b.ftcarry); }
}

template <int D>
void launch_smoother_scan(const bn_plan* plan, const Bufs& b, hipStream_t st) {
  const PlanHost& p = plan->host;
  if (!needs_scan(p)) return;
  const size_t ldsS = sizeof(agg_t) * SElem<D, CT>::NF * kScan;
  { BN_TIMED(KID_SSCAN_TILE, st);
  if (plan->view.n_groups > 0)
    hipLaunchKernelGGL((k_sscan_tile<D>), dim3(plan->view.n_groups), dim3(kScan), ldsS, st, plan->view, b.sagg,
                       b.stagg); }
  { BN_TIMED(KID_SSCAN_SERIES, st);
  hipLaunchKernelGGL((k_sscan_series<D>), dim3(p.n_series), dim3(kScan), ldsS, st, plan->view, b.stagg, b.stcarry); }
}

template <int K1, int K2>
int run_filter_smoother(const bn_plan* plan, HyperView hv, const float* dt, const float* y, const float* r,
                        const uint8_t* mask, float* pmean, float* pvar, double* loglik, int32_t* status, void* ws,
                        hipStream_t st) {
  constexpr int D = K1 + K2;
  const PlanHost& p = plan->host;
  Bufs b = ws_bufs(p, D, ws);
  launch_filter_scan<K1, K2>(plan, hv, dt, y, r, b, st);
  FilterPassOut fo;
  // the filter pass keeps the predicted moments: the smoother's Bryson-Frazier
  // recursion runs on them without factorising a covariance per step
  fo.state = StateIO{b.fbuf, b.fbuf + (int64_t)D * p.n_slots, true};
  fo.predict = true;
  fo.hrow = (hrow_t*)(b.fbuf + (int64_t)(D + D * (D + 1) / 2) * p.n_slots);
  fo.sagg = b.sagg;
  fo.ell = b.ell;
  fo.status = status != nullptr ? b.chst : nullptr;
  {
    BN_TIMED(KID_FILTER_PASS, st);
    hipLaunchKernelGGL((k_filter_pass<K1, K2, true>), lane_grid(p), dim3(kBlock), 0, st, plan->view, hv, dt, y, r,
                       mask, cs(b.fagg), cs(b.ftcarry), fo);
  }
  launch_smoother_scan<D>(plan, b, st);
  {
    BN_TIMED(KID_SMOOTHER_MBF, st);
    hipLaunchKernelGGL((k_smoother_mbf<K1, K2>), lane_grid(p), dim3(kBlock), 0, st, plan->view, hv, dt, y, r,
                       cs(b.sagg), cs(b.stcarry), CStateIO{b.fbuf, b.fbuf + (int64_t)D * p.n_slots, true}, fo.hrow, pmean,
                       pvar, fo.status);
  }
  {
    BN_TIMED(KID_SERIES_SUM, st);
    hipLaunchKernelGGL(k_series_sum, dim3(p.n_series), dim3(kBlock), 0, st, plan->view, b.ell, loglik,
                       (const int32_t*)fo.status, status);
  }
  BN_HIP_CHECK(hipGetLastError());
  return BN_OK;
}

template <int K1, int K2>
int run_kalman_filter(const bn_plan* plan, HyperView hv, const float* dt, const float* y, const float* r,
                      const uint8_t* mask, double* means, double* covs, double* loglik, uint32_t flags,
                      int32_t* status, void* ws, hipStream_t st) {
  constexpr int D = K1 + K2;
  const PlanHost& p = plan->host;
  Bufs b = ws_bufs(p, D, ws);
  launch_filter_scan<K1, K2>(plan, hv, dt, y, r, b, st);
  FilterPassOut fo;
  fo.state = StateIO{means, covs, false};
  fo.predict = (flags & BN_RETURN_PREDICT) != 0;
  fo.sagg = SoA{nullptr, 0};
  fo.ell = b.ell;
  fo.status = status != nullptr ? b.chst : nullptr;
  hipLaunchKernelGGL((k_filter_pass<K1, K2, false>), lane_grid(p), dim3(kBlock), 0, st, plan->view, hv, dt, y, r,
                     mask, cs(b.fagg), cs(b.ftcarry), fo);
  hipLaunchKernelGGL(k_series_sum, dim3(p.n_series), dim3(kBlock), 0, st, plan->view, b.ell, loglik,
                     (const int32_t*)fo.status, status);
  BN_HIP_CHECK(hipGetLastError());
  return BN_OK;
}

template <int K1, int K2>
int run_loglik_grad(const bn_plan* plan, HyperView hv, const float* dt, const float* y, const float* r,
                    const uint8_t* mask, double* loglik, double* grad, void* ws, void* gws, hipStream_t st) {
  constexpr int D = K1 + K2;
  const PlanHost& p = plan->host;
  Bufs b = ws_bufs(p, D, ws);
  launch_filter_scan<K1, K2>(plan, hv, dt, y, r, b, st);
  GradOut go{(double*)gws, p.n_chunks_pad, b.ell};
  constexpr int64_t NZ = GradDims<K1, K2>::NZ, NH = GradDims<K1, K2>::NH;
  hipLaunchKernelGGL((k_grad_chunk<K1, K2, -1>), dim3((unsigned)((NZ * p.n_chunks_pad + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, st, plan->view, hv, dt, y, r, mask, cs(b.fagg), cs(b.ftcarry), go);
  hipLaunchKernelGGL((k_grad_chunk<K1, K2, 2>), dim3((unsigned)((NH * p.n_chunks_pad + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, st, plan->view, hv, dt, y, r, mask, cs(b.fagg), cs(b.ftcarry), go);
  constexpr int NT = GradDims<K1, K2>::NF <= 128 ? 128 : 256;
  int32_t max_nc = 1;
  for (int32_t v : p.series_nchunks) max_nc = std::max(max_nc, v);
  int32_t span = 1;  // chunks between consecutive elements of the current level
  while ((max_nc + span - 1) / span > kGradGroup) {
    const int32_t ne = (max_nc + span - 1) / span;
    const unsigned groups = (unsigned)((ne + kGradGroup - 1) / kGradGroup);
    hipLaunchKernelGGL((k_grad_group<K1, K2>), dim3((unsigned)p.n_series, groups), dim3(NT), 0, st, plan->view,
                       (double*)gws, (int64_t)p.n_chunks_pad, span);
    span *= kGradGroup;
  }
  hipLaunchKernelGGL((k_grad_total<K1, K2>), dim3((unsigned)p.n_series), dim3(NT), 0, st, plan->view,
                     (const double*)gws, (int64_t)p.n_chunks_pad, span, grad);
  hipLaunchKernelGGL(k_series_sum, dim3(p.n_series), dim3(kBlock), 0, st, plan->view, b.ell, loglik, nullptr,
                       (int32_t*)nullptr);
  BN_HIP_CHECK(hipGetLastError());
  return BN_OK;
}

template <int K1, int K2>
int run_predict(const bn_plan* plan, HyperView hv, PredictIO io, hipStream_t st) {
  if (io.n_test == 0) return BN_OK;
  const int64_t n = io.n_test;
  // the sort keys are slots (32 bits) and the values test indices (32 bits)
  const bool sorted = plan->host.n_slots < (int64_t)UINT32_MAX && n < (int64_t)UINT32_MAX;
  unsigned bits = 1;
  while (bits < 32 && ((uint64_t)1 << bits) <= (uint64_t)plan->host.n_slots) ++bits;
  size_t sort_bytes = 0;
  if (sorted)
    BN_HIP_CHECK(rocprim::radix_sort_pairs(nullptr, sort_bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                           (const uint32_t*)nullptr, (uint32_t*)nullptr, (size_t)n, 0, bits, st));
  // stream-ordered scratch: loc [n] int64, keys and indices [n] x 2 uint32, sort space
  const size_t off_keys = (size_t)align256(n * 8), off_idx = off_keys + (size_t)align256(n * 4 * 2);
  const size_t off_sort = off_idx + (size_t)align256(n * 4 * 2);
  const size_t total = sorted ? off_sort + (size_t)align256((int64_t)sort_bytes) : (size_t)align256(n * 8);
  char* buf = nullptr;
  BN_HIP_CHECK(hipMallocAsync((void**)&buf, total, st));
  int64_t* loc = (int64_t*)buf;
  uint32_t* keys = sorted ? (uint32_t*)(buf + off_keys) : nullptr;
  uint32_t* idx = sorted ? (uint32_t*)(buf + off_idx) : nullptr;
  const dim3 grid((unsigned)((n + kBlock - 1) / kBlock));
  hipLaunchKernelGGL(k_predict_locate, grid, dim3(kBlock), 0, st, plan->view, io, keys, idx, loc);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && sorted) {
    size_t sb = sort_bytes;
    e = rocprim::radix_sort_pairs(buf + off_sort, sb, keys, keys + n, idx, idx + n, (size_t)n, 0, bits, st);
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL((k_predict<K1, K2>), grid, dim3(kBlock), 0, st, plan->view, hv, io,
                       sorted ? (const uint32_t*)(idx + n) : nullptr, (const int64_t*)loc);
    e = hipGetLastError();
  }
  const hipError_t ef = hipFreeAsync(buf, st);
  if (e != hipSuccess) return fail(BN_EHIP, std::string("bn_predict: ") + hipGetErrorString(e));
  BN_HIP_CHECK(ef);
  return BN_OK;
}

template <int K1, int K2>
int run_rts_smoother(const bn_plan* plan, HyperView hv, const float* dts, const double* fmean, const double* fcov,
                     double* means, double* covs, double* gains, uint32_t flags, void* ws, hipStream_t st) {
  constexpr int D = K1 + K2;
  const PlanHost& p = plan->host;
  Bufs b = ws_bufs(p, D, ws);
  CStateIO fin{fmean, fcov, false};
  if (needs_scan(p)) {
    hipLaunchKernelGGL((k_smoother_reduce<K1, K2>), lane_grid(p), dim3(kBlock), 0, st, plan->view, hv, dts, fin,
                       b.sagg);
    launch_smoother_scan<D>(plan, b, st);
  }
  SmootherPassIO<double> io;
  io.fin = fin;
  io.dt = dts;
  io.shifted = true;
  io.out_mean = means;
  io.out_cov = covs;
  io.gains = gains;
  io.full = (flags & BN_RETURN_FULL) != 0;
  hipLaunchKernelGGL((k_smoother_pass<K1, K2, double>), lane_grid(p), dim3(kBlock), 0, st, plan->view, hv,
                     cs(b.sagg), cs(b.stcarry), io);
  BN_HIP_CHECK(hipGetLastError());
  return BN_OK;
}

template <int K1, int K2>
int run_discretise(HyperView hv, int64_t n, const float* dt, float* A, float* Q, hipStream_t st) {
  if (n <= 0) return BN_OK;
  hipLaunchKernelGGL((k_discretise<K1, K2>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, hv, n, dt, A, Q);
  BN_HIP_CHECK(hipGetLastError());
  return BN_OK;
}

// ---------------------------------------------------------------------------
// k_pairs_filter: Kalman filter over pairs of neighbouring states (ops.py:383-426, the
// filter of the sparse Markov model): pair (a, b) -> (b, A b), P- = [[Pbb + 1e-32 I,
// Pbb A'], [A Pbb, A Pbb A' + Q]], observed whole with a full 2d x 2d noise per step.  One
// 64-thread block per series, the 2d x 2d algebra in LDS (2d <= 12), FP64.  Outputs the
// pair moments after every step and the per-series log likelihood.
// ---------------------------------------------------------------------------
template <int K1, int K2>
__global__ __launch_bounds__(64) void k_pairs_filter(HyperView hv, const int64_t* __restrict__ off,
                                                     const double* __restrict__ dt, const double* __restrict__ y,
                                                     const double* __restrict__ R, double* __restrict__ mo,
                                                     double* __restrict__ Po, double* __restrict__ ll) {
  constexpr int D = K1 + K2, E = 2 * D;
  __shared__ double P[E][E], Pm[E][E], L[E][E], X[E][E + 1];
  __shared__ double m[E], mm[E], sv[E], A[D][D], Q[D][D];
  __shared__ double s_ld;
  const int b = blockIdx.x, tid = threadIdx.x;
  const int64_t o0 = off[b], n_steps = off[b + 1] - o0;
  const Model<K1, K2> mdl = load_model<K1, K2>(hv, b);
  for (int e = tid; e < E * E; e += 64) {
    const int i = e / E, j = e % E;
    double v = 0.0;
    if (i < D && j < D) v = mdl.Pinf[i][j];
    if (i >= D && j >= D) v = mdl.Pinf[i - D][j - D];
    P[i][j] = v;
  }
  for (int e = tid; e < E; e += 64) m[e] = 0.0;
  double ell = 0.0;
  __syncthreads();
  for (int64_t n = 0; n < n_steps; ++n) {
    const int64_t s = o0 + n;
    if (tid == 0) {
      double Aa[D][D], Qq[D][D];
      mdl.discretise(dt[s], Aa, Qq);
      for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) { A[i][j] = Aa[i][j]; Q[i][j] = Qq[i][j]; }
    }
    __syncthreads();
    for (int i = tid; i < E; i += 64) {
      double v = 0.0;
      if (i < D) v = m[D + i];
      else
        for (int q = 0; q < D; ++q) v = fma(A[i - D][q], m[D + q], v);
      mm[i] = v;
    }
    for (int e = tid; e < E * E; e += 64) {
      const int i = e / E, j = e % E;
      double v;
      if (i < D && j < D) {
        v = P[D + i][D + j] + (i == j ? 1e-32 : 0.0);
      } else if (i < D) {
        v = 0.0;
        for (int q = 0; q < D; ++q) v = fma(P[D + i][D + q], A[j - D][q], v);
      } else if (j < D) {
        v = 0.0;
        for (int q = 0; q < D; ++q) v = fma(A[i - D][q], P[D + q][D + j], v);
      } else {
        v = Q[i - D][j - D];
        for (int q = 0; q < D; ++q) {
          double t = 0.0;
          for (int r = 0; r < D; ++r) t = fma(P[D + q][D + r], A[j - D][r], t);
          v = fma(A[i - D][q], t, v);
        }
      }
      Pm[i][j] = v;
    }
    __syncthreads();
    // S = P- + R_n ; right-hand sides [P- | y - m-]
    for (int e = tid; e < E * E; e += 64) {
      const int i = e / E, j = e % E;
      L[i][j] = Pm[i][j] + R[s * E * E + e];
      X[i][j] = Pm[i][j];
    }
    for (int i = tid; i < E; i += 64) {
      const double v = y[s * E + i] - mm[i];
      sv[i] = v;
      X[i][E] = v;
    }
    __syncthreads();
    if (tid == 0) {  // Cholesky (lower) of S; a non-positive pivot gives NaN, as the reference
      double ld = 0.0;
      for (int j = 0; j < E; ++j) {
        double d = L[j][j];
        for (int q = 0; q < j; ++q) d -= L[j][q] * L[j][q];
        d = sqrt(d);
        L[j][j] = d;
        ld += 2.0 * log(d);
        for (int i = j + 1; i < E; ++i) {
          double v = L[i][j];
          for (int q = 0; q < j; ++q) v -= L[i][q] * L[j][q];
          L[i][j] = v / d;
        }
      }
      s_ld = ld;
    }
    __syncthreads();
    if (tid <= E) {  // X := S^-1 X, column tid
      const int c = tid;
      for (int i = 0; i < E; ++i) {
        double v = X[i][c];
        for (int q = 0; q < i; ++q) v -= L[i][q] * X[q][c];
        X[i][c] = v / L[i][i];
      }
      for (int i = E - 1; i >= 0; --i) {
        double v = X[i][c];
        for (int q = i + 1; q < E; ++q) v -= L[q][i] * X[q][c];
        X[i][c] = v / L[i][i];
      }
    }
    __syncthreads();
    // m = m- + P- S^-1 v ; P = P- - P- S^-1 P-
    for (int i = tid; i < E; i += 64) {
      double v = mm[i];
      for (int q = 0; q < E; ++q) v = fma(Pm[i][q], X[q][E], v);
      m[i] = v;
      mo[s * E + i] = v;
    }
    for (int e = tid; e < E * E; e += 64) {
      const int i = e / E, j = e % E;
      double v = Pm[i][j];
      for (int q = 0; q < E; ++q) v = fma(-Pm[i][q], X[q][j], v);
      P[i][j] = v;
      Po[s * E * E + e] = v;
    }
    if (tid == 0) {
      double quad = 0.0;
      for (int i = 0; i < E; ++i) quad = fma(sv[i], X[i][E], quad);
      ell += -0.5 * (quad + s_ld + E * kLog2Pi);
    }
    __syncthreads();
  }
  if (tid == 0) ll[b] = ell;
}

template <int K1, int K2>
int run_pairs_filter(HyperView hv, int32_t n_series, const int64_t* off, const double* dt, const double* y,
                     const double* R, double* mo, double* Po, double* ll, hipStream_t st) {
  hipLaunchKernelGGL((k_pairs_filter<K1, K2>), dim3(n_series), dim3(64), 0, st, hv, off, dt, y, R, mo, Po, ll);
  BN_HIP_CHECK(hipGetLastError());
  return BN_OK;
}

// Calls F<K1, K2>::run(args...) for the descriptor's component kinds.
#define BN_DISPATCH(desc, FN, ...)                                              \
  [&]() -> int {                                                                \
    const int k1_ = (desc)->kind[0];                                            \
    const int k2_ = (desc)->n_comp > 1 ? (desc)->kind[1] : 0;                   \
    switch (k1_ * 4 + k2_) {                                                    \
      case 1 * 4 + 0: return FN<1, 0>(__VA_ARGS__);                             \
      case 2 * 4 + 0: return FN<2, 0>(__VA_ARGS__);                             \
      case 3 * 4 + 0: return FN<3, 0>(__VA_ARGS__);                             \
      case 1 * 4 + 1: return FN<1, 1>(__VA_ARGS__);                             \
      case 1 * 4 + 2: return FN<1, 2>(__VA_ARGS__);                             \
      case 1 * 4 + 3: return FN<1, 3>(__VA_ARGS__);                             \
      case 2 * 4 + 1: return FN<2, 1>(__VA_ARGS__);                             \
      case 2 * 4 + 2: return FN<2, 2>(__VA_ARGS__);                             \
      case 2 * 4 + 3: return FN<2, 3>(__VA_ARGS__);                             \
      case 3 * 4 + 1: return FN<3, 1>(__VA_ARGS__);                             \
      case 3 * 4 + 2: return FN<3, 2>(__VA_ARGS__);                             \
      case 3 * 4 + 3: return FN<3, 3>(__VA_ARGS__);                             \
      default: return fail(BN_EUNSUPPORTED, "unsupported kernel combination");  \
    }                                                                           \
  }()

int check_plan(const bn_plan* plan) {
  if (plan == nullptr) return fail(BN_EINVAL, "plan is NULL");
  return BN_OK;
}

}  // namespace

namespace bn {
int set_error(int code, const std::string& msg) { return fail(code, msg); }

double* site_partials(const bn_plan* plan, const bn_kernel_desc* kern, void* workspace, int which) {
  if (plan == nullptr || workspace == nullptr || check_desc(kern) != BN_OK) return nullptr;
  WsLayout w = ws_layout(plan->host, state_dim(kern));
  return (double*)((char*)workspace + (which == 0 ? w.sp1 : w.sp2));
}
}  // namespace bn

// ---- per-kernel timing (measurement only; not thread-safe) -----------------
namespace bn {
namespace {
const char* const kKernelNames[KID_COUNT] = {
    "k_filter_reduce", "k_fscan_tile", "k_fscan_series", "k_filter_pass", "k_smoother_reduce",
    "k_sscan_tile", "k_sscan_series", "k_smoother_pass", "k_series_sum", "k_st_conditional",
    "k_st_sites", "k_st_filter", "k_st_smoother", "k_st_sum", "k_smoother_mbf", "k_mf_pseudo", "k_mf_filter",
    "k_mf_sdiag", "k_mf_blocks", "k_mf_smoother", "k_ih_tied", "k_ih_dare", "k_ih_sinv", "k_ih_filter",
    "k_ih_smoother"};
struct TimingRec { int32_t kid; hipEvent_t a, b; };
bool g_timing = false;
std::vector<TimingRec> g_recs;
std::vector<hipEvent_t> g_pool;

hipEvent_t pool_event() {
  if (!g_pool.empty()) { hipEvent_t e = g_pool.back(); g_pool.pop_back(); return e; }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}
}  // namespace

thread_local bool g_timing_suppressed = false;  // set while a no-allocation entry point runs

TimedLaunch::TimedLaunch(int32_t k, hipStream_t s) : kid(k), st(s) {
  if (g_timing && !g_timing_suppressed && (a = pool_event()) != nullptr) (void)hipEventRecord(a, st);
}

TimedLaunch::~TimedLaunch() {
  if (a == nullptr) return;
  hipEvent_t b = pool_event();
  if (b == nullptr) { g_pool.push_back(a); return; }
  (void)hipEventRecord(b, st);
  g_recs.push_back(TimingRec{kid, a, b});
}
}  // namespace bn

static_assert(ST_INNOV_NONPD == BN_STATUS_INNOVATION_NONPD && ST_NONFINITE == BN_STATUS_NONFINITE &&
                  ST_SMOOTHER_NONPD == BN_STATUS_SMOOTHER_NONPD,
              "lane status bits match the C ABI's");

// ---------------------------------------------------------------------------
// Scratch of the no-allocation fused pass (bn_filter_smoother_noalloc): the
// workspace that bn_filter_smoother takes from its caller lives here instead, in
// a file-scope device array of BN_NOALLOC_SCRATCH_BYTES that every GPU the
// library runs on holds until the process ends.  In host code `scratch` names a
// host-side stand-in, not device memory: kernels get the address that
// hipGetSymbolAddress returns for the current device, looked up once per device.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[BN_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (NULL and an error message on failure).
void* scratch_ptr() {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) { fail(BN_EHIP, std::string("hipGetDevice: ") + hipGetErrorString(e)); return nullptr; }
  if (dev < 0 || dev >= kMaxDevices) { fail(BN_EUNSUPPORTED, "device ordinal beyond 64"); return nullptr; }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      fail(BN_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}
}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* bn_last_error(void) { return g_err.c_str(); }

int bn_version(void) { return 1; }

int bn_timing_enable(int32_t enable) {
  for (auto& r : g_recs) { g_pool.push_back(r.a); g_pool.push_back(r.b); }
  g_recs.clear();
  g_timing = enable != 0;
  return BN_OK;
}

int bn_timing_read(int32_t kernel_id, double* total_ms, int32_t* count) {
  if (kernel_id < 0 || kernel_id >= KID_COUNT || !total_ms || !count) return fail(BN_EINVAL, "bad timing query");
  double tot = 0.0;
  int32_t n = 0;
  for (auto& r : g_recs) {
    if (r.kid != kernel_id) continue;
    BN_HIP_CHECK(hipEventSynchronize(r.b));
    float ms = 0.f;
    BN_HIP_CHECK(hipEventElapsedTime(&ms, r.a, r.b));
    tot += ms;
    ++n;
  }
  *total_ms = tot;
  *count = n;
  return BN_OK;
}

const char* bn_kernel_name(int32_t kernel_id) {
  return (kernel_id >= 0 && kernel_id < KID_COUNT) ? kKernelNames[kernel_id] : "";
}

int32_t bn_num_kernel_ids(void) { return KID_COUNT; }

int bn_plan_create(const int64_t* offsets_host, int32_t n_series, int32_t chunk_len, bn_plan** out) {
  if (offsets_host == nullptr || out == nullptr) return fail(BN_EINVAL, "NULL argument to bn_plan_create");
  *out = nullptr;
  bn_plan* p = new bn_plan();
  std::string err;
  if (!build_plan(offsets_host, n_series, chunk_len, p->host, err)) {
    delete p;
    return fail(BN_EINVAL, err);
  }
  const PlanHost& h = p->host;
  // one device allocation for every table, 256-byte aligned pieces
  struct Piece { const void* src; size_t bytes; size_t off; };
  Piece pcs[16] = {
      {h.chunk_series.data(), h.chunk_series.size() * 4, 0}, {h.chunk_len_.data(), h.chunk_len_.size() * 4, 0},
      {h.chunk_flags.data(), h.chunk_flags.size(), 0},       {h.chunk_tile.data(), h.chunk_tile.size() * 4, 0},
      {h.chunk_start.data(), h.chunk_start.size() * 8, 0},   {h.tile_first.data(), h.tile_first.size() * 4, 0},
      {h.tile_count.data(), h.tile_count.size() * 4, 0},     {h.tile_series.data(), h.tile_series.size() * 4, 0},
      {h.tile_flags.data(), h.tile_flags.size(), 0},         {h.series_first_tile.data(), h.series_first_tile.size() * 4, 0},
      {h.series_ntiles.data(), h.series_ntiles.size() * 4, 0}, {h.series_first_chunk.data(), h.series_first_chunk.size() * 4, 0},
      {h.series_nchunks.data(), h.series_nchunks.size() * 4, 0}, {h.group_tiles.data(), h.group_tiles.size() * 4, 0},
      {h.group_w.data(), h.group_w.size(), 0},               {h.wave_maxlen.data(), h.wave_maxlen.size() * 4, 0}};
  size_t total = 0;
  for (auto& pc : pcs) {
    pc.off = total;
    total += (pc.bytes + 255) & ~size_t(255);
  }
  if (total == 0) total = 256;
  hipError_t e = hipMalloc(&p->dev, total);
  if (e != hipSuccess) {
    delete p;
    return fail(BN_ENOMEM, std::string("hipMalloc(plan): ") + hipGetErrorString(e));
  }
  std::vector<char> staging(total, 0);
  for (auto& pc : pcs)
    if (pc.bytes) std::memcpy(staging.data() + pc.off, pc.src, pc.bytes);
  e = hipMemcpy(p->dev, staging.data(), total, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(p->dev);
    delete p;
    return fail(BN_EHIP, std::string("hipMemcpy(plan): ") + hipGetErrorString(e));
  }
  char* d = (char*)p->dev;
  PlanView& v = p->view;
  v.n_series = h.n_series;
  v.n_chunks = h.n_chunks;
  v.n_chunks_pad = h.n_chunks_pad;
  v.n_tiles = h.n_tiles;
  v.chunk_len = h.chunk_len;
  v.gshift = h.gshift;
  v.n_slots = h.n_slots;
  v.chunk_series = (const int32_t*)(d + pcs[0].off);
  v.chunk_len_ = (const int32_t*)(d + pcs[1].off);
  v.chunk_flags = (const uint8_t*)(d + pcs[2].off);
  v.chunk_tile = (const int32_t*)(d + pcs[3].off);
  v.chunk_start = (const int64_t*)(d + pcs[4].off);
  v.tile_first = (const int32_t*)(d + pcs[5].off);
  v.tile_count = (const int32_t*)(d + pcs[6].off);
  v.tile_series = (const int32_t*)(d + pcs[7].off);
  v.tile_flags = (const uint8_t*)(d + pcs[8].off);
  v.series_first_tile = (const int32_t*)(d + pcs[9].off);
  v.series_ntiles = (const int32_t*)(d + pcs[10].off);
  v.series_first_chunk = (const int32_t*)(d + pcs[11].off);
  v.series_nchunks = (const int32_t*)(d + pcs[12].off);
  v.n_groups = (int32_t)h.group_w.size();
  v.group_tiles = (const int32_t*)(d + pcs[13].off);
  v.group_w = (const uint8_t*)(d + pcs[14].off);
  v.wave_maxlen = (const int32_t*)(d + pcs[15].off);
  *out = p;
  return BN_OK;
}

int bn_plan_destroy(bn_plan* plan) {
  if (plan == nullptr) return BN_OK;
  if (plan->dev) (void)hipFree(plan->dev);
  delete plan;
  return BN_OK;
}

int64_t bn_plan_num_slots(const bn_plan* plan) { return plan ? plan->host.n_slots : -1; }
int32_t bn_plan_num_chunks(const bn_plan* plan) { return plan ? plan->host.n_chunks : -1; }
int32_t bn_plan_num_tiles(const bn_plan* plan) { return plan ? plan->host.n_tiles : -1; }

int64_t bn_workspace_bytes(const bn_plan* plan, const bn_kernel_desc* kern) {
  if (plan == nullptr || check_desc(kern) != BN_OK) return -1;
  return ws_layout(plan->host, state_dim(kern)).total;
}

int bn_plan_gather_f32(const bn_plan* plan, const float* src, float* dst, int32_t width, float pad, void* stream) {
  if (check_plan(plan)) return BN_EINVAL;
  if (width < 1) return fail(BN_EINVAL, "width must be >= 1");
  const int64_t n = plan->host.n_slots;
  hipLaunchKernelGGL(k_gather_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, plan->view,
                     src, dst, width, pad);
  BN_HIP_CHECK(hipGetLastError());
  return BN_OK;
}

int bn_plan_scatter_f32(const bn_plan* plan, const float* src, float* dst, int32_t width, void* stream) {
  if (check_plan(plan)) return BN_EINVAL;
  if (width < 1) return fail(BN_EINVAL, "width must be >= 1");
  const int64_t n = plan->host.n_slots;
  hipLaunchKernelGGL(k_scatter_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, plan->view,
                     src, dst, width);
  BN_HIP_CHECK(hipGetLastError());
  return BN_OK;
}

int bn_plan_gather_f64(const bn_plan* plan, const double* src, double* dst, int32_t width, double pad, void* stream) {
  if (check_plan(plan)) return BN_EINVAL;
  if (width < 1) return fail(BN_EINVAL, "width must be >= 1");
  const int64_t n = plan->host.n_slots;
  hipLaunchKernelGGL(k_gather_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, plan->view,
                     src, dst, width, pad);
  BN_HIP_CHECK(hipGetLastError());
  return BN_OK;
}

int bn_plan_scatter_f64(const bn_plan* plan, const double* src, double* dst, int32_t width, void* stream) {
  if (check_plan(plan)) return BN_EINVAL;
  if (width < 1) return fail(BN_EINVAL, "width must be >= 1");
  const int64_t n = plan->host.n_slots;
  hipLaunchKernelGGL(k_scatter_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, plan->view,
                     src, dst, width);
  BN_HIP_CHECK(hipGetLastError());
  return BN_OK;
}

int bn_plan_gather_u8(const bn_plan* plan, const uint8_t* src, uint8_t* dst, uint8_t pad, void* stream) {
  if (check_plan(plan)) return BN_EINVAL;
  const int64_t n = plan->host.n_slots;
  hipLaunchKernelGGL(k_gather_u8, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, plan->view,
                     src, dst, pad);
  BN_HIP_CHECK(hipGetLastError());
  return BN_OK;
}

int bn_pairs_filter(const bn_kernel_desc* kern, const double* lengthscale, const double* variance,
                    int32_t hyper_stride, int32_t n_series, const int64_t* offsets, const double* dt, const double* y,
                    const double* noise, double* pair_means, double* pair_covs, double* loglik, void* stream) {
  if (int rc = check_desc(kern)) return rc;
  if (n_series < 1 || !offsets || !dt || !y || !noise || !pair_means || !pair_covs || !loglik || !lengthscale ||
      !variance)
    return fail(BN_EINVAL, "bad argument to bn_pairs_filter");
  HyperView hv{lengthscale, variance, hyper_stride};
  return BN_DISPATCH(kern, run_pairs_filter, hv, n_series, offsets, dt, y, noise, pair_means, pair_covs, loglik,
                     (hipStream_t)stream);
}

int bn_discretise(const bn_kernel_desc* kern, const double* lengthscale, const double* variance, int64_t n,
                      const float* dt, float* A, float* Q, void* stream) {
  if (int rc = check_desc(kern)) return rc;
  HyperView hv{lengthscale, variance, 0};
  return BN_DISPATCH(kern, run_discretise, hv, n, dt, A, Q, (hipStream_t)stream);
}

int bn_filter_smoother(const bn_plan* plan, const bn_kernel_desc* kern, const double* lengthscale,
                           const double* variance, int32_t hyper_stride, const float* dt, const float* y,
                           const float* r, const uint8_t* mask, float* post_mean, float* post_var, double* loglik,
                           int32_t* status, void* workspace, void* stream) {
  if (check_plan(plan)) return BN_EINVAL;
  if (int rc = check_desc(kern)) return rc;
  if (!dt || !y || !r || !post_mean || !post_var || !loglik || !workspace || !lengthscale || !variance)
    return fail(BN_EINVAL, "NULL array argument to bn_filter_smoother");
  if (hyper_stride != 0 && hyper_stride != kern->n_comp) return fail(BN_EINVAL, "hyper_stride must be 0 or n_comp");
  HyperView hv{lengthscale, variance, hyper_stride};
  return BN_DISPATCH(kern, run_filter_smoother, plan, hv, dt, y, r, mask, post_mean, post_var, loglik, status,
                     workspace, (hipStream_t)stream);
}

int bn_filter_smoother_noalloc(const bn_plan* plan, const bn_kernel_desc* kern, const double* lengthscale,
                               const double* variance, int32_t hyper_stride, const float* dt, const float* y,
                               const float* r, const uint8_t* mask, float* post_mean, float* post_var,
                               double* loglik, int32_t* status, void* stream) {
  if (check_plan(plan)) return BN_EINVAL;
  if (int rc = check_desc(kern)) return rc;
  if (!dt || !y || !r || !post_mean || !post_var || !loglik || !lengthscale || !variance)
    return fail(BN_EINVAL, "NULL array argument to bn_filter_smoother_noalloc");
  if (hyper_stride != 0 && hyper_stride != kern->n_comp) return fail(BN_EINVAL, "hyper_stride must be 0 or n_comp");
  const int64_t need = ws_layout(plan->host, state_dim(kern)).total;
  if (need > (int64_t)BN_NOALLOC_SCRATCH_BYTES)
    return fail(BN_EINVAL, "bn_filter_smoother_noalloc: this plan needs " + std::to_string(need) +
                               " bytes of scratch, more than the " + std::to_string((int64_t)BN_NOALLOC_SCRATCH_BYTES) +
                               " held by the library; use bn_filter_smoother with a workspace");
  void* ws = scratch_ptr();
  if (ws == nullptr) return BN_EHIP;
  HyperView hv{lengthscale, variance, hyper_stride};
  g_timing_suppressed = true;
  const int rc = BN_DISPATCH(kern, run_filter_smoother, plan, hv, dt, y, r, mask, post_mean, post_var, loglik,
                             status, ws, (hipStream_t)stream);
  g_timing_suppressed = false;
  return rc;
}

int bn_kalman_filter(const bn_plan* plan, const bn_kernel_desc* kern, const double* lengthscale,
                         const double* variance, int32_t hyper_stride, const float* dt, const float* y,
                         const float* r, const uint8_t* mask, double* means, double* covs, double* loglik,
                         uint32_t flags, int32_t* status, void* workspace, void* stream) {
  if (check_plan(plan)) return BN_EINVAL;
  if (int rc = check_desc(kern)) return rc;
  if (!dt || !y || !r || !means || !covs || !loglik || !workspace || !lengthscale || !variance)
    return fail(BN_EINVAL, "NULL array argument to bn_kalman_filter");
  if (hyper_stride != 0 && hyper_stride != kern->n_comp) return fail(BN_EINVAL, "hyper_stride must be 0 or n_comp");
  HyperView hv{lengthscale, variance, hyper_stride};
  return BN_DISPATCH(kern, run_kalman_filter, plan, hv, dt, y, r, mask, means, covs, loglik, flags, status,
                     workspace, (hipStream_t)stream);
}

int bn_filter_loglik(const bn_plan* plan, const bn_kernel_desc* kern, const double* lengthscale,
                     const double* variance, int32_t hyper_stride, const float* dt, const float* y, const float* r,
                     const uint8_t* mask, double* loglik, void* workspace, void* stream) {
  if (check_plan(plan)) return BN_EINVAL;
  if (int rc = check_desc(kern)) return rc;
  if (!dt || !y || !r || !loglik || !workspace || !lengthscale || !variance)
    return fail(BN_EINVAL, "NULL array argument to bn_filter_loglik");
  if (hyper_stride != 0 && hyper_stride != kern->n_comp) return fail(BN_EINVAL, "hyper_stride must be 0 or n_comp");
  HyperView hv{lengthscale, variance, hyper_stride};
  return BN_DISPATCH(kern, run_kalman_filter, plan, hv, dt, y, r, mask, (double*)nullptr, (double*)nullptr, loglik,
                     0u, (int32_t*)nullptr, workspace, (hipStream_t)stream);
}

int64_t bn_grad_workspace_bytes(const bn_plan* plan, const bn_kernel_desc* kern) {
  if (plan == nullptr || check_desc(kern) != BN_OK) return -1;
  const int D = state_dim(kern);
  const int64_t NZ = D + D * (D + 1) / 2, NH = 2 * kern->n_comp;
  return align256((NZ * NZ + NH * NZ + NZ + NH) * plan->host.n_chunks_pad * 8);
}

int bn_filter_loglik_grad(const bn_plan* plan, const bn_kernel_desc* kern, const double* lengthscale,
                          const double* variance, int32_t hyper_stride, const float* dt, const float* y,
                          const float* r, const uint8_t* mask, double* loglik, double* grad, void* workspace,
                          void* grad_workspace, void* stream) {
  if (check_plan(plan)) return BN_EINVAL;
  if (int rc = check_desc(kern)) return rc;
  if (!dt || !y || !r || !loglik || !grad || !workspace || !grad_workspace || !lengthscale || !variance)
    return fail(BN_EINVAL, "NULL array argument to bn_filter_loglik_grad");
  if (hyper_stride != 0 && hyper_stride != kern->n_comp) return fail(BN_EINVAL, "hyper_stride must be 0 or n_comp");
  HyperView hv{lengthscale, variance, hyper_stride};
  return BN_DISPATCH(kern, run_loglik_grad, plan, hv, dt, y, r, mask, loglik, grad, workspace, grad_workspace,
                     (hipStream_t)stream);
}

int bn_rts_smoother(const bn_plan* plan, const bn_kernel_desc* kern, const double* lengthscale,
                        const double* variance, int32_t hyper_stride, const float* dt, const double* filter_mean,
                        const double* filter_cov, double* means, double* covs, double* gains, uint32_t flags,
                        void* workspace, void* stream) {
  if (check_plan(plan)) return BN_EINVAL;
  if (int rc = check_desc(kern)) return rc;
  if (!dt || !filter_mean || !filter_cov || !means || !covs || !workspace || !lengthscale || !variance)
    return fail(BN_EINVAL, "NULL array argument to bn_rts_smoother");
  if (hyper_stride != 0 && hyper_stride != kern->n_comp) return fail(BN_EINVAL, "hyper_stride must be 0 or n_comp");
  HyperView hv{lengthscale, variance, hyper_stride};
  return BN_DISPATCH(kern, run_rts_smoother, plan, hv, dt, filter_mean, filter_cov, means, covs, gains, flags,
                     workspace, (hipStream_t)stream);
}

int bn_predict(const bn_plan* plan, const bn_kernel_desc* kern, const double* lengthscale, const double* variance,
               int32_t hyper_stride, const int64_t* train_offsets, const double* x_train, const double* means,
               const double* covs, const double* gains, const int64_t* test_offsets, const double* x_test,
               int64_t n_test, double* out_mean, double* out_var, void* stream) {
  if (check_plan(plan)) return BN_EINVAL;
  if (int rc = check_desc(kern)) return rc;
  if (n_test < 0) return fail(BN_EINVAL, "n_test must be >= 0");
  if (!lengthscale || !variance || !train_offsets || !x_train || !means || !covs || !gains || !test_offsets ||
      (n_test > 0 && (!x_test || !out_mean || !out_var)))
    return fail(BN_EINVAL, "NULL array argument to bn_predict");
  if (hyper_stride != 0 && hyper_stride != kern->n_comp) return fail(BN_EINVAL, "hyper_stride must be 0 or n_comp");
  HyperView hv{lengthscale, variance, hyper_stride};
  PredictIO io{train_offsets, x_train, means, covs, gains, test_offsets, x_test, n_test, out_mean, out_var};
  return BN_DISPATCH(kern, run_predict, plan, hv, io, (hipStream_t)stream);
}

}  // extern "C"
