// Host build of the HIP library's per-lane bodies (bayesnewton_amd/csrc/bn_lanes.h,
// bn_core.h) and plan construction (bn_plan_host.h).  It runs the fused filter +
// smoother pass of bn_filter_smoother with plain loops over the lanes and sequential
// scans in place of the LDS tile / series scans (the combine operators are associative,
// so only the rounding order differs), and checks:
//   * the posterior means / variances and per-series log-likelihoods against the
//     oracle's FP64 C filter + smoother (oracle/kfrts.c) on ragged batches (empty
//     series, one-step series, chunk lengths from 1 to the sequential filter, masks,
//     three kernels);
//   * the per-series status bits on forced failures: a negative noise variance (a
//     negative site precision), NaN observations, a non-PD smoother, and that the
//     other series of the batch keep status 0 and their oracle values.
// tests/test_cpu_lanes.py builds it with -fsanitize=address,undefined and runs it.
// Prints one line per case; exit status 0 when every check holds.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "../../bayesnewton_amd/csrc/bn_lanes.h"
#include "../../bayesnewton_amd/csrc/bn_plan_host.h"

extern "C" int oracle_filter_smoother(int n_comp, const int* kinds, const double* ls, const double* var,
                                      int hyper_stride, const int64_t* offsets, int B, const double* dt,
                                      const double* y, const double* r, const uint8_t* mask, double* post_mean,
                                      double* post_var, double* loglik, int nthreads);

using namespace bn;

namespace {

int g_fail = 0;

void check(bool ok, const std::string& what) {
  if (!ok) {
    ++g_fail;
    std::printf("FAIL %s\n", what.c_str());
  }
}

PlanView view_of(const PlanHost& p) {
  PlanView v;
  v.n_series = p.n_series;
  v.n_chunks = p.n_chunks;
  v.n_chunks_pad = p.n_chunks_pad;
  v.n_tiles = p.n_tiles;
  v.chunk_len = p.chunk_len;
  v.gshift = p.gshift;
  v.n_slots = p.n_slots;
  v.chunk_series = p.chunk_series.data();
  v.chunk_len_ = p.chunk_len_.data();
  v.chunk_flags = p.chunk_flags.data();
  v.chunk_tile = p.chunk_tile.data();
  v.chunk_start = p.chunk_start.data();
  v.tile_first = p.tile_first.data();
  v.tile_count = p.tile_count.data();
  v.tile_series = p.tile_series.data();
  v.tile_flags = p.tile_flags.data();
  v.series_first_tile = p.series_first_tile.data();
  v.series_ntiles = p.series_ntiles.data();
  v.series_first_chunk = p.series_first_chunk.data();
  v.series_nchunks = p.series_nchunks.data();
  v.n_groups = (int32_t)p.group_w.size();
  v.group_tiles = p.group_tiles.data();
  v.group_w = p.group_w.data();
  v.wave_maxlen = p.wave_maxlen.data();
  return v;
}

template <class T>
std::vector<T> gather(const PlanHost& p, const std::vector<T>& src, T pad) {
  PlanView v = view_of(p);
  std::vector<T> out((size_t)p.n_slots, pad);
  for (int32_t j = 0; j < p.n_chunks; ++j)
    for (int32_t k = 0; k < p.chunk_len_[j]; ++k) out[(size_t)v.slot(j, k)] = src[(size_t)(p.chunk_start[j] + k)];
  return out;
}

template <class T>
std::vector<T> scatter(const PlanHost& p, const std::vector<T>& src) {
  PlanView v = view_of(p);
  std::vector<T> out((size_t)p.n_steps);
  for (int32_t j = 0; j < p.n_chunks; ++j)
    for (int32_t k = 0; k < p.chunk_len_[j]; ++k) out[(size_t)(p.chunk_start[j] + k)] = src[(size_t)v.slot(j, k)];
  return out;
}

template <class E>
E load_e(const std::vector<agg_t>& a, int64_t stride, int64_t i) {
  E e;
  soa_load(e, CSoA{a.data(), stride}, i);
  return e;
}
template <class E>
void store_e(const E& e, std::vector<agg_t>& a, int64_t stride, int64_t i) {
  soa_store(e, SoA{a.data(), stride}, i);
}

// The fused pass of run_filter_smoother (bn_markov.hip) with host loops.
template <int K1, int K2>
void host_pass(const PlanHost& p, const std::vector<double>& ls, const std::vector<double>& var, int32_t stride,
               const std::vector<float>& dt, const std::vector<float>& y, const std::vector<float>& r,
               const std::vector<uint8_t>& mask, std::vector<float>& pm, std::vector<float>& pv,
               std::vector<double>& ll, std::vector<int32_t>& status) {
  constexpr int D = K1 + K2;
  constexpr int NS = D * (D + 1) / 2;
  using FE = FElem<D, double>;
  using SE = SElem<D, double>;
  const PlanView pl = view_of(p);
  const int64_t nc = p.n_chunks_pad, nt = p.n_tiles > 0 ? p.n_tiles : 1, ns = p.n_slots;
  const HyperView hv{ls.data(), var.data(), stride};
  std::vector<agg_t> fagg((size_t)(FE::NF * nc)), ftagg((size_t)(FE::NF * nt)), ftcarry((size_t)(FE::NF * nt));
  std::vector<agg_t> sagg((size_t)(SE::NF * nc)), stagg((size_t)(SE::NF * nt)), stcarry((size_t)(SE::NF * nt));
  std::vector<state_t> fbuf((size_t)((D + NS + D + 1) * ns));
  std::vector<double> ell((size_t)nc);
  std::vector<int32_t> chst((size_t)nc, -1);
  auto single = [&](int32_t j) { return (p.chunk_flags[j] & (CH_FIRST | CH_LAST)) == (CH_FIRST | CH_LAST); };
  auto single_tile = [&](int32_t t) {
    return (p.tile_flags[t] & (TL_FIRST | TL_LAST)) == (TL_FIRST | TL_LAST) && p.tile_count[t] == 1;
  };
  if (p.needs_scan) {
    for (int32_t j = 0; j < p.n_chunks; ++j)
      if (!single(j)) lane_filter_reduce<K1, K2, double>(pl, hv, dt.data(), y.data(), r.data(), SoA{fagg.data(), nc}, j);
    for (int32_t t = 0; t < p.n_tiles; ++t) {  // tile-local inclusive prefixes
      if (single_tile(t)) continue;
      const int32_t f = p.tile_first[t], c = p.tile_count[t];
      for (int32_t i = 1; i < c; ++i) {
        FE out;
        fcombine<D, double>(load_e<FE>(fagg, nc, f + i - 1), load_e<FE>(fagg, nc, f + i), out);
        store_e(out, fagg, nc, f + i);
      }
      store_e(load_e<FE>(fagg, nc, f + c - 1), ftagg, nt, t);
    }
    for (int32_t b = 0; b < p.n_series; ++b) {  // exclusive tile carries
      const int32_t n = p.series_ntiles[b], ft = p.series_first_tile[b];
      if (n <= 1) continue;
      FE run = load_e<FE>(ftagg, nt, ft);
      for (int32_t t = 1; t < n; ++t) {
        store_e(run, ftcarry, nt, ft + t);
        FE out;
        fcombine<D, double>(run, load_e<FE>(ftagg, nt, ft + t), out);
        run = out;
      }
    }
  }
  FilterPassOut fo;
  fo.state = StateIO{fbuf.data(), fbuf.data() + (int64_t)D * ns, true};
  fo.predict = true;
  fo.hrow = reinterpret_cast<hrow_t*>(fbuf.data() + (int64_t)(D + NS) * ns);
  fo.sagg = SoA{sagg.data(), nc};
  fo.ell = ell.data();
  fo.status = chst.data();
  for (int32_t j = 0; j < p.n_chunks; ++j)
    lane_filter_pass<K1, K2, double, true>(pl, hv, dt.data(), y.data(), r.data(), mask.data(), CSoA{fagg.data(), nc},
                                     CSoA{ftcarry.data(), nt}, fo, j);
  if (p.needs_scan) {
    for (int32_t t = 0; t < p.n_tiles; ++t) {  // tile-local inclusive suffixes
      if (single_tile(t)) continue;
      const int32_t f = p.tile_first[t], c = p.tile_count[t];
      for (int32_t i = c - 2; i >= 0; --i) {
        SE out;
        scombine<D, double>(load_e<SE>(sagg, nc, f + i), load_e<SE>(sagg, nc, f + i + 1), out);
        store_e(out, sagg, nc, f + i);
      }
      store_e(load_e<SE>(sagg, nc, f), stagg, nt, t);
    }
    for (int32_t b = 0; b < p.n_series; ++b) {  // exclusive carries from the end
      const int32_t n = p.series_ntiles[b], ft = p.series_first_tile[b];
      if (n <= 1) continue;
      SE run = load_e<SE>(stagg, nt, ft + n - 1);
      for (int32_t t = n - 2; t >= 0; --t) {
        store_e(run, stcarry, nt, ft + t);
        SE out;
        scombine<D, double>(load_e<SE>(stagg, nt, ft + t), run, out);
        run = out;
      }
    }
  }
  pm.assign((size_t)ns, 0.f);
  pv.assign((size_t)ns, 0.f);
  for (int32_t j = 0; j < p.n_chunks; ++j)
    lane_smoother_mbf<K1, K2, double>(pl, hv, dt.data(), y.data(), r.data(), CSoA{sagg.data(), nc},
                                      CSoA{stcarry.data(), nt}, CStateIO{fbuf.data(), fbuf.data() + (int64_t)D * ns, true},
                                      fo.hrow, pm.data(), pv.data(), chst.data(), j);
  ll.assign((size_t)p.n_series, 0.0);
  status.assign((size_t)p.n_series, 0);
  for (int32_t b = 0; b < p.n_series; ++b)
    for (int32_t i = 0; i < p.series_nchunks[b]; ++i) {
      const int32_t j = p.series_first_chunk[b] + i;
      ll[b] += ell[j];
      status[b] |= chst[j];
    }
}

struct Case {
  std::string name;
  int k1, k2;
  std::vector<int64_t> lens;
  int32_t chunk_len;
  uint64_t seed;
};

struct Data {
  std::vector<int64_t> off;
  std::vector<double> ls, var, dt, y, r;
  std::vector<uint8_t> mask;
};

Data make_data(const Case& c) {
  std::mt19937_64 rng(c.seed);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  std::normal_distribution<double> Nn(0.0, 1.0);
  Data d;
  d.off.push_back(0);
  for (int64_t n : c.lens) d.off.push_back(d.off.back() + n);
  const int B = (int)c.lens.size(), nc = c.k2 > 0 ? 2 : 1;
  for (int b = 0; b < B; ++b)
    for (int q = 0; q < nc; ++q) {
      d.ls.push_back(0.5 + 4.0 * U(rng));
      d.var.push_back(0.3 + U(rng));
    }
  for (int b = 0; b < B; ++b) {
    double x = 0.0;
    for (int64_t k = 0; k < c.lens[b]; ++k) {
      const double step = k == 0 ? 0.0 : 0.002 + 0.2 * U(rng);
      x += step;
      // FP32 inputs, as the library takes them; the oracle sees the same rounded values
      d.dt.push_back((double)(float)step);
      d.y.push_back((double)(float)(std::sin(0.8 * x) + 0.3 * Nn(rng)));
      d.r.push_back((double)(float)(0.01 + 0.3 * U(rng)));
      d.mask.push_back(U(rng) < 0.1 ? 1 : 0);
    }
  }
  return d;
}

template <int K1, int K2>
bool run_host(const Case& c, const Data& d, std::vector<float>& pm, std::vector<float>& pv,
              std::vector<double>& ll, std::vector<int32_t>& st) {
  PlanHost p;
  std::string err;
  if (!build_plan(d.off.data(), (int32_t)c.lens.size(), c.chunk_len, p, err)) {
    std::printf("plan error: %s\n", err.c_str());
    return false;
  }
  std::vector<float> dt(d.dt.begin(), d.dt.end()), y(d.y.begin(), d.y.end()), r(d.r.begin(), d.r.end());
  auto dtt = gather(p, dt, 0.f), yt = gather(p, y, 0.f), rt = gather(p, r, 1.f);
  auto mt = gather(p, d.mask, (uint8_t)1);
  std::vector<float> pmt, pvt;
  host_pass<K1, K2>(p, d.ls, d.var, K2 > 0 ? 2 : 1, dtt, yt, rt, mt, pmt, pvt, ll, st);
  pm = scatter(p, pmt);
  pv = scatter(p, pvt);
  return true;
}

bool run_any(const Case& c, const Data& d, std::vector<float>& pm, std::vector<float>& pv, std::vector<double>& ll,
             std::vector<int32_t>& st) {
  if (c.k1 == 3 && c.k2 == 2) return run_host<3, 2>(c, d, pm, pv, ll, st);
  if (c.k1 == 2 && c.k2 == 0) return run_host<2, 0>(c, d, pm, pv, ll, st);
  if (c.k1 == 1 && c.k2 == 3) return run_host<1, 3>(c, d, pm, pv, ll, st);
  std::printf("unsupported kernel in the harness\n");
  return false;
}

void oracle(const Case& c, const Data& d, std::vector<double>& om, std::vector<double>& ov, std::vector<double>& oll) {
  const int kinds[2] = {c.k1, c.k2};
  const int nc = c.k2 > 0 ? 2 : 1;
  om.assign(d.dt.size(), 0.0);
  ov.assign(d.dt.size(), 0.0);
  oll.assign(c.lens.size(), 0.0);
  oracle_filter_smoother(nc, kinds, d.ls.data(), d.var.data(), nc, d.off.data(), (int)c.lens.size(), d.dt.data(),
                         d.y.data(), d.r.data(), d.mask.data(), om.data(), ov.data(), oll.data(), 1);
}

// max errors over the steps of the series where `use[b]`
void compare(const Case& c, const Data& d, const std::vector<float>& pm, const std::vector<float>& pv,
             const std::vector<double>& ll, const std::vector<bool>& use, double* em, double* ev, double* el) {
  std::vector<double> om, ov, oll;
  oracle(c, d, om, ov, oll);
  double mmax = 1e-300;
  for (size_t b = 0; b < use.size(); ++b)
    if (use[b])
      for (int64_t i = d.off[b]; i < d.off[b + 1]; ++i) mmax = std::fmax(mmax, std::fabs(om[i]));
  *em = *ev = *el = 0.0;
  for (size_t b = 0; b < use.size(); ++b) {
    if (!use[b]) continue;
    for (int64_t i = d.off[b]; i < d.off[b + 1]; ++i) {
      *em = std::fmax(*em, std::fabs(pm[i] - om[i]) / mmax);
      *ev = std::fmax(*ev, std::fabs(pv[i] - ov[i]) / std::fabs(ov[i]));
      if (!(std::isfinite((double)pm[i]) && std::isfinite((double)pv[i]))) *em = *ev = INFINITY;
    }
    if (c.lens[b] > 0) *el = std::fmax(*el, std::fabs(ll[b] - oll[b]) / std::fabs(oll[b]));
    else if (ll[b] != 0.0) *el = INFINITY;
  }
}

}  // namespace

int main() {
  const std::vector<Case> cases = {
      {"sum52_32_ragged_chunk16", 3, 2, {700, 0, 1, 2600, 33, 0, 129}, 16, 1},
      {"sum52_32_chunk1", 3, 2, {40, 3, 0, 17}, 1, 2},
      {"sum52_32_sequential", 3, 2, {300, 1, 0, 450}, 4096, 3},
      {"m32_many_tiles", 2, 0, {20000, 5, 900}, 2, 4},
      {"sum12_52_chunk7", 1, 3, {1000, 64, 65, 2}, 7, 5},
  };
  for (const Case& c : cases) {
    Data d = make_data(c);
    std::vector<float> pm, pv;
    std::vector<double> ll;
    std::vector<int32_t> st;
    if (!run_any(c, d, pm, pv, ll, st)) { ++g_fail; continue; }
    double em, ev, el;
    compare(c, d, pm, pv, ll, std::vector<bool>(c.lens.size(), true), &em, &ev, &el);
    int32_t st_or = 0;
    for (int32_t v : st) st_or |= v;
    std::printf("case %s: mean %.2e var %.2e loglik %.2e status %d\n", c.name.c_str(), em, ev, el, st_or);
    // FP32 outputs of FP64 recurrences against the FP64 oracle
    check(em < 1e-5 && ev < 1e-5 && el < 1e-10, c.name + " vs oracle");
    check(st_or == 0, c.name + " status of a healthy batch");
  }
  // forced failures in one batch: series 0 healthy, 1 a negative noise variance (a
  // negative site precision), 2 a NaN observation on an unmasked step, 3 healthy
  {
    Case c{"forced_failures", 3, 2, {500, 400, 300, 600}, 32, 9};
    Data d = make_data(c);
    const int64_t s1 = d.off[1] + 123, s2 = d.off[2] + 77;
    d.r[s1] = -5.0;
    d.mask[s1] = 0;
    d.y[s2] = NAN;
    d.mask[s2] = 0;
    std::vector<float> pm, pv;
    std::vector<double> ll;
    std::vector<int32_t> st;
    if (!run_any(c, d, pm, pv, ll, st)) { ++g_fail; }
    std::printf("case forced_failures: status %d %d %d %d, loglik %g %g %g %g\n", st[0], st[1], st[2], st[3], ll[0],
                ll[1], ll[2], ll[3]);
    check(st[0] == 0 && st[3] == 0, "healthy series keep status 0");
    check((st[1] & ST_INNOV_NONPD) != 0, "negative noise variance sets the innovation bit");
    check((st[2] & ST_NONFINITE) != 0, "NaN observation sets the non-finite bit");
    check(!std::isfinite(ll[2]), "NaN observation propagates into the log-likelihood (as in the reference)");
    double em, ev, el;
    compare(c, d, pm, pv, ll, {true, false, false, true}, &em, &ev, &el);
    std::printf("case forced_failures healthy series: mean %.2e var %.2e loglik %.2e\n", em, ev, el);
    check(em < 1e-5 && ev < 1e-5 && el < 1e-10, "healthy series of a failing batch vs oracle");
  }
  // a non-PD smoother: a huge negative noise variance late in a series makes the filter's
  // covariance indefinite; the smoothed variances it feeds are then not all positive
  {
    Case c{"smoother_nonpd", 2, 0, {200, 200}, 16, 11};
    Data d = make_data(c);
    for (int64_t k = 150; k < 200; ++k) d.r[d.off[0] + k] = -1e-3;
    std::vector<float> pm, pv;
    std::vector<double> ll;
    std::vector<int32_t> st;
    if (!run_any(c, d, pm, pv, ll, st)) { ++g_fail; }
    std::printf("case smoother_nonpd: status %d %d\n", st[0], st[1]);
    check(st[0] != 0 && st[1] == 0, "failing series flagged, the other not");
  }
  // plan construction (bn_plan_host.h): refusals, and every step of every series in
  // exactly one in-range slot with no slot shared
  {
    PlanHost p;
    std::string err;
    const int64_t bad1[3] = {0, 5, 3}, none[3] = {0, 0, 0}, bad0[2] = {1, 4};
    check(!build_plan(bad1, 2, 4, p, err), "decreasing offsets refused");
    check(!build_plan(none, 2, 4, p, err), "a batch without steps refused");
    check(!build_plan(bad0, 1, 4, p, err), "offsets[0] != 0 refused");
    const int64_t big[2] = {0, (int64_t)kTile * kTile * 3 + 1};
    check(!build_plan(big, 1, 3, p, err), "a series beyond kTile^2 chunks refused");
    check(!build_plan(big, 1, 0, p, err), "chunk_len 0 refused");
    const std::vector<int64_t> off = {0, 0, 1, 1, 130, 1000, 1000, 1077};
    for (int32_t cl : {1, 3, 64, 129, 5000}) {
      check(build_plan(off.data(), (int32_t)off.size() - 1, cl, p, err), "plan built");
      PlanView v = view_of(p);
      std::vector<int> seen((size_t)p.n_slots, 0);
      std::vector<int> step_seen((size_t)p.n_steps, 0);
      bool ok = true;
      for (int32_t j = 0; j < p.n_chunks; ++j)
        for (int32_t k = 0; k < p.chunk_len_[j]; ++k) {
          const int64_t s = v.slot(j, k), t = p.chunk_start[j] + k;
          ok = ok && s >= 0 && s < p.n_slots && t >= 0 && t < p.n_steps;
          if (ok) { ++seen[(size_t)s]; ++step_seen[(size_t)t]; }
        }
      for (int x : seen) ok = ok && x <= 1;
      for (int x : step_seen) ok = ok && x == 1;
      for (int32_t b = 0; b < p.n_series; ++b)
        ok = ok && (p.series_nchunks[b] == 0) == (off[b + 1] == off[b]);
      check(ok, "plan slots cover every step once (chunk_len " + std::to_string(cl) + ")");
      // scan groups: every tile that needs a scan in exactly one group, no other tile in
      // any, a tile of a group of width w holds at most kTile / w chunks
      bool gok = p.group_tiles.size() == (size_t)kGroupMax * p.group_w.size() && v.n_groups == (int32_t)p.group_w.size();
      std::vector<int> in_group((size_t)p.n_tiles, 0);
      for (size_t g = 0; gok && g < p.group_w.size(); ++g) {
        const int w = p.group_w[g];
        gok = gok && (w == 1 || w == 2 || w == 4 || w == 8);
        for (int h = 0; gok && h < kGroupMax; ++h) {
          const int32_t t = p.group_tiles[g * kGroupMax + h];
          if (t < 0) continue;
          gok = gok && h < w && t < p.n_tiles && p.tile_count[t] <= kTile / w;
          if (gok) ++in_group[(size_t)t];
        }
      }
      for (int32_t t = 0; gok && t < p.n_tiles; ++t) {
        const bool one = (p.tile_flags[t] & (TL_FIRST | TL_LAST)) == (TL_FIRST | TL_LAST) && p.tile_count[t] == 1;
        gok = in_group[(size_t)t] == (one ? 0 : 1);
      }
      check(gok, "scan groups cover every tile once (chunk_len " + std::to_string(cl) + ")");
      bool wok = p.wave_maxlen.size() == (size_t)(p.n_chunks + 63) / 64;
      for (int32_t j = 0; wok && j < p.n_chunks; ++j) {
        int32_t mx = 0;
        for (int32_t u = j & ~63; u < std::min(p.n_chunks, (j & ~63) + 64); ++u) mx = std::max(mx, p.chunk_len_[u]);
        wok = p.wave_maxlen[(size_t)(j >> 6)] == mx;
      }
      check(wok, "wave_maxlen is the longest chunk of each 64 (chunk_len " + std::to_string(cl) + ")");
    }
    // many short series pack eight tiles to a group, long ones one
    {
      std::vector<int64_t> o2(41, 0);
      for (int b = 1; b <= 40; ++b) o2[b] = o2[b - 1] + (b <= 36 ? 10 : 1000);
      check(build_plan(o2.data(), 40, 1, p, err), "plan built (mixed lengths)");
      int n8 = 0, n1 = 0;
      for (uint8_t w : p.group_w) { n8 += w == 8; n1 += w == 1; }
      std::printf("mixed lengths: %d tiles in %zu groups (%d of width 8, %d of width 1)\n", p.n_tiles, p.group_w.size(),
                  n8, n1);
      // 36 one-tile series of 10 chunks (width 8: 5 groups), 4 series of 8 tiles whose
      // last has 104 chunks (width 1)
      check(n8 == 5 && n1 == 32 && p.group_w.size() == 37, "short series share groups");
    }
  }
  // log_fast (bn_core.h) against the library log: at most 4 ulp on finite x > 0 (three
  // ranges: the full exponent range, the Student-t argument 1 + r^2 / nu, around 1, and a
  // subnormal), and log's values for zero, negatives, infinities and NaN
  {
    std::mt19937_64 g(7);
    double worst = 0.0;
    auto ulps = [](double a, double b) {
      const double sp = std::nextafter(std::fabs(b), INFINITY) - std::fabs(b);
      return b == 0.0 ? std::fabs(a) : std::fabs(a - b) / sp;
    };
    std::uniform_real_distribution<double> u1(-700.0, 700.0), u2(0.0, 40.0), u3(0.5, 2.0);
    for (int i = 0; i < 200000; ++i) {
      const double xs[3] = {std::exp(u1(g)), 1.0 + u2(g), u3(g)};
      for (double x : xs) worst = std::max(worst, ulps(log_fast(x), std::log(x)));
    }
    std::printf("log_fast: worst %.2f ulp\n", worst);
    worst = std::max(worst, ulps(log_fast(1e-310), std::log(1e-310)));  // subnormal
    bool same = log_fast(0.0) == -INFINITY && std::isnan(log_fast(-1.0)) && log_fast(INFINITY) == INFINITY &&
                std::isnan(log_fast(-INFINITY)) && std::isnan(log_fast(NAN)) && log_fast(1.0) == 0.0;
    check(worst <= 4.0 && same, "log_fast within 4 ulp of log, library values off the normal range");
  }
  // strided_sum (bn_core.h): every element of a thread's stride counted once, at lengths
  // around the 8-element rounds (integers, so every order is exact), and a[t] itself when
  // the thread holds one element
  {
    std::vector<double> a(5000);
    for (size_t i = 0; i < a.size(); ++i) a[i] = (double)((i * 7919) % 1009) - 500.0;
    bool ok = true;
    for (int32_t n : {0, 1, 255, 256, 257, 2047, 2048, 2049, 4999}) {
      for (int32_t t : {0, 1, 100, 255}) {
        double want = 0.0;
        for (int32_t i = t; i < n; i += 256) want += a[i];
        ok = ok && strided_sum(a.data(), n, t, 256) == want;
      }
    }
    ok = ok && strided_sum(a.data(), 200, 7, 256) == a[7];
    check(ok, "strided_sum counts each strided element once");
    // strided_or: one set bit per element, so a dropped or extra element shows
    std::vector<int32_t> fl(4000);
    for (size_t i = 0; i < fl.size(); ++i) fl[i] = (int32_t)(1u << (i / 256 % 31));
    bool oks = true;
    for (int32_t n : {0, 1, 300, 2047, 2048, 2049, 3999}) {
      for (int32_t t : {0, 5, 255}) {
        int32_t want = 0;
        for (int32_t i = t; i < n; i += 256) want |= fl[i];
        oks = oks && strided_or(fl.data(), n, t, 256) == want;
      }
    }
    check(oks, "strided_or covers each strided element");
  }
  std::printf(g_fail ? "LANES_HOST_FAILED %d\n" : "LANES_HOST_OK\n", g_fail);
  return g_fail ? 1 : 0;
}
