// Per-pair entry points behind the reference's per-pair Python API:
//   pair enumeration       (ld_pipeline.iter_pairs_within_distance)
//   counts for pair lists  (ld_pipeline.compute_counts_for_pairs /
//                           compute_genotype_counts_for_pairs)
//   per-pair estimators    (ld_statistics.dd / dz / pi2, the 2-pop and
//                           multi-pop batches) and r / r^2 / D'.
// These are not the binned hot path; they serve callers that want per-pair
// values.  Estimators divide by the falling factorials exactly as the
// reference does (numerator / denominator with a den > 0 guard).
#include "pgld_internal.h"
#include "stats_hap.inc"
#include "stats_geno.inc"

namespace pgld {

// ---------------------------------------------------------------------------
// Pair enumeration
// ---------------------------------------------------------------------------
__global__ void pair_bounds_kernel(const double* __restrict__ pos, int64_t V, double max_dist,
                                   int64_t* __restrict__ jmax) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= V) return;
  const double ub = pos[i] + max_dist;
  int64_t lo = i + 1, hi = V;  // positions are sorted: every j <= i has pos <= ub
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (pos[mid] <= ub) lo = mid + 1; else hi = mid;
  }
  jmax[i] = lo;
}

// offs[k] = number of pairs of anchors v_lo .. v_lo+k-1 (exclusive prefix, m+1 entries)
__global__ void emit_pairs_kernel(const int64_t* __restrict__ offs, int64_t v_lo, int64_t m,
                                  int64_t n_pairs, int32_t* __restrict__ idx_i,
                                  int32_t* __restrict__ idx_j) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_pairs) return;
  int64_t lo = 0, hi = m;  // largest k with offs[k] <= t
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (offs[mid] <= t) lo = mid; else hi = mid;
  }
  const int64_t i = v_lo + lo;
  idx_i[t] = (int32_t)i;
  idx_j[t] = (int32_t)(i + 1 + (t - offs[lo]));
}

// ---------------------------------------------------------------------------
// Counts for explicit pair lists, from the packed panels (dot8 on 4-bit values).
// ---------------------------------------------------------------------------
__device__ __forceinline__ const uint4* site_half(const int8_t* panels, int64_t block_bytes,
                                                  int64_t pop_off, int ks, int f, int s,
                                                  int64_t site, int half) {
  const int8_t* frag = panels + (site >> 5) * block_bytes + pop_off +
                       ((int64_t)f * ks + s) * PGLD_FRAG_BYTES;
  return reinterpret_cast<const uint4*>(frag + ((site & 31) + 32 * half) * 16);
}

// Dot product of 32 4-bit indicator values (E2M1 1.0 = 2, so each matching
// pair contributes 4; the caller divides the total by 4).
__device__ __forceinline__ int dot32(uint4 a, uint4 b, int acc) {
  acc = __builtin_amdgcn_udot8(a.x, b.x, acc, false);
  acc = __builtin_amdgcn_udot8(a.y, b.y, acc, false);
  acc = __builtin_amdgcn_udot8(a.z, b.z, acc, false);
  acc = __builtin_amdgcn_udot8(a.w, b.w, acc, false);
  return acc;
}

// Feature products P[fa][fb] = sum_k F_fa(i)[k] * F_fb(j)[k] for all feature pairs.
template <int NF>
__device__ __forceinline__ void feature_products(const int8_t* panels, int64_t block_bytes,
                                                 int64_t pop_off, int ks, int64_t i, int64_t j,
                                                 int (&prod)[NF][NF]) {
#pragma unroll
  for (int a = 0; a < NF; ++a)
#pragma unroll
    for (int b = 0; b < NF; ++b) prod[a][b] = 0;
  for (int s = 0; s < ks; ++s) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      uint4 fi[NF], fj[NF];
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        fi[f] = *site_half(panels, block_bytes, pop_off, ks, f, s, i, h);
        fj[f] = *site_half(panels, block_bytes, pop_off, ks, f, s, j, h);
      }
#pragma unroll
      for (int a = 0; a < NF; ++a)
#pragma unroll
        for (int b = 0; b < NF; ++b) prod[a][b] = dot32(fi[a], fj[b], prod[a][b]);
    }
  }
#pragma unroll
  for (int a = 0; a < NF; ++a)
#pragma unroll
    for (int b = 0; b < NF; ++b) prod[a][b] >>= 2;
}

// 0/1 nibble masks of 4-bit E2M1 feature values: where the value is 1.0
// (0x2), -1.0 (0xA), or either.
__device__ __forceinline__ uint32_t nib_pos(uint32_t w) { return (w >> 1) & ~(w >> 3) & 0x11111111u; }
__device__ __forceinline__ uint32_t nib_neg(uint32_t w) { return (w >> 1) & (w >> 3) & 0x11111111u; }
__device__ __forceinline__ uint32_t nib_one(uint32_t w) { return (w >> 1) & 0x11111111u; }
__device__ __forceinline__ uint4 map4(uint4 a, uint32_t (*f)(uint32_t)) {
  return make_uint4(f(a.x), f(a.y), f(a.z), f(a.w));
}
// Dot product of 32 0/1 nibbles.
__device__ __forceinline__ int dot32m(uint4 a, uint4 b, int acc) {
  acc = __builtin_amdgcn_udot8(a.x, b.x, acc, false);
  acc = __builtin_amdgcn_udot8(a.y, b.y, acc, false);
  acc = __builtin_amdgcn_udot8(a.z, b.z, acc, false);
  acc = __builtin_amdgcn_udot8(a.w, b.w, acc, false);
  return acc;
}

// The centred moments of genotype sites i and j from the packed features
// u (signed: the products split into the +1 and -1 parts), q and v.
__device__ __forceinline__ GenoMoments geno_pair_moments(const int8_t* panels, int64_t block_bytes,
                                                         int64_t pop_off, int ks, bool missing,
                                                         int64_t i, int64_t j) {
  int pp = 0, nn = 0, pn = 0, np_ = 0;          // u.u parts
  int pq = 0, nq = 0, qp = 0, qn = 0, qq = 0;   // u.q, q.u, q.q parts
  int pv = 0, nv = 0, qv = 0, vp = 0, vn = 0, vq = 0, vv = 0;
  for (int s = 0; s < ks; ++s) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const uint4 ui = *site_half(panels, block_bytes, pop_off, ks, 0, s, i, h);
      const uint4 uj = *site_half(panels, block_bytes, pop_off, ks, 0, s, j, h);
      const uint4 Qi = map4(*site_half(panels, block_bytes, pop_off, ks, 1, s, i, h), nib_one);
      const uint4 Qj = map4(*site_half(panels, block_bytes, pop_off, ks, 1, s, j, h), nib_one);
      const uint4 Pi = map4(ui, nib_pos), Ni = map4(ui, nib_neg);
      const uint4 Pj = map4(uj, nib_pos), Nj = map4(uj, nib_neg);
      pp = dot32m(Pi, Pj, pp); nn = dot32m(Ni, Nj, nn);
      pn = dot32m(Pi, Nj, pn); np_ = dot32m(Ni, Pj, np_);
      pq = dot32m(Pi, Qj, pq); nq = dot32m(Ni, Qj, nq);
      qp = dot32m(Qi, Pj, qp); qn = dot32m(Qi, Nj, qn);
      qq = dot32m(Qi, Qj, qq);
      if (missing) {
        const uint4 Vi = map4(*site_half(panels, block_bytes, pop_off, ks, 2, s, i, h), nib_one);
        const uint4 Vj = map4(*site_half(panels, block_bytes, pop_off, ks, 2, s, j, h), nib_one);
        pv = dot32m(Pi, Vj, pv); nv = dot32m(Ni, Vj, nv); qv = dot32m(Qi, Vj, qv);
        vp = dot32m(Vi, Pj, vp); vn = dot32m(Vi, Nj, vn); vq = dot32m(Vi, Qj, vq);
        vv = dot32m(Vi, Vj, vv);
      }
    }
  }
  GenoMoments m;
  m.ab11 = pp + nn - pn - np_;
  m.ab12 = pq - nq;
  m.ab21 = qp - qn;
  m.ab22 = qq;
  m.a1 = pv - nv; m.a2 = qv; m.b1 = vp - vn; m.b2 = vq; m.n0 = vv;
  return m;
}

struct CountArgs {
  const int8_t* panels;
  int64_t block_bytes, pop_off;
  int ks, missing, K, mode;
  const int32_t* site_stats;  // [3][V] of this pop
  int64_t V;
  const int32_t* idx_i;
  const int32_t* idx_j;
  int64_t N;
  int32_t* counts;   // [N][4] or [N][9]
  int32_t* nvalid;   // [N] or null
};

__global__ void __launch_bounds__(256) count_pairs_kernel(CountArgs a) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.N) return;
  const int64_t i = a.idx_i[t], j = a.idx_j[t];
  if (a.mode == PGLD_MODE_HAP) {
    int hh, hv, vh, vv;
    if (a.missing) {
      int pr[2][2];
      feature_products<2>(a.panels, a.block_bytes, a.pop_off, a.ks, i, j, pr);
      hh = pr[0][0]; hv = pr[0][1]; vh = pr[1][0]; vv = pr[1][1];
    } else {
      int pr[1][1];
      feature_products<1>(a.panels, a.block_bytes, a.pop_off, a.ks, i, j, pr);
      hh = pr[0][0]; hv = a.site_stats[i]; vh = a.site_stats[j]; vv = a.K;
    }
    int32_t* c = a.counts + 4 * t;
    c[0] = hh; c[1] = hv - hh; c[2] = vh - hh; c[3] = vv - hv - vh + hh;
    if (a.nvalid) a.nvalid[t] = vv;
  } else {
    // centred features u = g - 1, q = u^2, v = [g >= 0]: the pair's moments,
    // then its class counts by the exact inverse transform
    GenoMoments m = geno_pair_moments(a.panels, a.block_bytes, a.pop_off, a.ks, a.missing != 0,
                                      i, j);
    if (!a.missing) {
      const int* st = a.site_stats;
      m.a1 = 2 * st[a.V + i] + st[i] - a.K; m.a2 = a.K - st[i];
      m.b1 = 2 * st[a.V + j] + st[j] - a.K; m.b2 = a.K - st[j]; m.n0 = a.K;
    }
    int mc[9];
    geno_class_counts(m, mc);   // moments order n22 .. n00
    int32_t* c = a.counts + 9 * t;
#pragma unroll
    for (int k = 0; k < 9; ++k) c[k] = mc[8 - k];   // n00, n01, ..., n22
    if (a.nvalid) a.nvalid[t] = m.n0;
  }
}

// ---------------------------------------------------------------------------
// Per-pair haplotype estimators on FP64 counts (reference division form).
// ---------------------------------------------------------------------------
__device__ __forceinline__ HapPop hap_pop_div(double c1, double c2, double c3, double c4,
                                              double n) {
  HapPop p;
  p.c1 = c1; p.c2 = c2; p.c3 = c3; p.c4 = c4;
  hap_basis(p);
  const double f1 = n, f2 = f1 * (n - 1.0), f3 = f2 * (n - 2.0), f4 = f3 * (n - 3.0);
  p.r1 = f1 > 0.0 ? 1.0 / f1 : 0.0;
  p.r2 = f2 > 0.0 ? 1.0 / f2 : 0.0;
  p.r3 = f3 > 0.0 ? 1.0 / f3 : 0.0;
  p.r4 = f4 > 0.0 ? 1.0 / f4 : 0.0;
  return p;
}

// call codes: (kind, p0, p1, p2, p3) with kind 0 = dd, 1 = dz, 2 = pi2
__device__ double hap_raw_call(const HapPop* P, int kind, int i, int j, int k, int l) {
  if (kind == 0) return i == j ? hap_dd_same(P[i]) : hap_dd_pair(P[i], P[j]);
  if (kind == 1) {
    if (i == j && j == k) return hap_dz_same(P[i]);
    if (i == j) return hap_dz_iik(P[i], P[k]);
    if (i == k) return hap_dz_iji(P[i], P[j]);
    if (j == k) return hap_dz_ijj(P[i], P[j]);
    return hap_dz_ijk(P[i], P[j], P[k]);
  }
  int cnt[PGLD_MAX_POPS] = {0, 0, 0, 0};
  cnt[i]++; cnt[j]++; cnt[k]++; cnt[l]++;
  int nu = 0, mc = 0, t3 = -1, s1 = -1;
  for (int q = 0; q < PGLD_MAX_POPS; ++q) {
    nu += cnt[q] > 0;
    mc = cnt[q] > mc ? cnt[q] : mc;
    if (cnt[q] == 3) t3 = q;
    if (cnt[q] == 1) s1 = q;
  }
  if (nu == 1) return hap_pi2_same(P[i]);
  if (mc == 3) return hap_pi2_triple(P[t3], P[s1]);
  if (i == j && k == l) return hap_pi2_iikk(P[i], P[k]);
  if (i == j) return hap_pi2_iikl(P[i], P[k], P[l]);
  if (k == l) return hap_pi2_ijkk(P[i], P[j], P[k]);
  if ((i == k && j == l) || (i == l && j == k)) return hap_pi2_ijij(P[i], P[j]);
  if (nu == 3) {
    if (i == k) return hap_pi2_shared(P[i], P[j], P[l]);
    if (i == l) return hap_pi2_shared(P[i], P[j], P[k]);
    if (j == k) return hap_pi2_shared(P[j], P[i], P[l]);
    return hap_pi2_shared(P[j], P[i], P[k]);
  }
  return hap_pi2_alldiff(P[i], P[j], P[k], P[l]);
}

struct HapCallArgs {
  const double* counts;    // [N][stride], pop p at columns 4p..4p+3
  int64_t stride;
  const double* nvalid;    // [N][P] or null
  int64_t N;
  int P;
  int n_calls;
  int calls[64][5];
  double* out;             // [N][n_calls]
};

__device__ __forceinline__ void load_pops(const double* counts, int64_t stride, const double* nv,
                                          int P, int64_t t, HapPop* pp) {
  for (int p = 0; p < P; ++p) {
    const double* c = counts + t * stride + 4 * p;
    const double given = nv ? nv[t * P + p] : -1.0;  // < 0: use the sum of the counts
    const double n = given >= 0.0 ? given : (c[0] + c[1] + c[2] + c[3]);
    pp[p] = hap_pop_div(c[0], c[1], c[2], c[3], n);
  }
}

__global__ void __launch_bounds__(256) hap_calls_kernel(HapCallArgs a) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.N) return;
  HapPop pp[PGLD_MAX_POPS];
  load_pops(a.counts, a.stride, a.nvalid, a.P, t, pp);
  for (int c = 0; c < a.n_calls; ++c)
    a.out[t * a.n_calls + c] =
        hap_raw_call(pp, a.calls[c][0], a.calls[c][1], a.calls[c][2], a.calls[c][3], a.calls[c][4]);
}

template <int P>
__global__ void __launch_bounds__(256) hap_set_kernel(const double* counts, int64_t stride,
                                                      const double* nvalid, int64_t N,
                                                      double* out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N) return;
  HapPop pp[P];
  load_pops(counts, stride, nvalid, P, t, pp);
  double x[P == 1 ? HAP_NSTATS_P1 : P == 2 ? HAP_NSTATS_P2 : P == 3 ? HAP_NSTATS_P3 : HAP_NSTATS_P4];
  if constexpr (P == 1) hap_stats_P1(pp, x);
  if constexpr (P == 2) hap_stats_P2(pp, x);
  if constexpr (P == 3) hap_stats_P3(pp, x);
  if constexpr (P == 4) hap_stats_P4(pp, x);
  constexpr int NS = sizeof(x) / sizeof(double);
  for (int q = 0; q < NS; ++q) out[t * NS + q] = x[q];
}

// ---------------------------------------------------------------------------
// Per-pair genotype estimators on FP64 nine-way counts.
// ---------------------------------------------------------------------------
// c: the nine counts in the reference's order n00, n01, ..., n22 (first digit
// the genotype at locus i); n: valid individuals.
__device__ __forceinline__ GenoPop geno_pop_div(const double* c, double n) {
  GenoPop p;
#pragma unroll
  for (int m = 0; m < 9; ++m) p.g[m] = c[8 - m];  // moments order n22 .. n00
  double y[9];
  geno_centered(p.g, y);
  geno_set_centered(p, y);
  p.n = n;
  const double f1 = n, f2 = f1 * (n - 1.0), f3 = f2 * (n - 2.0), f4 = f3 * (n - 3.0);
  p.r1 = f1 > 0.0 ? 1.0 / f1 : 0.0;
  p.r2 = f2 > 0.0 ? 1.0 / f2 : 0.0;
  p.r3 = f3 > 0.0 ? 1.0 / f3 : 0.0;
  p.r4 = f4 > 0.0 ? 1.0 / f4 : 0.0;
  return p;
}

// Case dispatch of compute_all_{dd,dz,pi2}_geno (pg_gpu/genotype_kernels.py:
// 561-878), populations passed in call order as the reference does.
__device__ double geno_raw_call(const GenoPop* P, int kind, int i, int j, int k, int l) {
  if (kind == 0) return i == j ? geno_dd_same(P[i]) : geno_dd_pair(P[i], P[j]);
  if (kind == 1) {
    if (i == j && j == k) return geno_dz_same(P[i]);
    if (i == j) return geno_dz_iik(P[i], P[k]);
    if (i == k) return geno_dz_iji(P[i], P[j]);
    if (j == k) return geno_dz_ijj(P[i], P[j]);
    return geno_dz_ijk(P[i], P[j], P[k]);
  }
  int cnt[PGLD_MAX_POPS] = {0, 0, 0, 0};
  cnt[i]++; cnt[j]++; cnt[k]++; cnt[l]++;
  int nu = 0, mc = 0, t3 = -1, s1 = -1;
  for (int q = 0; q < PGLD_MAX_POPS; ++q) {
    nu += cnt[q] > 0;
    mc = cnt[q] > mc ? cnt[q] : mc;
    if (cnt[q] == 3) t3 = q;
    if (cnt[q] == 1) s1 = q;
  }
  if (nu == 1) return geno_pi2_same(P[i]);
  if (mc == 3)
    return (i == t3 && j == t3) ? geno_pi2_iiij(P[t3], P[s1]) : geno_pi2_ijii(P[t3], P[s1]);
  if (i == j && k == l) return geno_pi2_iikk(P[i], P[k]);
  if (i == j) return geno_pi2_iikl(P[i], P[k], P[l]);
  if (k == l) return geno_pi2_ijkk(P[i], P[j], P[k]);
  if ((i == k && j == l) || (i == l && j == k)) return geno_pi2_ijij(P[i], P[j]);
  if (nu == 3) {
    if (i == k) return geno_pi2_shared(P[i], P[j], P[l]);
    if (i == l) return geno_pi2_shared(P[i], P[j], P[k]);
    if (j == k) return geno_pi2_shared(P[j], P[i], P[l]);
    return geno_pi2_shared(P[j], P[i], P[k]);
  }
  return geno_pi2_alldiff(P[i], P[j], P[k], P[l]);
}

__device__ __forceinline__ void load_geno_pops(const double* counts, int64_t stride,
                                               const double* nv, int P, int64_t t, GenoPop* pp) {
  for (int p = 0; p < P; ++p) {
    const double* c = counts + t * stride + 9 * p;
    const double given = nv ? nv[t * P + p] : -1.0;  // < 0: use the sum of the counts
    double n = 0.0;
    for (int m = 0; m < 9; ++m) n += c[m];
    pp[p] = geno_pop_div(c, given >= 0.0 ? given : n);
  }
}

__global__ void __launch_bounds__(256) geno_calls_kernel(HapCallArgs a) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.N) return;
  GenoPop pp[PGLD_MAX_POPS];
  load_geno_pops(a.counts, a.stride, a.nvalid, a.P, t, pp);
  for (int c = 0; c < a.n_calls; ++c)
    a.out[t * a.n_calls + c] =
        geno_raw_call(pp, a.calls[c][0], a.calls[c][1], a.calls[c][2], a.calls[c][3], a.calls[c][4]);
}

template <int P>
__global__ void __launch_bounds__(256) geno_set_kernel(const double* counts, int64_t stride,
                                                       const double* nvalid, int64_t N,
                                                       double* out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N) return;
  GenoPop pp[P];
  load_geno_pops(counts, stride, nvalid, P, t, pp);
  double x[P == 1 ? GENO_NSTATS_P1 : P == 2 ? GENO_NSTATS_P2 : P == 3 ? GENO_NSTATS_P3
                                                                     : GENO_NSTATS_P4];
  if constexpr (P == 1) geno_stats_P1(pp, x);
  if constexpr (P == 2) geno_stats_P2(pp, x);
  if constexpr (P == 3) geno_stats_P3(pp, x);
  if constexpr (P == 4) geno_stats_P4(pp, x);
  constexpr int NS = sizeof(x) / sizeof(double);
  for (int q = 0; q < NS; ++q) out[t * NS + q] = x[q];
}

// sigma_D^2 = DD / pi2 of one population (ld_statistics_genotype.py:1518-1598,
// genotype_kernels.py:91-169): NaN where n < 4 or pi2 <= 0.
__global__ void __launch_bounds__(256) sigma_d2_geno_kernel(const double* counts, int64_t stride,
                                                            int64_t N, double* out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N) return;
  GenoPop pp[1];
  load_geno_pops(counts, stride, nullptr, 1, t, pp);
  pp[0].r4 = 1.0;  // the common 1/(n)_4 cancels: take the ratio of the numerators
  const double dd = geno_dd_same(pp[0]), pi2 = geno_pi2_same(pp[0]);
  out[t] = (pp[0].n >= 4.0 && pi2 > 0.0) ? dd / pi2 : __builtin_nan("");
}

// r, r^2, D' (haplotype_kernels.py:20-62 arithmetic)
__global__ void r_family_kernel(const double* counts, int64_t stride, const double* nvalid,
                                int64_t N, int which, double* out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N) return;
  const double* c = counts + t * stride;
  const double a = c[0], b = c[1], cc = c[2], d = c[3];
  const double n = nvalid ? nvalid[t] : a + b + cc + d;
  const double pA = (a + b) / n, pB = (a + cc) / n;
  const double D = (a * d - b * cc) / (n * n);
  double v;
  if (which == 2) {
    const double qA = 1.0 - pA, qB = 1.0 - pB;
    double dmax;
    if (D >= 0.0) {
      const double v1 = pA * qB, v2 = qA * pB;
      dmax = v1 < v2 ? v1 : v2;
    } else {
      const double v1 = pA * pB, v2 = qA * qB;
      dmax = v1 < v2 ? v1 : v2;
    }
    v = dmax > 0.0 ? D / dmax : __builtin_nan("");
  } else {
    const double den = pA * (1.0 - pA) * pB * (1.0 - pB);
    v = den > 0.0 ? (which == 0 ? D / sqrt(den) : D * D / den) : __builtin_nan("");
  }
  out[t] = v;
}

static unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace pgld

using namespace pgld;

extern "C" int pgld_pair_bounds(const double* d_pos, int64_t n_sites, double max_dist,
                                int64_t* d_jmax, void* stream) {
  if (n_sites < 0) return fail(PGLD_EINVAL, "negative n_sites");
  if (n_sites == 0) return PGLD_OK;
  hipLaunchKernelGGL(pair_bounds_kernel, dim3(blocks_for(n_sites)), dim3(256), 0,
                     as_stream(stream), d_pos, n_sites, max_dist, d_jmax);
  PGLD_LAUNCH_CHECK();
  return PGLD_OK;
}

extern "C" int pgld_emit_pairs(const int64_t* d_offsets, int64_t v_lo, int64_t n_anchors,
                               int64_t n_pairs, int32_t* d_idx_i, int32_t* d_idx_j, void* stream) {
  if (n_anchors < 1 || n_pairs < 0) return fail(PGLD_EINVAL, "bad pair chunk");
  if (n_pairs == 0) return PGLD_OK;
  hipLaunchKernelGGL(emit_pairs_kernel, dim3(blocks_for(n_pairs)), dim3(256), 0,
                     as_stream(stream), d_offsets, v_lo, n_anchors, n_pairs, d_idx_i, d_idx_j);
  PGLD_LAUNCH_CHECK();
  return PGLD_OK;
}

extern "C" int pgld_count_pairs(const pgld_panel_desc* desc, const int8_t* d_panels,
                                const int32_t* d_site_stats, int32_t pop, const int32_t* d_idx_i,
                                const int32_t* d_idx_j, int64_t n_pairs, int32_t* d_counts,
                                int32_t* d_nvalid, void* stream) {
  PanelLayout L;
  int rc = make_layout(desc, &L);
  if (rc) return rc;
  if (pop < 0 || pop >= desc->n_pops) return fail(PGLD_EINVAL, "pop out of range");
  if (n_pairs < 0) return fail(PGLD_EINVAL, "negative n_pairs");
  if (n_pairs == 0) return PGLD_OK;
  CountArgs a;
  a.panels = d_panels;
  a.block_bytes = L.block_bytes;
  a.pop_off = L.pop_off[pop];
  a.ks = L.ksteps[pop];
  a.missing = L.missing[pop];
  a.K = L.pop_size[pop];
  a.mode = desc->mode;
  a.site_stats = d_site_stats + (int64_t)pop * 3 * L.n_sites;
  a.V = L.n_sites;
  a.idx_i = d_idx_i;
  a.idx_j = d_idx_j;
  a.N = n_pairs;
  a.counts = d_counts;
  a.nvalid = d_nvalid;
  hipLaunchKernelGGL(count_pairs_kernel, dim3(blocks_for(n_pairs)), dim3(256), 0,
                     as_stream(stream), a);
  PGLD_LAUNCH_CHECK();
  return PGLD_OK;
}

extern "C" int pgld_hap_calls(const double* d_counts, int64_t stride, int32_t n_pops,
                              const double* d_nvalid, int64_t n_pairs, const int32_t* calls,
                              int32_t n_calls, double* d_out, void* stream) {
  if (n_pops < 1 || n_pops > PGLD_MAX_POPS) return fail(PGLD_EINVAL, "1-4 populations supported");
  if (stride < 4 * n_pops) return fail(PGLD_EINVAL, "count rows narrower than 4 * n_pops");
  if (n_calls < 1 || n_calls > 64) return fail(PGLD_EINVAL, "1..64 calls per launch");
  HapCallArgs a;
  a.counts = d_counts;
  a.stride = stride;
  a.nvalid = d_nvalid;
  a.N = n_pairs;
  a.P = n_pops;
  a.n_calls = n_calls;
  a.out = d_out;
  for (int c = 0; c < n_calls; ++c) {
    const int kind = calls[5 * c];
    if (kind < 0 || kind > 2) return fail(PGLD_EINVAL, "call kind must be 0 (dd), 1 (dz), 2 (pi2)");
    const int arity = kind + 2;
    for (int q = 0; q < 5; ++q) a.calls[c][q] = calls[5 * c + q];
    for (int q = 1; q <= arity; ++q)
      if (a.calls[c][q] < 0 || a.calls[c][q] >= n_pops)
        return fail(PGLD_EINVAL, "population index out of range");
    for (int q = arity + 1; q < 5; ++q) a.calls[c][q] = 0;
  }
  if (n_pairs == 0) return PGLD_OK;
  hipLaunchKernelGGL(hap_calls_kernel, dim3(blocks_for(n_pairs)), dim3(256), 0, as_stream(stream), a);
  PGLD_LAUNCH_CHECK();
  return PGLD_OK;
}

extern "C" int pgld_hap_stat_set(const double* d_counts, int64_t stride, int32_t n_pops,
                                 const double* d_nvalid, int64_t n_pairs, double* d_out,
                                 void* stream) {
  if (n_pops < 1 || n_pops > PGLD_MAX_POPS) return fail(PGLD_EINVAL, "1-4 populations supported");
  if (stride < 4 * n_pops) return fail(PGLD_EINVAL, "count rows narrower than 4 * n_pops");
  if (n_pairs == 0) return PGLD_OK;
  const dim3 g(blocks_for(n_pairs)), b(256);
  hipStream_t s = as_stream(stream);
  switch (n_pops) {
    case 1: hipLaunchKernelGGL(hap_set_kernel<1>, g, b, 0, s, d_counts, stride, d_nvalid, n_pairs, d_out); break;
    case 2: hipLaunchKernelGGL(hap_set_kernel<2>, g, b, 0, s, d_counts, stride, d_nvalid, n_pairs, d_out); break;
    case 3: hipLaunchKernelGGL(hap_set_kernel<3>, g, b, 0, s, d_counts, stride, d_nvalid, n_pairs, d_out); break;
    default: hipLaunchKernelGGL(hap_set_kernel<4>, g, b, 0, s, d_counts, stride, d_nvalid, n_pairs, d_out); break;
  }
  PGLD_LAUNCH_CHECK();
  return PGLD_OK;
}

extern "C" int pgld_r_family(const double* d_counts, int64_t stride, const double* d_nvalid,
                             int64_t n_pairs, int32_t which, double* d_out, void* stream) {
  if (which < 0 || which > 2) return fail(PGLD_EINVAL, "which must be 0 (r), 1 (r^2), 2 (D')");
  if (stride < 4) return fail(PGLD_EINVAL, "count rows narrower than 4");
  if (n_pairs == 0) return PGLD_OK;
  hipLaunchKernelGGL(r_family_kernel, dim3(blocks_for(n_pairs)), dim3(256), 0, as_stream(stream),
                     d_counts, stride, d_nvalid, n_pairs, which, d_out);
  PGLD_LAUNCH_CHECK();
  return PGLD_OK;
}

static int check_calls(const int32_t* calls, int32_t n_calls, int32_t n_pops, HapCallArgs* a) {
  if (n_calls < 1 || n_calls > 64) return fail(PGLD_EINVAL, "1..64 calls per launch");
  for (int c = 0; c < n_calls; ++c) {
    const int kind = calls[5 * c];
    if (kind < 0 || kind > 2) return fail(PGLD_EINVAL, "call kind must be 0 (dd), 1 (dz), 2 (pi2)");
    const int arity = kind + 2;
    for (int q = 0; q < 5; ++q) a->calls[c][q] = calls[5 * c + q];
    for (int q = 1; q <= arity; ++q)
      if (a->calls[c][q] < 0 || a->calls[c][q] >= n_pops)
        return fail(PGLD_EINVAL, "population index out of range");
    for (int q = arity + 1; q < 5; ++q) a->calls[c][q] = 0;
  }
  a->n_calls = n_calls;
  return PGLD_OK;
}

extern "C" int pgld_geno_calls(const double* d_counts, int64_t stride, int32_t n_pops,
                               const double* d_nvalid, int64_t n_pairs, const int32_t* calls,
                               int32_t n_calls, double* d_out, void* stream) {
  if (n_pops < 1 || n_pops > PGLD_MAX_POPS) return fail(PGLD_EINVAL, "1-4 populations supported");
  if (stride < 9 * n_pops) return fail(PGLD_EINVAL, "count rows narrower than 9 * n_pops");
  HapCallArgs a;
  int rc = check_calls(calls, n_calls, n_pops, &a);
  if (rc) return rc;
  a.counts = d_counts;
  a.stride = stride;
  a.nvalid = d_nvalid;
  a.N = n_pairs;
  a.P = n_pops;
  a.out = d_out;
  if (n_pairs == 0) return PGLD_OK;
  hipLaunchKernelGGL(geno_calls_kernel, dim3(blocks_for(n_pairs)), dim3(256), 0, as_stream(stream),
                     a);
  PGLD_LAUNCH_CHECK();
  return PGLD_OK;
}

extern "C" int pgld_geno_stat_set(const double* d_counts, int64_t stride, int32_t n_pops,
                                  const double* d_nvalid, int64_t n_pairs, double* d_out,
                                  void* stream) {
  if (n_pops < 1 || n_pops > PGLD_MAX_POPS) return fail(PGLD_EINVAL, "1-4 populations supported");
  if (stride < 9 * n_pops) return fail(PGLD_EINVAL, "count rows narrower than 9 * n_pops");
  if (n_pairs == 0) return PGLD_OK;
  const dim3 g(blocks_for(n_pairs)), b(256);
  hipStream_t s = as_stream(stream);
  switch (n_pops) {
    case 1: hipLaunchKernelGGL(geno_set_kernel<1>, g, b, 0, s, d_counts, stride, d_nvalid, n_pairs, d_out); break;
    case 2: hipLaunchKernelGGL(geno_set_kernel<2>, g, b, 0, s, d_counts, stride, d_nvalid, n_pairs, d_out); break;
    case 3: hipLaunchKernelGGL(geno_set_kernel<3>, g, b, 0, s, d_counts, stride, d_nvalid, n_pairs, d_out); break;
    default: hipLaunchKernelGGL(geno_set_kernel<4>, g, b, 0, s, d_counts, stride, d_nvalid, n_pairs, d_out); break;
  }
  PGLD_LAUNCH_CHECK();
  return PGLD_OK;
}

extern "C" int pgld_sigma_d2_geno(const double* d_counts, int64_t stride, int64_t n_pairs,
                                  double* d_out, void* stream) {
  if (stride < 9) return fail(PGLD_EINVAL, "count rows narrower than 9");
  if (n_pairs == 0) return PGLD_OK;
  hipLaunchKernelGGL(sigma_d2_geno_kernel, dim3(blocks_for(n_pairs)), dim3(256), 0,
                     as_stream(stream), d_counts, stride, n_pairs, d_out);
  PGLD_LAUNCH_CHECK();
  return PGLD_OK;
}

