// Plain per-pair FP64 comparator and the dense count-block kernel.
//
// plain_kernel is the straightforward GPU form of the reference's chunk
// loop: one workgroup per anchor site i, one thread per partner j, a loop over
// the population's sample rows of the raw int8 matrix to form the four
// haplotype (or nine genotype) counts (the reference does this with gathers
// and boolean reductions, pg_gpu/ld_pipeline.py:155-172 and :216-250), the
// same FP64 estimators as the fused kernel, and per-workgroup bin partials.  It reports the same
// fixed-point output as the fused kernel so the two can be compared exactly.
#include "pgld_internal.h"
#include "stats_hap.inc"
#include "stats_geno.inc"

namespace pgld {

int validate_binned(const pgld_panel_desc* desc, const pgld_binned_args* g, PanelLayout* L);

struct PlainParams {
  const int8_t* mat;
  int64_t ld;
  const int32_t* rows;
  int32_t row_offset[PGLD_MAX_POPS + 1];
  const double* pos;
  const double* coord;
  int32_t coord_is_pos;
  double max_dist;
  const double* edges;
  int32_t n_bins;
  const double* rtab;
  unsigned long long* out;
  int64_t V;
};

// Per-pair counting straight from the int8 matrix, and the statistic set.
template <int MODE, int P> struct PlainSet;
template <int P> struct PlainSet<PGLD_MODE_HAP, P> {
  static constexpr int NS = P == 1 ? HAP_NSTATS_P1 : P == 2 ? HAP_NSTATS_P2 : P == 3 ? HAP_NSTATS_P3
                                                                                     : HAP_NSTATS_P4;
  using Pop = HapPop;
  __device__ static Pop pop(const PlainParams& a, const int32_t* srows, int p, int64_t i,
                            int64_t j) {
    int n11 = 0, n10 = 0, n01 = 0, n00 = 0;
    for (int r = a.row_offset[p]; r < a.row_offset[p + 1]; ++r) {
      const int8_t* row = a.mat + (int64_t)srows[r] * a.ld;
      const int x = row[i], y = row[j];
      n11 += (x == 1) & (y == 1);
      n10 += (x == 1) & (y == 0);
      n01 += (x == 0) & (y == 1);
      n00 += (x == 0) & (y == 0);
    }
    return hap_pop(n11, n10, n01, n00, n11 + n10 + n01 + n00, a.rtab);
  }
  __device__ static void eval(const Pop* p, double* o) {
    if constexpr (P == 1) hap_stats_P1(p, o);
    if constexpr (P == 2) hap_stats_P2(p, o);
    if constexpr (P == 3) hap_stats_P3(p, o);
    if constexpr (P == 4) hap_stats_P4(p, o);
  }
  __device__ static void r2(const Pop& p, double* o) { hap_r2(p, o); }
};
template <int P> struct PlainSet<PGLD_MODE_GENO, P> {
  static constexpr int NS = P == 1 ? GENO_NSTATS_P1 : P == 2 ? GENO_NSTATS_P2 : P == 3 ? GENO_NSTATS_P3
                                                                                        : GENO_NSTATS_P4;
  using Pop = GenoPop;
  __device__ static Pop pop(const PlainParams& a, const int32_t* srows, int p, int64_t i,
                            int64_t j) {
    // class of (x, y) in moments order: n22, n21, n20, n12, n11, n10, n02, n01, n00
    int c[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int r = a.row_offset[p]; r < a.row_offset[p + 1]; ++r) {
      const int8_t* row = a.mat + (int64_t)srows[r] * a.ld;
      const int x = row[i], y = row[j];
      if (x >= 0 && y >= 0) c[(2 - x) * 3 + (2 - y)] += 1;
    }
    int y[9];
    geno_centered(c, y);  // y[0] = valid individuals
    return geno_pop(c, y, y[0], a.rtab);
  }
  __device__ static void eval(const Pop* p, double* o) {
    if constexpr (P == 1) geno_stats_P1(p, o);
    if constexpr (P == 2) geno_stats_P2(p, o);
    if constexpr (P == 3) geno_stats_P3(p, o);
    if constexpr (P == 4) geno_stats_P4(p, o);
  }
  __device__ static void r2(const Pop& p, double* o) { geno_r2(p, o); }
};

template <int NS>
__device__ __forceinline__ void plain_flush(unsigned long long* slimb, int bin, const double* acc,
                                            int count) {
  unsigned long long* dst = slimb + (int64_t)bin * (NS + 1) * 3;
  for (int q = 0; q < NS; ++q) {
    long long x2, x1, x0;
    pgld_to_limbs(acc[q], x2, x1, x0);
    atomicAdd(dst + 3 * q + 0, (unsigned long long)x0);
    atomicAdd(dst + 3 * q + 1, (unsigned long long)x1);
    atomicAdd(dst + 3 * q + 2, (unsigned long long)x2);
  }
  atomicAdd(dst + 3 * NS, (unsigned long long)count);
}

template <int MODE, int P, bool R2>
__global__ void __launch_bounds__(256) plain_kernel(PlainParams a) {
  using S = PlainSet<MODE, P>;
  constexpr int NS = S::NS + (R2 ? 2 : 0);  // output row width
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  double* sedges = reinterpret_cast<double*>(smem);
  unsigned long long* slimb = reinterpret_cast<unsigned long long*>(sedges + PGLD_MAX_BINS + 8);
  int32_t* srows = reinterpret_cast<int32_t*>(slimb + (size_t)a.n_bins * (NS + 1) * 3);
  __shared__ int64_t s_jmax;
  const int nb = a.n_bins;
  const int64_t i = blockIdx.x;
  const int t = threadIdx.x;
  const int nrows = a.row_offset[P];
  for (int e = t; e <= nb; e += blockDim.x) sedges[e] = a.edges[e];
  for (int e = t; e < nb * (NS + 1) * 3; e += blockDim.x) slimb[e] = 0ull;
  for (int e = t; e < nrows; e += blockDim.x) srows[e] = a.rows[e];
  const double pos_i = a.pos[i];
  const double ub = pos_i + a.max_dist;
  if (t == 0) {  // first index with pos > ub (right-sided searchsorted)
    int64_t lo = i + 1, hi = a.V;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (a.pos[mid] <= ub) lo = mid + 1; else hi = mid;
    }
    s_jmax = lo;
  }
  __syncthreads();
  const int64_t jmax = s_jmax;
  const double co_i = a.coord[i];
  int cur = -1, count = 0, guess = 0;
  double acc[NS];
  for (int q = 0; q < NS; ++q) acc[q] = 0.0;
  for (int64_t j = i + 1 + t; j < jmax; j += blockDim.x) {
    const double d = a.coord_is_pos ? (a.pos[j] - pos_i) : fabs(a.coord[j] - co_i);
    const int b = pgld_bin_scan(d, sedges, nb, guess, nb - 1);
    if (b < 0 || b >= nb) continue;
    guess = b;
    typename S::Pop pp[P];
    for (int p = 0; p < P; ++p) pp[p] = S::pop(a, srows, p, i, j);
    double x[NS];
    S::eval(pp, x);
    if constexpr (R2) S::r2(pp[0], x + S::NS);
    if (b != cur) {
      if (cur >= 0) plain_flush<NS>(slimb, cur, acc, count);
      cur = b;
      count = 0;
      for (int q = 0; q < NS; ++q) acc[q] = 0.0;
    }
    for (int q = 0; q < NS; ++q) acc[q] += x[q];
    ++count;
  }
  if (cur >= 0) plain_flush<NS>(slimb, cur, acc, count);
  __syncthreads();
  for (int e = t; e < nb * (NS + 1); e += blockDim.x) {
    unsigned long long* L = slimb + 3 * e;
    if (e % (NS + 1) == NS) {
      if (L[0]) atomicAdd(a.out + 3 * e, L[0]);
      continue;
    }
    long long l0 = (long long)L[0], l1 = (long long)L[1], l2 = (long long)L[2];
    pgld_normalize_limbs(l2, l1, l0);
    if (l0) atomicAdd(a.out + 3 * e + 0, (unsigned long long)l0);
    if (l1) atomicAdd(a.out + 3 * e + 1, (unsigned long long)l1);
    if (l2) atomicAdd(a.out + 3 * e + 2, (unsigned long long)l2);
  }
}

template <int MODE, int P, bool R2 = false>
static int launch_plain(const PanelLayout& L, const pgld_binned_args* g, hipStream_t s) {
  constexpr int NS = PlainSet<MODE, P>::NS + (R2 ? 2 : 0);
  PlainParams a;
  a.mat = g->d_mat;
  a.ld = g->ld;
  a.rows = g->d_rows;
  for (int p = 0; p <= PGLD_MAX_POPS; ++p) a.row_offset[p] = p <= P ? g->row_offset[p] : 0;
  a.pos = g->d_pos;
  a.coord = g->d_coord ? g->d_coord : g->d_pos;
  a.coord_is_pos = g->coord_is_pos || g->d_coord == nullptr || g->d_coord == g->d_pos;
  a.max_dist = g->max_dist;
  a.edges = g->d_edges;
  a.n_bins = g->n_bins;
  a.rtab = g->d_rtab;
  a.out = reinterpret_cast<unsigned long long*>(g->d_out);
  a.V = L.n_sites;
  for (int p = 0; p < P; ++p)
    if (a.row_offset[p + 1] - a.row_offset[p] != L.pop_size[p])
      return fail(PGLD_EINVAL, "row_offset disagrees with pop_size");
  const size_t lds = (PGLD_MAX_BINS + 8) * sizeof(double) +
                     (size_t)g->n_bins * (NS + 1) * 3 * sizeof(unsigned long long) +
                     (size_t)a.row_offset[P] * sizeof(int32_t);
  if (lds > 160 * 1024) return fail(PGLD_EINVAL, "plain kernel: too many sample rows for LDS");
  PGLD_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(plain_kernel<MODE, P, R2>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  if (L.n_sites < 2) return PGLD_OK;
  hipLaunchKernelGGL((plain_kernel<MODE, P, R2>), dim3((unsigned)(L.n_sites - 1)), dim3(256), lds, s, a);
  PGLD_LAUNCH_CHECK();
  return PGLD_OK;
}


// ---------------------------------------------------------------------------
// Ragged windows, plain form: one workgroup per anchor site i, one thread per
// partner j in i's window, per-pair sample loops; workgroup tree reduction.
// ---------------------------------------------------------------------------
struct PlainWinParams {
  PlainParams p;
  const int64_t* win;  // [n_windows][2], sorted, non-overlapping
  int64_t n_windows;
};

template <int MODE>
__global__ void __launch_bounds__(256) plain_window_kernel(PlainWinParams w) {
  using S = PlainSet<MODE, 1>;
  constexpr int NS = S::NS + 2;
  __shared__ double red[256];
  __shared__ int64_t s_win;
  const PlainParams& a = w.p;
  const int64_t i = blockIdx.x;
  const int t = threadIdx.x;
  if (t == 0) {  // last window with lo <= i
    int64_t lo = 0, hi = w.n_windows;
    while (hi - lo > 1) {
      const int64_t mid = (lo + hi) >> 1;
      if (w.win[2 * mid] <= i) lo = mid; else hi = mid;
    }
    s_win = (w.n_windows > 0 && w.win[2 * lo] <= i && i < w.win[2 * lo + 1]) ? lo : -1;
  }
  __syncthreads();
  const int64_t win = s_win;
  if (win < 0) return;  // uniform across the workgroup
  const int64_t hi = w.win[2 * win + 1];
  double acc[NS];
  for (int q = 0; q < NS; ++q) acc[q] = 0.0;
  int count = 0;
  for (int64_t j = i + 1 + t; j < hi; j += blockDim.x) {
    typename S::Pop pp[1];
    pp[0] = S::pop(a, a.rows, 0, i, j);
    double x[NS];
    S::eval(pp, x);
    S::r2(pp[0], x + S::NS);
    for (int q = 0; q < NS; ++q) acc[q] += x[q];
    ++count;
  }
  unsigned long long* dst = a.out + (int64_t)win * (NS + 1) * 3;
  for (int q = 0; q <= NS; ++q) {
    red[t] = q < NS ? acc[q] : (double)count;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
      if (t < h) red[t] += red[t + h];
      __syncthreads();
    }
    if (t == 0) {
      const double v = red[0];
      if (q < NS) {
        long long x2, x1, x0;
        pgld_to_limbs(v, x2, x1, x0);
        if (x0) atomicAdd(dst + 3 * q + 0, (unsigned long long)x0);
        if (x1) atomicAdd(dst + 3 * q + 1, (unsigned long long)x1);
        if (x2) atomicAdd(dst + 3 * q + 2, (unsigned long long)x2);
      } else if (v > 0.0) {
        atomicAdd(dst + 3 * NS, (unsigned long long)v);
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Dense counts of one population for a rectangle of sites (MFMA path).
// One wave per 32x32 block; the same operand panels as the fused kernel.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(64) count_block_hap_kernel(
    const int8_t* __restrict__ panels, int64_t block_bytes, int64_t pop_off, int ks, int missing,
    const int32_t* __restrict__ cnt1, int K, int64_t V, int64_t i0, int64_t ni, int64_t j0,
    int64_t nj, int32_t* __restrict__ out) {
  const int lane = threadIdx.x;
  const int64_t rb = i0 / 32 + blockIdx.y, cb = j0 / 32 + blockIdx.x;
  const v4i* A = reinterpret_cast<const v4i*>(panels + rb * block_bytes + pop_off) + lane;
  const v4i* B = reinterpret_cast<const v4i*>(panels + cb * block_bytes + pop_off) + lane;
  v16f hh = zero16f(), hv = zero16f(), vh = zero16f(), vv = zero16f();
  for (int s = 0; s < ks; ++s) {
    const v4i ah = A[s * 64], bh = B[s * 64];
    hh = pgld_mfma_fp4(ah, bh, hh);
    if (missing) {
      const v4i av = A[(ks + s) * 64], bv = B[(ks + s) * 64];
      hv = pgld_mfma_fp4(ah, bv, hv);
      vh = pgld_mfma_fp4(av, bh, vh);
      vv = pgld_mfma_fp4(av, bv, vv);
    }
  }
  const int64_t j = cb * 32 + (lane & 31);
  for (int r = 0; r < 16; ++r) {
    const int64_t i = rb * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    if (i >= i0 + ni || j >= j0 + nj || i >= V || j >= V) continue;
    int hvr, vhr, vvr;
    if (missing) {
      hvr = (int)hv[r]; vhr = (int)vh[r]; vvr = (int)vv[r];
    } else {
      hvr = cnt1[i]; vhr = cnt1[j]; vvr = K;
    }
    const int n11 = (int)hh[r];
    const int64_t o = (i - i0) * nj + (j - j0);
    const int64_t plane = ni * nj;
    out[o] = n11;
    out[plane + o] = hvr - n11;
    out[2 * plane + o] = vhr - n11;
    out[3 * plane + o] = vvr - hvr - vhr + n11;
  }
}

// Genotype variant: nine planes in moments class order n22 .. n00.
__global__ void __launch_bounds__(64) count_block_geno_kernel(
    const int8_t* __restrict__ panels, int64_t block_bytes, int64_t pop_off, int ks, int missing,
    const int32_t* __restrict__ st, int K, int64_t V, int64_t i0, int64_t ni, int64_t j0,
    int64_t nj, int32_t* __restrict__ out) {
  const int lane = threadIdx.x;
  const int64_t rb = i0 / 32 + blockIdx.y, cb = j0 / 32 + blockIdx.x;
  const v4i* A = reinterpret_cast<const v4i*>(panels + rb * block_bytes + pop_off) + lane;
  const v4i* B = reinterpret_cast<const v4i*>(panels + cb * block_bytes + pop_off) + lane;
  const int nf = missing ? 3 : 2;
  v16f pr[3][3];
  for (int x = 0; x < 3; ++x)
    for (int y = 0; y < 3; ++y) pr[x][y] = zero16f();
  for (int s = 0; s < ks; ++s) {
    v4i fa[3], fb[3];
    for (int f = 0; f < nf; ++f) { fa[f] = A[(f * ks + s) * 64]; fb[f] = B[(f * ks + s) * 64]; }
    for (int x = 0; x < nf; ++x)
      for (int y = 0; y < nf; ++y) pr[x][y] = pgld_mfma_fp4(fa[x], fb[y], pr[x][y]);
  }
  const int64_t j = cb * 32 + (lane & 31);
  for (int r = 0; r < 16; ++r) {
    const int64_t i = rb * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    if (i >= i0 + ni || j >= j0 + nj || i >= V || j >= V) continue;
    // the centred moments from the feature products (u, q, v), then the
    // class counts by the exact inverse transform
    GenoMoments m;
    m.ab11 = (int)pr[0][0][r]; m.ab12 = (int)pr[0][1][r];
    m.ab21 = (int)pr[1][0][r]; m.ab22 = (int)pr[1][1][r];
    if (missing) {
      m.a1 = (int)pr[0][2][r]; m.a2 = (int)pr[1][2][r];
      m.b1 = (int)pr[2][0][r]; m.b2 = (int)pr[2][1][r]; m.n0 = (int)pr[2][2][r];
    } else {
      m.a1 = 2 * st[V + i] + st[i] - K; m.a2 = K - st[i];
      m.b1 = 2 * st[V + j] + st[j] - K; m.b2 = K - st[j]; m.n0 = K;
    }
    int c[9];
    geno_class_counts(m, c);   // moments order
    const int64_t o = (i - i0) * nj + (j - j0);
    const int64_t plane = ni * nj;
    for (int m = 0; m < 9; ++m) out[m * plane + o] = c[m];
  }
}

}  // namespace pgld

using namespace pgld;

extern "C" int pgld_binned_plain(const pgld_panel_desc* desc, const pgld_binned_args* args,
                                 void* stream) {
  PanelLayout L;
  int rc = validate_binned(desc, args, &L);
  if (rc) return rc;
  if (!args->d_mat || !args->d_rows) return fail(PGLD_EINVAL, "plain kernel needs d_mat and d_rows");
  hipStream_t s = as_stream(stream);
  if (args->with_r2) {  // one population (checked by validate_binned)
    return desc->mode == PGLD_MODE_HAP ? launch_plain<PGLD_MODE_HAP, 1, true>(L, args, s)
                                       : launch_plain<PGLD_MODE_GENO, 1, true>(L, args, s);
  }
  if (desc->mode == PGLD_MODE_HAP) {
    switch (desc->n_pops) {
      case 1: return launch_plain<PGLD_MODE_HAP, 1>(L, args, s);
      case 2: return launch_plain<PGLD_MODE_HAP, 2>(L, args, s);
      case 3: return launch_plain<PGLD_MODE_HAP, 3>(L, args, s);
      case 4: return launch_plain<PGLD_MODE_HAP, 4>(L, args, s);
      default: break;
    }
  } else {
    switch (desc->n_pops) {
      case 1: return launch_plain<PGLD_MODE_GENO, 1>(L, args, s);
      case 2: return launch_plain<PGLD_MODE_GENO, 2>(L, args, s);
      case 3: return launch_plain<PGLD_MODE_GENO, 3>(L, args, s);
      case 4: return launch_plain<PGLD_MODE_GENO, 4>(L, args, s);
      default: break;
    }
  }
  return fail(PGLD_EUNSUPPORTED, "plain kernel: this (mode, n_pops) is not built yet");
}

extern "C" int pgld_count_block(const pgld_panel_desc* desc, const int8_t* d_panels,
                                const int32_t* d_site_stats, int32_t pop, int64_t i0, int64_t ni,
                                int64_t j0, int64_t nj, int32_t* d_counts, void* stream) {
  PanelLayout L;
  int rc = make_layout(desc, &L);
  if (rc) return rc;
  if (pop < 0 || pop >= desc->n_pops) return fail(PGLD_EINVAL, "pop out of range");
  if (i0 % 32 || j0 % 32 || i0 < 0 || j0 < 0 || ni < 0 || nj < 0)
    return fail(PGLD_EINVAL, "i0/j0 must be non-negative multiples of 32");
  if (i0 + ni > L.n_sites || j0 + nj > L.n_sites) return fail(PGLD_EINVAL, "block out of range");
  if (ni == 0 || nj == 0) return PGLD_OK;
  dim3 grid((unsigned)((nj + 31) / 32), (unsigned)((ni + 31) / 32));
  hipLaunchKernelGGL(desc->mode == PGLD_MODE_HAP ? count_block_hap_kernel : count_block_geno_kernel, grid, dim3(64), 0, as_stream(stream), d_panels,
                     L.block_bytes, L.pop_off[pop], L.ksteps[pop], L.missing[pop],
                     d_site_stats + (int64_t)pop * 3 * L.n_sites, L.pop_size[pop], L.n_sites, i0,
                     ni, j0, nj, d_counts);
  PGLD_LAUNCH_CHECK();
  return PGLD_OK;
}

extern "C" int pgld_windows_plain(const pgld_panel_desc* desc, const pgld_window_args* args,
                                  void* stream) {
  PanelLayout L;
  int rc = make_layout(desc, &L);
  if (rc) return rc;
  if (!args) return fail(PGLD_EINVAL, "null window args");
  if (desc->n_pops != 1) return fail(PGLD_EUNSUPPORTED, "window sums are built for one population");
  if (!args->d_mat || !args->d_rows || !args->d_rtab || !args->d_windows || !args->d_out)
    return fail(PGLD_EINVAL, "null device pointer in window args");
  if (args->n_windows <= 0 || L.n_sites == 0) return PGLD_OK;
  PlainWinParams w{};
  w.p.mat = args->d_mat;
  w.p.ld = args->ld;
  w.p.rows = args->d_rows;
  w.p.row_offset[0] = 0;
  w.p.row_offset[1] = L.pop_size[0];
  w.p.rtab = args->d_rtab;
  w.p.out = reinterpret_cast<unsigned long long*>(args->d_out);
  w.p.V = L.n_sites;
  w.win = args->d_windows;
  w.n_windows = args->n_windows;
  hipStream_t s = as_stream(stream);
  if (desc->mode == PGLD_MODE_HAP)
    hipLaunchKernelGGL(plain_window_kernel<PGLD_MODE_HAP>, dim3((unsigned)L.n_sites), dim3(256), 0,
                       s, w);
  else
    hipLaunchKernelGGL(plain_window_kernel<PGLD_MODE_GENO>, dim3((unsigned)L.n_sites), dim3(256),
                       0, s, w);
  PGLD_LAUNCH_CHECK();
  return PGLD_OK;
}

