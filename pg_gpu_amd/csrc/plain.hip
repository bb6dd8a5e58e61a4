// Plain per-pair FP64 comparator and the dense count-block kernel.
//
// plain_hap_kernel is the straightforward GPU form of the reference's chunk
// loop: one workgroup per anchor site i, one thread per partner j, a loop over
// the population's sample rows of the raw int8 matrix to form the four
// haplotype counts (the reference does this with gathers and boolean
// reductions, pg_gpu/ld_pipeline.py:155-172), the same FP64 estimators as the
// fused kernel, and per-workgroup bin partials.  It reports the same
// fixed-point output as the fused kernel so the two can be compared exactly.
#include "pgld_internal.h"
#include "stats_hap.inc"

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

template <int P> struct PlainSet;
template <> struct PlainSet<1> {
  static constexpr int NS = HAP_NSTATS_P1;
  __device__ static void eval(const HapPop* p, double* o) { hap_stats_P1(p, o); }
};
template <> struct PlainSet<2> {
  static constexpr int NS = HAP_NSTATS_P2;
  __device__ static void eval(const HapPop* p, double* o) { hap_stats_P2(p, o); }
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

template <int P>
__global__ void __launch_bounds__(256) plain_hap_kernel(PlainParams a) {
  constexpr int NS = PlainSet<P>::NS;
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
    HapPop pp[P];
    for (int p = 0; p < P; ++p) {
      int n11 = 0, n10 = 0, n01 = 0, n00 = 0;
      for (int r = a.row_offset[p]; r < a.row_offset[p + 1]; ++r) {
        const int8_t* row = a.mat + (int64_t)srows[r] * a.ld;
        const int x = row[i], y = row[j];
        n11 += (x == 1) & (y == 1);
        n10 += (x == 1) & (y == 0);
        n01 += (x == 0) & (y == 1);
        n00 += (x == 0) & (y == 0);
      }
      pp[p] = hap_pop(n11, n10, n01, n00, n11 + n10 + n01 + n00, a.rtab);
    }
    double x[NS];
    PlainSet<P>::eval(pp, x);
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

template <int P>
static int launch_plain_hap(const PanelLayout& L, const pgld_binned_args* g, hipStream_t s) {
  constexpr int NS = PlainSet<P>::NS;
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
  PGLD_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(plain_hap_kernel<P>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  if (L.n_sites < 2) return PGLD_OK;
  hipLaunchKernelGGL(plain_hap_kernel<P>, dim3((unsigned)(L.n_sites - 1)), dim3(256), lds, s, a);
  PGLD_LAUNCH_CHECK();
  return PGLD_OK;
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
  v16i hh, hv, vh, vv;
  for (int q = 0; q < 16; ++q) hh[q] = hv[q] = vh[q] = vv[q] = 0;
  for (int s = 0; s < ks; ++s) {
    const v4i ah = A[s * 64], bh = B[s * 64];
    hh = __builtin_amdgcn_mfma_i32_32x32x32_i8(ah, bh, hh, 0, 0, 0);
    if (missing) {
      const v4i av = A[(ks + s) * 64], bv = B[(ks + s) * 64];
      hv = __builtin_amdgcn_mfma_i32_32x32x32_i8(ah, bv, hv, 0, 0, 0);
      vh = __builtin_amdgcn_mfma_i32_32x32x32_i8(av, bh, vh, 0, 0, 0);
      vv = __builtin_amdgcn_mfma_i32_32x32x32_i8(av, bv, vv, 0, 0, 0);
    }
  }
  const int64_t j = cb * 32 + (lane & 31);
  for (int r = 0; r < 16; ++r) {
    const int64_t i = rb * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    if (i >= i0 + ni || j >= j0 + nj || i >= V || j >= V) continue;
    int hvr, vhr, vvr;
    if (missing) {
      hvr = hv[r]; vhr = vh[r]; vvr = vv[r];
    } else {
      hvr = cnt1[i]; vhr = cnt1[j]; vvr = K;
    }
    const int n11 = hh[r];
    const int64_t o = (i - i0) * nj + (j - j0);
    const int64_t plane = ni * nj;
    out[o] = n11;
    out[plane + o] = hvr - n11;
    out[2 * plane + o] = vhr - n11;
    out[3 * plane + o] = vvr - hvr - vhr + n11;
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
  if (desc->mode == PGLD_MODE_HAP) {
    switch (desc->n_pops) {
      case 1: return launch_plain_hap<1>(L, args, s);
      case 2: return launch_plain_hap<2>(L, args, s);
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
  if (desc->mode != PGLD_MODE_HAP)
    return fail(PGLD_EUNSUPPORTED, "count_block: genotype mode not built yet");
  if (ni == 0 || nj == 0) return PGLD_OK;
  dim3 grid((unsigned)((nj + 31) / 32), (unsigned)((ni + 31) / 32));
  hipLaunchKernelGGL(count_block_hap_kernel, grid, dim3(64), 0, as_stream(stream), d_panels,
                     L.block_bytes, L.pop_off[pop], L.ksteps[pop], L.missing[pop],
                     d_site_stats + (int64_t)pop * 3 * L.n_sites, L.pop_size[pop], L.n_sites, i0,
                     ni, j0, nj, d_counts);
  PGLD_LAUNCH_CHECK();
  return PGLD_OK;
}
