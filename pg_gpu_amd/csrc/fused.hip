// Fused binned two-locus LD for haplotype data on gfx950.
//
// One persistent workgroup (4 waves) walks a list of 64x64-site tiles of the
// distance band.  Per tile:
//   1. counting -- each wave multiplies the packed int8 indicator panels of a
//      32-row x 32-column sub-tile on v_mfma_i32_32x32x32_i8 (int32
//      accumulation, exact): H.H, H.V, V.H, V.V give n11, n10, n01, n00 of
//      every pair (only H.H when the population has no missing calls), and
//      stores them as int16 in LDS;
//   2. epilogue -- each lane takes one column j and 16 rows i of the tile,
//      evaluates the unbiased DD/Dz/pi2 estimators in FP64 registers,
//      digitizes the pair distance and adds the values to a running sum for
//      the current bin (bins only decrease along the lane's rows, so runs are
//      long); a run is flushed when the bin changes and at the tile end, with
//      a wave reduction when the whole wave is in one bin.
//   3. flushes convert the FP64 run sums to 3-limb fixed point and add them
//      with integer LDS atomics; at exit each workgroup adds its limbs to the
//      global result with integer atomics.  Integer addition is associative,
//      so the bin sums are bitwise reproducible whatever the schedule.
// No pair list and no per-pair statistic ever reaches HBM.
#include "pgld_internal.h"
#include "stats_hap.inc"

namespace pgld {

struct FusedParams {
  const int8_t* panels;
  const int32_t* site_stats;  // [P][3][V]
  const double* pos;
  const double* coord;
  int32_t coord_is_pos;
  double max_dist;
  const double* edges;
  int32_t n_bins;
  const double* rtab;
  const int32_t* tiles;
  int64_t n_tiles;
  unsigned long long* out;
  int64_t V;
  int64_t block_bytes;
  int64_t pop_off[PGLD_MAX_POPS];
  int32_t ksteps[PGLD_MAX_POPS];
  int32_t missing[PGLD_MAX_POPS];
  int32_t K[PGLD_MAX_POPS];
};

template <int P> struct HapSet;
template <> struct HapSet<1> {
  static constexpr int NS = HAP_NSTATS_P1;
  __device__ static void eval(const HapPop* p, double* o) { hap_stats_P1(p, o); }
};
template <> struct HapSet<2> {
  static constexpr int NS = HAP_NSTATS_P2;
  __device__ static void eval(const HapPop* p, double* o) { hap_stats_P2(p, o); }
};

constexpr int TILE = 64;

__device__ __forceinline__ v16i zero16() {
  v16i z;
#pragma unroll
  for (int q = 0; q < 16; ++q) z[q] = 0;
  return z;
}

// Row (within a 32x32 MFMA output block) of accumulator register r of `lane`.
__device__ __forceinline__ int mfma_row(int lane, int r) {
  return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
}

// Counts of one population for the 32x32 sub-tile (rb, cb) into LDS
// cnt[q][row][col] (q = n11, n10, n01, n00), int16.
__device__ __forceinline__ void gemm_hap_pop(const FusedParams& a, int p, int64_t rb, int64_t cb,
                                             int lane, int16_t* cnt, int roff, int coff) {
  const int ks = a.ksteps[p];
  const v4i* A = reinterpret_cast<const v4i*>(a.panels + rb * a.block_bytes + a.pop_off[p]) + lane;
  const v4i* B = reinterpret_cast<const v4i*>(a.panels + cb * a.block_bytes + a.pop_off[p]) + lane;
  const int col = lane & 31;
  int16_t* c1p = cnt;
  int16_t* c2p = cnt + TILE * TILE;
  int16_t* c3p = cnt + 2 * TILE * TILE;
  int16_t* c4p = cnt + 3 * TILE * TILE;
  if (a.missing[p]) {
    const v4i* Av = A + ks * 64;
    const v4i* Bv = B + ks * 64;
    v16i hh = zero16(), hv = zero16(), vh = zero16(), vv = zero16();
#pragma unroll 2
    for (int s = 0; s < ks; ++s) {
      const v4i ah = A[s * 64], av = Av[s * 64], bh = B[s * 64], bv = Bv[s * 64];
      hh = __builtin_amdgcn_mfma_i32_32x32x32_i8(ah, bh, hh, 0, 0, 0);
      hv = __builtin_amdgcn_mfma_i32_32x32x32_i8(ah, bv, hv, 0, 0, 0);
      vh = __builtin_amdgcn_mfma_i32_32x32x32_i8(av, bh, vh, 0, 0, 0);
      vv = __builtin_amdgcn_mfma_i32_32x32x32_i8(av, bv, vv, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int o = (roff + mfma_row(lane, r)) * TILE + coff + col;
      const int n11 = hh[r];
      c1p[o] = (int16_t)n11;
      c2p[o] = (int16_t)(hv[r] - n11);
      c3p[o] = (int16_t)(vh[r] - n11);
      c4p[o] = (int16_t)(vv[r] - hv[r] - vh[r] + n11);
    }
  } else {
    v16i hh = zero16();
#pragma unroll 4
    for (int s = 0; s < ks; ++s)
      hh = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[s * 64], B[s * 64], hh, 0, 0, 0);
    const int32_t* cnt1 = a.site_stats + (int64_t)p * 3 * a.V;
    const int64_t j = cb * 32 + col;
    const int kj = j < a.V ? cnt1[j] : 0;
    const int K = a.K[p];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int rr = mfma_row(lane, r);
      const int64_t i = rb * 32 + rr;
      const int ki = i < a.V ? cnt1[i] : 0;
      const int o = (roff + rr) * TILE + coff + col;
      const int n11 = hh[r];
      c1p[o] = (int16_t)n11;
      c2p[o] = (int16_t)(ki - n11);
      c3p[o] = (int16_t)(kj - n11);
      c4p[o] = (int16_t)(K - ki - kj + n11);
    }
  }
}

template <int NS>
__device__ __forceinline__ void flush_lane(unsigned long long* slimb, int bin, const double* acc,
                                           int count) {
  unsigned long long* dst = slimb + (int64_t)bin * (NS + 1) * 3;
#pragma unroll
  for (int q = 0; q < NS; ++q) {
    long long x2, x1, x0;
    pgld_to_limbs(acc[q], x2, x1, x0);
    atomicAdd(dst + 3 * q + 0, (unsigned long long)x0);
    atomicAdd(dst + 3 * q + 1, (unsigned long long)x1);
    atomicAdd(dst + 3 * q + 2, (unsigned long long)x2);
  }
  atomicAdd(dst + 3 * NS, (unsigned long long)count);
}

template <int NS>
__device__ __forceinline__ void flush_wave(unsigned long long* slimb, int lane, int cur,
                                           const double* acc, int count) {
  int bmax = cur;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) bmax = max(bmax, __shfl_xor(bmax, off, 64));
  const bool uniform = __all(cur < 0 || cur == bmax);
  if (bmax >= 0 && uniform) {
    const bool mine = (cur == bmax);
    unsigned long long* dst = slimb + (int64_t)bmax * (NS + 1) * 3;
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      const double v = pgld_wave_sum(mine ? acc[q] : 0.0);
      if (lane == 0) {
        long long x2, x1, x0;
        pgld_to_limbs(v, x2, x1, x0);
        atomicAdd(dst + 3 * q + 0, (unsigned long long)x0);
        atomicAdd(dst + 3 * q + 1, (unsigned long long)x1);
        atomicAdd(dst + 3 * q + 2, (unsigned long long)x2);
      }
    }
    const int c = pgld_wave_sum_int(mine ? count : 0);
    if (lane == 0) atomicAdd(dst + 3 * NS, (unsigned long long)c);
  } else if (cur >= 0) {
    flush_lane<NS>(slimb, cur, acc, count);
  }
}

// Adds the workgroup's LDS limbs to the global result (carry-normalised).
template <int NS>
__device__ __forceinline__ void flush_global(unsigned long long* slimb, unsigned long long* out,
                                             int n_bins) {
  for (int e = threadIdx.x; e < n_bins * (NS + 1); e += blockDim.x) {
    unsigned long long* L = slimb + 3 * e;
    if (e % (NS + 1) == NS) {
      if (L[0]) atomicAdd(out + 3 * e, L[0]);
      continue;
    }
    long long l0 = (long long)L[0], l1 = (long long)L[1], l2 = (long long)L[2];
    pgld_normalize_limbs(l2, l1, l0);
    if (l0) atomicAdd(out + 3 * e + 0, (unsigned long long)l0);
    if (l1) atomicAdd(out + 3 * e + 1, (unsigned long long)l1);
    if (l2) atomicAdd(out + 3 * e + 2, (unsigned long long)l2);
  }
}

template <int P>
__global__ void __launch_bounds__(256, 2) fused_hap_kernel(FusedParams a) {
  constexpr int NS = HapSet<P>::NS;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  int16_t* cnt = reinterpret_cast<int16_t*>(smem);  // [P][4][64][64]
  double* sedges = reinterpret_cast<double*>(smem + (size_t)P * 4 * TILE * TILE * 2);
  unsigned long long* slimb = reinterpret_cast<unsigned long long*>(sedges + PGLD_MAX_BINS + 8);
  const int nb = a.n_bins;
  const int t = threadIdx.x, lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  for (int e = t; e <= nb; e += blockDim.x) sedges[e] = a.edges[e];
  for (int e = t; e < nb * (NS + 1) * 3; e += blockDim.x) slimb[e] = 0ull;
  __syncthreads();

  // XCD-aware split: workgroups are dealt round-robin to the 8 XCDs, so give
  // each XCD a contiguous stretch of the (row-major) tile list; neighbouring
  // tiles share row panels, which then stay in that XCD's L2.
  const int nx = 8;
  const int xcd = blockIdx.x % nx;
  const int per = gridDim.x / nx;
  const int64_t seg_lo = a.n_tiles * xcd / nx, seg_hi = a.n_tiles * (xcd + 1) / nx;

  for (int64_t tile = seg_lo + blockIdx.x / nx; tile < seg_hi; tile += per) {
    const int64_t RB = a.tiles[2 * tile], CB = a.tiles[2 * tile + 1];
    {  // ---- counting on MFMA
      const int sr = wave >> 1, sc = wave & 1;
      const int64_t rb = 2 * RB + sr, cb = 2 * CB + sc;
      if (cb >= rb && rb * 32 < a.V && cb * 32 < a.V) {
#pragma unroll
        for (int p = 0; p < P; ++p)
          gemm_hap_pop(a, p, rb, cb, lane, cnt + (size_t)p * 4 * TILE * TILE, sr * 32, sc * 32);
      }
    }
    __syncthreads();
    {  // ---- estimator epilogue + binning
      const int64_t j = CB * TILE + lane;
      const bool jok = j < a.V;
      const double pos_j = jok ? a.pos[j] : 0.0;
      const double co_j = jok ? a.coord[j] : 0.0;
      int cur = -1, count = 0, guess = 0;
      double acc[NS];
#pragma unroll
      for (int q = 0; q < NS; ++q) acc[q] = 0.0;
      for (int s = 0; s < 16; ++s) {
        const int row = wave * 16 + s;
        const int64_t i = RB * TILE + row;
        bool ok = jok && (i < j);
        int b = -1;
        if (ok) {
          const double pi = a.pos[i];
          ok = pos_j <= pi + a.max_dist;
          const double d = a.coord_is_pos ? (pos_j - pi) : fabs(co_j - a.coord[i]);
          b = pgld_bin_scan(d, sedges, nb, guess, nb - 1);
          ok = ok && b >= 0 && b < nb;
          guess = b < 0 ? 0 : (b >= nb ? nb - 1 : b);
        }
        if (ok) {
          HapPop pp[P];
#pragma unroll
          for (int p = 0; p < P; ++p) {
            const int16_t* c = cnt + (size_t)p * 4 * TILE * TILE + row * TILE + lane;
            const int n11 = c[0], n10 = c[TILE * TILE], n01 = c[2 * TILE * TILE],
                      n00 = c[3 * TILE * TILE];
            pp[p] = hap_pop(n11, n10, n01, n00, n11 + n10 + n01 + n00, a.rtab);
          }
          double x[NS];
          HapSet<P>::eval(pp, x);
          if (b != cur) {
            if (cur >= 0) flush_lane<NS>(slimb, cur, acc, count);
            cur = b;
            count = 0;
#pragma unroll
            for (int q = 0; q < NS; ++q) acc[q] = 0.0;
          }
#pragma unroll
          for (int q = 0; q < NS; ++q) acc[q] += x[q];
          ++count;
        }
      }
      flush_wave<NS>(slimb, lane, cur, acc, count);
    }
    __syncthreads();
  }
  __syncthreads();
  flush_global<NS>(slimb, a.out, nb);
}

static int g_num_cus = 0;

static int num_cus() {
  if (g_num_cus == 0) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess)
      g_num_cus = v;
    if (g_num_cus <= 0) g_num_cus = 256;
  }
  return g_num_cus;
}

template <int P>
static int launch_fused_hap(const PanelLayout& L, const pgld_binned_args* g, hipStream_t s) {
  constexpr int NS = HapSet<P>::NS;
  FusedParams a;
  a.panels = g->d_panels;
  a.site_stats = g->d_site_stats;
  a.pos = g->d_pos;
  a.coord = g->d_coord ? g->d_coord : g->d_pos;
  a.coord_is_pos = g->coord_is_pos || g->d_coord == nullptr || g->d_coord == g->d_pos;
  a.max_dist = g->max_dist;
  a.edges = g->d_edges;
  a.n_bins = g->n_bins;
  a.rtab = g->d_rtab;
  a.tiles = g->d_tiles;
  a.n_tiles = g->n_tiles;
  a.out = reinterpret_cast<unsigned long long*>(g->d_out);
  a.V = L.n_sites;
  a.block_bytes = L.block_bytes;
  for (int p = 0; p < PGLD_MAX_POPS; ++p) {
    a.pop_off[p] = L.pop_off[p];
    a.ksteps[p] = L.ksteps[p];
    a.missing[p] = L.missing[p];
    a.K[p] = L.pop_size[p];
  }
  const size_t lds = (size_t)P * 4 * TILE * TILE * 2 + (PGLD_MAX_BINS + 8) * sizeof(double) +
                     (size_t)g->n_bins * (NS + 1) * 3 * sizeof(unsigned long long);
  PGLD_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(fused_hap_kernel<P>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  int grid = g->grid > 0 ? g->grid : 2 * num_cus();
  grid = ((grid + 7) / 8) * 8;
  const int64_t need = ((g->n_tiles + 7) / 8) * 8;
  if (grid > need) grid = (int)need;
  if (grid < 8) grid = 8;
  hipLaunchKernelGGL(fused_hap_kernel<P>, dim3(grid), dim3(256), lds, s, a);
  PGLD_LAUNCH_CHECK();
  return PGLD_OK;
}

int validate_binned(const pgld_panel_desc* desc, const pgld_binned_args* g, PanelLayout* L) {
  int rc = make_layout(desc, L);
  if (rc) return rc;
  if (!g) return fail(PGLD_EINVAL, "null binned args");
  if (g->n_bins < 1 || g->n_bins > PGLD_MAX_BINS)
    return fail(PGLD_EINVAL, "n_bins must be in [1, 64]");
  if (!g->d_pos || !g->d_edges || !g->d_rtab || !g->d_out)
    return fail(PGLD_EINVAL, "null device pointer in binned args");
  return PGLD_OK;
}

}  // namespace pgld

using namespace pgld;

extern "C" int pgld_n_stats(int32_t mode, int32_t n_pops) {
  if (n_pops < 1 || n_pops > 4) return -1;
  static const int hap[5] = {0, HAP_NSTATS_P1, HAP_NSTATS_P2, HAP_NSTATS_P3, HAP_NSTATS_P4};
  if (mode == PGLD_MODE_HAP) return hap[n_pops];
  return -1;
}

extern "C" int pgld_binned_fused(const pgld_panel_desc* desc, const pgld_binned_args* args,
                                 void* stream) {
  PanelLayout L;
  int rc = validate_binned(desc, args, &L);
  if (rc) return rc;
  if (!args->d_panels || !args->d_site_stats || (!args->d_tiles && args->n_tiles))
    return fail(PGLD_EINVAL, "fused kernel needs panels, site stats and a tile list");
  if (args->n_tiles == 0) return PGLD_OK;
  hipStream_t s = as_stream(stream);
  if (desc->mode == PGLD_MODE_HAP) {
    switch (desc->n_pops) {
      case 1: return launch_fused_hap<1>(L, args, s);
      case 2: return launch_fused_hap<2>(L, args, s);
      default: break;
    }
  }
  return fail(PGLD_EUNSUPPORTED, "fused kernel: this (mode, n_pops) is not built yet");
}
