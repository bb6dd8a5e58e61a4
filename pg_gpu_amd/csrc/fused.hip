// Fused binned two-locus LD for haplotype data on gfx950.
//
// Every wave owns whole 32 x 32 site tiles of the distance band (row block rb,
// column block cb >= rb) and processes each one end to end, with no
// workgroup barrier inside the tile loop:
//   1. counting -- v_mfma_i32_32x32x32_i8 multiplies the packed int8
//      indicator panels (exact int32 accumulation): H.H, H.V, V.H, V.V give
//      n11, n10, n01, n00 of the tile's 1024 pairs (only H.H when the
//      population has no missing calls; the marginals come from site
//      tallies).  The counts are packed to 16-bit pairs in registers.
//   2. epilogue -- lane l holds column j = l & 31 and 16 rows i of the tile
//      (the MFMA output layout).  It evaluates the unbiased DD/Dz/pi2
//      estimators in FP64 registers, reading 1/(n)_k from an LDS table, and
//      sums them per distance bin.  Tiles that lie inside the cutoff, off the
//      diagonal and within one bin take a branch-free path; the others test
//      each pair (i < j, cutoff, digitize) and keep a running sum per bin
//      (the bin only decreases along a lane's rows), flushed on change.
//   3. at the tile end the wave's FP64 sums are reduced across lanes and
//      added to the workgroup's LDS bins as 3-limb fixed point with integer
//      atomics; at exit each workgroup adds its limbs to the global result.
//      Integer addition is associative, so the bin sums are bitwise
//      reproducible whatever the schedule.
// With two waves per SIMD one wave's MFMA/loads overlap the other's FP64
// epilogue.  No pair list and no per-pair statistic ever reaches HBM.
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
  const double* rtab;          // 4 doubles per n in [0, max pop size]
  const int32_t* tiles;        // (rb, cb) pairs of 32-site blocks
  int64_t n_tiles;
  unsigned long long* out;
  int64_t V;
  int64_t block_bytes;
  int64_t pop_off[PGLD_MAX_POPS];
  int32_t ksteps[PGLD_MAX_POPS];
  int32_t missing[PGLD_MAX_POPS];
  int32_t K[PGLD_MAX_POPS];
  int32_t rt_lo[PGLD_MAX_POPS];   // smallest n kept in the LDS table of pop p
  int32_t rt_off[PGLD_MAX_POPS];  // first entry of pop p in the LDS table
  int32_t rt_len;                 // total LDS table entries
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

__device__ __forceinline__ v16i zero16() {
  v16i z;
#pragma unroll
  for (int q = 0; q < 16; ++q) z[q] = 0;
  return z;
}

// Row (within a 32x32 MFMA output block) of accumulator register r of `lane`;
// the column is lane & 31.
__device__ __forceinline__ int mfma_row(int lane, int r) {
  return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
}

// Counts of population p for tile (rb, cb), packed per slot as
// w0 = n11 | n10 << 16, w1 = n01 | n00 << 16.
__device__ __forceinline__ void tile_counts(const FusedParams& a, int p, int64_t rb, int64_t cb,
                                            int lane, uint32_t (&w0)[16], uint32_t (&w1)[16]) {
  const int ks = a.ksteps[p];
  const v4i* A = reinterpret_cast<const v4i*>(a.panels + rb * a.block_bytes + a.pop_off[p]) + lane;
  const v4i* B = reinterpret_cast<const v4i*>(a.panels + cb * a.block_bytes + a.pop_off[p]) + lane;
  if (a.missing[p]) {
    const v4i* Av = A + ks * 64;
    const v4i* Bv = B + ks * 64;
    v16i hh = zero16(), hv = zero16(), vh = zero16(), vv = zero16();
    // Operand loads run PF k-steps ahead of the MFMAs (a register ring), so
    // the L2 latency of one step overlaps the multiplies of the previous ones.
    constexpr int PF = 4;
    v4i ah[PF], av[PF], bh[PF], bv[PF];
#pragma unroll
    for (int d = 0; d < PF; ++d) {
      if (d < ks) {
        ah[d] = A[d * 64]; av[d] = Av[d * 64]; bh[d] = B[d * 64]; bv[d] = Bv[d * 64];
      }
    }
    for (int s0 = 0; s0 < ks; s0 += PF) {
#pragma unroll
      for (int d = 0; d < PF; ++d) {
        const int s = s0 + d;
        if (s < ks) {
          hh = __builtin_amdgcn_mfma_i32_32x32x32_i8(ah[d], bh[d], hh, 0, 0, 0);
          hv = __builtin_amdgcn_mfma_i32_32x32x32_i8(ah[d], bv[d], hv, 0, 0, 0);
          vh = __builtin_amdgcn_mfma_i32_32x32x32_i8(av[d], bh[d], vh, 0, 0, 0);
          vv = __builtin_amdgcn_mfma_i32_32x32x32_i8(av[d], bv[d], vv, 0, 0, 0);
          if (s + PF < ks) {
            const int n = (s + PF) * 64;
            ah[d] = A[n]; av[d] = Av[n]; bh[d] = B[n]; bv[d] = Bv[n];
          }
        }
      }
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const uint32_t n11 = hh[r];
      w0[r] = n11 | ((uint32_t)(hv[r] - hh[r]) << 16);
      w1[r] = (uint32_t)(vh[r] - hh[r]) | ((uint32_t)(vv[r] - hv[r] - vh[r] + hh[r]) << 16);
    }
  } else {
    v16i hh = zero16();
    constexpr int PF = 4;
    v4i ah[PF], bh[PF];
#pragma unroll
    for (int d = 0; d < PF; ++d)
      if (d < ks) { ah[d] = A[d * 64]; bh[d] = B[d * 64]; }
    for (int s0 = 0; s0 < ks; s0 += PF) {
#pragma unroll
      for (int d = 0; d < PF; ++d) {
        const int s = s0 + d;
        if (s < ks) {
          hh = __builtin_amdgcn_mfma_i32_32x32x32_i8(ah[d], bh[d], hh, 0, 0, 0);
          if (s + PF < ks) { ah[d] = A[(s + PF) * 64]; bh[d] = B[(s + PF) * 64]; }
        }
      }
    }
    const int32_t* cnt1 = a.site_stats + (int64_t)p * 3 * a.V;
    const int64_t j = cb * 32 + (lane & 31);
    const int kj = j < a.V ? cnt1[j] : 0;
    const int K = a.K[p];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int64_t i = rb * 32 + mfma_row(lane, r);
      const int ki = i < a.V ? cnt1[i] : 0;
      const int n11 = hh[r];
      w0[r] = (uint32_t)n11 | ((uint32_t)(ki - n11) << 16);
      w1[r] = (uint32_t)(kj - n11) | ((uint32_t)(K - ki - kj + n11) << 16);
    }
  }
}

template <int P>
__device__ __forceinline__ void pair_stats(const FusedParams& a, const double* srt,
                                           const uint32_t (&w0)[P][16],
                                           const uint32_t (&w1)[P][16], int r, double* x) {
  HapPop pp[P];
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int n11 = w0[p][r] & 0xffff, n10 = w0[p][r] >> 16;
    const int n01 = w1[p][r] & 0xffff, n00 = w1[p][r] >> 16;
    const int n = n11 + n10 + n01 + n00;
    int e = n - a.rt_lo[p];
    e = e < 0 ? 0 : e;
    pp[p] = hap_pop(n11, n10, n01, n00, e, srt + 4 * a.rt_off[p]);
  }
  HapSet<P>::eval(pp, x);
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

// Sum the lanes whose running bin is `bin` and add it once (lane 0).
template <int NS>
__device__ __forceinline__ void flush_wave_bin(unsigned long long* slimb, int lane, int bin,
                                               bool mine, const double* acc, int count) {
  unsigned long long* dst = slimb + (int64_t)bin * (NS + 1) * 3;
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
}

template <int NS>
__device__ __forceinline__ void flush_wave(unsigned long long* slimb, int lane, int cur,
                                           const double* acc, int count) {
  int bmax = cur;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) bmax = max(bmax, __shfl_xor(bmax, off, 64));
  const bool uniform = __all(cur < 0 || cur == bmax);
  if (bmax >= 0 && uniform)
    flush_wave_bin<NS>(slimb, lane, bmax, cur == bmax, acc, count);
  else if (cur >= 0)
    flush_lane<NS>(slimb, cur, acc, count);
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

__device__ __forceinline__ double pair_dist(const FusedParams& a, int64_t i, int64_t j) {
  return a.coord_is_pos ? (a.pos[j] - a.pos[i]) : fabs(a.coord[j] - a.coord[i]);
}

template <int P>
__global__ void __launch_bounds__(256, 2) fused_hap_kernel(FusedParams a) {
  constexpr int NS = HapSet<P>::NS;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  double* sedges = reinterpret_cast<double*>(smem);
  double* srt = sedges + PGLD_MAX_BINS + 8;
  unsigned long long* slimb = reinterpret_cast<unsigned long long*>(srt + 4 * a.rt_len);
  const int nb = a.n_bins;
  const int t = threadIdx.x, lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  for (int e = t; e <= nb; e += blockDim.x) sedges[e] = a.edges[e];
  for (int p = 0; p < P; ++p) {
    const int len = a.K[p] - a.rt_lo[p] + 1;
    for (int e = t; e < 4 * len; e += blockDim.x)
      srt[4 * a.rt_off[p] + e] = a.rtab[4 * a.rt_lo[p] + e];
  }
  for (int e = t; e < nb * (NS + 1) * 3; e += blockDim.x) slimb[e] = 0ull;
  __syncthreads();

  // XCD-aware split: workgroups are dealt round-robin to the 8 XCDs, so each
  // XCD gets a contiguous stretch of the (row-major) tile list; neighbouring
  // tiles share row panels, which then stay in that XCD's L2.
  const int nx = 8;
  const int xcd = blockIdx.x % nx;
  const int waves_per_xcd = (gridDim.x / nx) * 4;
  const int64_t seg_lo = a.n_tiles * xcd / nx, seg_hi = a.n_tiles * (xcd + 1) / nx;
  const int col = lane & 31;

  for (int64_t tile = seg_lo + (blockIdx.x / nx) * 4 + wave; tile < seg_hi;
       tile += waves_per_xcd) {
    const int64_t rb = a.tiles[2 * tile], cb = a.tiles[2 * tile + 1];
    uint32_t w0[P][16], w1[P][16];
#pragma unroll
    for (int p = 0; p < P; ++p) tile_counts(a, p, rb, cb, lane, w0[p], w1[p]);

    // ---- tile classification (wave-uniform)
    const int64_t i0 = rb * 32, j0 = cb * 32;
    const int64_t i1 = min(i0 + 31, a.V - 1), j1 = min(j0 + 31, a.V - 1);
    bool fast = (cb > rb) && (j1 == j0 + 31) && (a.pos[j1] <= a.pos[i0] + a.max_dist);
    int bfast = -1;
    if (fast) {
      const int b0 = pgld_bin_scan(pair_dist(a, i1, j0), sedges, nb, 0, nb - 1);
      const int b1 = pgld_bin_scan(pair_dist(a, i0, j1), sedges, nb, b0 < 0 ? 0 : b0, nb - 1);
      fast = (b0 == b1) && b0 >= 0 && b0 < nb;
      bfast = b0;
    }
    double acc[NS];
#pragma unroll
    for (int q = 0; q < NS; ++q) acc[q] = 0.0;
    if (fast) {
      // every pair of the tile is valid and falls in bin bfast
#pragma unroll 1
      for (int r = 0; r < 16; ++r) {
        double x[NS];
        pair_stats<P>(a, srt, w0, w1, r, x);
#pragma unroll
        for (int q = 0; q < NS; ++q) acc[q] += x[q];
      }
      flush_wave_bin<NS>(slimb, lane, bfast, true, acc, 16);
    } else {
      const int64_t j = j0 + col;
      const bool jok = j < a.V;
      const double pos_j = jok ? a.pos[j] : 0.0;
      const double co_j = jok ? a.coord[j] : 0.0;
      int cur = -1, count = 0, guess = nb - 1;
#pragma unroll 1
      for (int r = 0; r < 16; ++r) {
        const int64_t i = i0 + mfma_row(lane, r);
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
          double x[NS];
          pair_stats<P>(a, srt, w0, w1, r, x);
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
  int off = 0;
  for (int p = 0; p < PGLD_MAX_POPS; ++p) {
    a.pop_off[p] = L.pop_off[p];
    a.ksteps[p] = L.ksteps[p];
    a.missing[p] = L.missing[p];
    a.K[p] = L.pop_size[p];
    int lo = 0;
    if (p < P) {
      lo = L.missing[p] ? g->n_lo[p] : L.pop_size[p];
      if (lo < 0 || lo > L.pop_size[p]) return fail(PGLD_EINVAL, "n_lo out of range");
    }
    a.rt_lo[p] = lo;
    a.rt_off[p] = off;
    if (p < P) off += L.pop_size[p] - lo + 1;
  }
  a.rt_len = off;
  const size_t lds = (PGLD_MAX_BINS + 8) * sizeof(double) + (size_t)off * 4 * sizeof(double) +
                     (size_t)g->n_bins * (NS + 1) * 3 * sizeof(unsigned long long);
  if (lds > 160 * 1024) return fail(PGLD_EINVAL, "fused kernel: LDS tables exceed 160 KiB");
  PGLD_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(fused_hap_kernel<P>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  int grid = g->grid > 0 ? g->grid : 2 * num_cus();
  grid = ((grid + 7) / 8) * 8;
  const int64_t need = ((g->n_tiles + 31) / 32) * 8;  // at least 4 tiles per wave
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
