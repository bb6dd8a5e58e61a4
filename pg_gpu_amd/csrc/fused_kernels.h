// Kernel templates and launchers of the fused binned LD kernel, its
// counting-stage twin and the ragged-window kernel, instantiated per traits
// class in fused_hap.hip and fused_geno.hip.
#pragma once
#include <type_traits>
#include "fused_tiles.h"

namespace pgld {

// GTAB: the reciprocal table is read from global memory (a.rtab, indexed by
// n) instead of LDS, for populations whose range of valid-sample counts
// makes the LDS table too large; a.rt_len is then 0 and rt_lo / rt_off 0.
template <class T, bool GTAB = false>
__global__ void __launch_bounds__(256, T::MINB) fused_kernel(FusedParams a) {
  constexpr int NS = T::NSO, GS = T::GS;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  double* sedges = reinterpret_cast<double*>(smem);
  double* stab = sedges + PGLD_MAX_BINS + 8;
  const double* srt = GTAB ? a.rtab : stab;
  const int nb = a.n_bins;
  double* sstage = stab + 4 * a.rt_len;                      // [4 waves][nb][GS]
  unsigned long long* slimb =
      reinterpret_cast<unsigned long long*>(sstage + 4 * nb * GS);  // [nb][NS+1][3]
  int* scnt = reinterpret_cast<int*>(slimb + nb * (NS + 1) * 3);    // [4 waves][nb]
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  for (int e = tid; e <= nb; e += blockDim.x) sedges[e] = a.edges[e];
  if (!GTAB) {
    for (int p = 0; p < T::NP; ++p) {
      const int len = a.K[p] - a.rt_lo[p] + 1;
      for (int e = tid; e < 4 * len; e += blockDim.x)
        stab[4 * a.rt_off[p] + e] = a.rtab[4 * a.rt_lo[p] + e];
    }
  }
  for (int e = tid; e < 4 * nb * GS; e += blockDim.x) sstage[e] = 0.0;
  for (int e = tid; e < nb * (NS + 1) * 3; e += blockDim.x) slimb[e] = 0ull;
  for (int e = tid; e < 4 * nb; e += blockDim.x) scnt[e] = 0;
  __syncthreads();
  // one bin edge per lane for the tile classification (pgld_bin_ballot)
  const double lane_edge = lane <= nb ? sedges[lane] : __builtin_inf();
  const double e64 = nb == 64 ? sedges[64] : 0.0;
  TileCtx t;
  t.fp = &a;
  t.sedges = sedges;
  t.srt = srt;
  t.stg = sstage + wave * nb * GS;
  t.scw = scnt + wave * nb;
  t.slimb = slimb;
  t.lane = lane;

  // XCD-aware split: workgroups are dealt round-robin to the 8 XCDs, so each
  // XCD gets a contiguous stretch of the tile list (ordered in L2-sized
  // squares of tiles); the waves of one XCD then share operand panels in
  // that XCD's L2.  With a scheduler buffer, after its first tile a wave
  // takes the next unclaimed tile of its XCD's stretch from an atomic
  // counter, so the waves of an XCD stay within one square of tiles
  // (a static stride lets fast and slow waves drift apart and the panels in
  // flight outgrow the L2).  The bin sums do not depend on the assignment.
  const int nx = 8;
  const int xcd = blockIdx.x % nx;
  const int waves_per_xcd = (gridDim.x / nx) * 4;
  const int64_t seg_lo = a.n_tiles * xcd / nx, seg_hi = a.n_tiles * (xcd + 1) / nx;
  uint32_t* ctr = a.sched ? a.sched + xcd * 32 : nullptr;

  // The next tile is claimed at the start of the current one and its list
  // entry loaded after the current counts: neither round trip is waited for
  // at a tile start.
  int64_t local = (blockIdx.x / nx) * 4 + wave;
  int32_t nrb = 0, ncb = 0;
  if (seg_lo + local < seg_hi) {
    nrb = a.tiles[2 * (seg_lo + local)];
    ncb = a.tiles[2 * (seg_lo + local) + 1];
  }
  while (seg_lo + local < seg_hi) {
    const int64_t rb = __builtin_amdgcn_readfirstlane(nrb), cb = __builtin_amdgcn_readfirstlane(ncb);
    uint32_t claim = 0;
    if (ctr && lane == 0) claim = atomicAdd(ctr, 1u);  // the next tile
    // corner coordinates for the classification, loaded before the counts
    const int64_t i0 = rb * 32, j0 = cb * 32;
    const int64_t i1 = min(i0 + 31, a.V - 1), j1 = min(j0 + 31, a.V - 1);
    const double d_min = coord_diff(a, i1, j0), d_max = coord_diff(a, i0, j1);
    const bool in_cut = a.pos[j1] <= a.pos[i0] + a.max_dist;
    typename T::Counts cts;
    T::counts(a, rb, cb, lane, cts);
    const int64_t next = ctr ? waves_per_xcd + (int64_t)__builtin_amdgcn_readfirstlane(__shfl(claim, 0, 64))
                             : local + waves_per_xcd;
    if (seg_lo + next < seg_hi) {
      nrb = a.tiles[2 * (seg_lo + next)];
      ncb = a.tiles[2 * (seg_lo + next) + 1];
    }

    // ---- tile classification (wave-uniform).  Coordinates are
    // non-decreasing along the sites, so the corner pairs (i1, j0) and
    // (i0, j1) bound the bins of every pair in the tile.
    t.i0 = i0;
    t.j0 = j0;
    t.bmin = pgld_bin_ballot(d_min, lane_edge, nb, e64);
    t.bmax = pgld_bin_ballot(d_max, lane_edge, nb, e64);
    t.clean = (cb > rb) && (j1 == j0 + 31) && in_cut && t.bmin >= 0 && t.bmax < nb;
    t.fast = t.clean && t.bmin == t.bmax;
    t.blo = t.bmin < 0 ? 0 : t.bmin;
    t.dual = !t.fast && t.blo < nb && (t.bmax >= nb ? nb - 1 : t.bmax) - t.blo <= 1;
    const uint32_t none[4] = {0u, 0u, 0u, 0u};
    tile_groups<T, 0>(a, t, cts, none);
    local = next;
  }
  // nothing may stay staged
  stage_drain<T, 0, NS, T::template gsz<0>()>(a, srt, t.stg, t.scw, slimb, lane, 0, nb - 1);
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

// Kernel parameters and LDS plan of a binned job (shared by both forms of the
// fused kernel).  fixed: bytes of the bin edges and reciprocal table;
// per_bin: bytes per bin of the staging arrays (nst statistics waves), limbs
// and pair counters; gtab: the table is read from global memory.
template <class T>
static int fused_params(const PanelLayout& L, const pgld_binned_args* g, int nst, FusedParams& a,
                        size_t& fixed, size_t& per_bin, bool& gtab) {
  constexpr int NS = T::NSO;
  constexpr int P = T::NP;
  a = FusedParams{};
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
  a.sched = g->d_sched;
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
  constexpr size_t kLds = 160 * 1024;
  fixed = (PGLD_MAX_BINS + 8) * sizeof(double) + (size_t)off * 4 * sizeof(double);
  per_bin = (size_t)nst * T::GS * sizeof(double) +
            (size_t)(NS + 1) * 3 * sizeof(unsigned long long) + (size_t)nst * sizeof(int);
  // Populations whose valid-sample counts span too wide a range for the LDS
  // table (large populations with heavily missing sites) read the global
  // table instead: every entry of a.rtab, indexed by n directly.
  gtab = fixed + per_bin > kLds;
  if (gtab) {
    for (int p = 0; p < PGLD_MAX_POPS; ++p) a.rt_lo[p] = a.rt_off[p] = 0;
    a.rt_len = 0;
    fixed = (PGLD_MAX_BINS + 8) * sizeof(double);
  }
  return PGLD_OK;
}

template <class T>
static int launch_fused_waves(const PanelLayout& L, const pgld_binned_args* g, hipStream_t s) {
  constexpr int NS = T::NSO;
  constexpr size_t kLds = 160 * 1024;
  FusedParams a;
  size_t fixed = 0, per_bin = 0;
  bool gtab = false;
  int rc = fused_params<T>(L, g, 4, a, fixed, per_bin, gtab);
  if (rc) return rc;
  // Bins whose per-workgroup limbs do not fit in LDS together are covered by
  // consecutive launches over ranges of bins (each re-counts the band; pairs
  // outside a launch's edges are skipped by its binning).  Those launches use
  // the static tile schedule: the caller's scheduler words serve one launch.
  const int bins_per = (int)std::min<size_t>((kLds - fixed) / per_bin, (size_t)g->n_bins);
  const int n_launch = (g->n_bins + bins_per - 1) / bins_per;
  int grid = g->grid > 0 ? g->grid : T::MINB * num_cus();
  grid = ((grid + 7) / 8) * 8;
  const int64_t need = ((g->n_tiles + 31) / 32) * 8;  // at least 4 tiles per wave
  if (grid > need) grid = (int)need;
  if (grid < 8) grid = 8;
  for (int b0 = 0; b0 < g->n_bins; b0 += bins_per) {
    FusedParams c = a;
    c.n_bins = std::min(bins_per, g->n_bins - b0);
    c.edges = g->d_edges + b0;
    c.out = a.out + (int64_t)b0 * (NS + 1) * 3;
    if (n_launch > 1) c.sched = nullptr;
    const size_t lds = fixed + (size_t)c.n_bins * per_bin;
    const void* fn = gtab ? reinterpret_cast<const void*>(fused_kernel<T, true>)
                          : reinterpret_cast<const void*>(fused_kernel<T, false>);
    PGLD_HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    if (gtab)
      hipLaunchKernelGGL((fused_kernel<T, true>), dim3(grid), dim3(256), lds, s, c);
    else
      hipLaunchKernelGGL((fused_kernel<T, false>), dim3(grid), dim3(256), lds, s, c);
    PGLD_LAUNCH_CHECK();
  }
  return PGLD_OK;
}

// ---------------------------------------------------------------------------
// Producer/consumer form of the fused kernel: one workgroup of 4 * MINB waves
// per CU.  Waves 0-3 ("counting waves") claim tiles, run the MFMA counting,
// classify each tile and, for tiles that are not single-bin, compute its bin
// words (tile_bins); they hand the tile to a statistics wave on the same SIMD
// through an LDS mailbox and go on to the next tile.  Waves 4 .. 4 MINB - 1
// ("statistics waves") only run the FP64 estimators and the binning of the
// tile in their mailbox.  The operand-load latency of the counting then
// overlaps the statistics of the other waves instead of stalling them, and a
// statistics wave needs no MFMA accumulators or operand registers.
//
// Mailbox m (statistics wave 4 + m, fed by counting wave m % 4): a state word
// (0 empty, 1 full, 2 closed), the tile descriptor, and the counts in the
// traits' mailbox layout followed by 4 bin words per lane.  Waits are bounded:
// a wave that gives up adds 2^62 to bin 0's pair count (the host rejects such
// a result) and leaves its loop, so every wave reaches the final barrier.
// ---------------------------------------------------------------------------
struct PcDesc {
  int32_t rb, cb, bmin, bmax, flags, blo, pad0, pad1;
};

template <class T>
struct PcShape {
  static constexpr int NCONS = 4 * (T::MINB - 1);                 // statistics waves
  static constexpr int THREADS = 64 * 4 * T::MINB;
  static constexpr int MB_WORDS = T::NWORDS * 16 * 64 + 4 * 64;   // per mailbox
  static constexpr size_t BYTES = 256 + (size_t)NCONS * sizeof(PcDesc) +
                                  (size_t)NCONS * MB_WORDS * sizeof(uint32_t);
};

// Polls mailbox state `st` until it is non-zero (for_full) or zero; returns
// the state, or -1 after about 0.3 s of polling.
__device__ __forceinline__ int pc_wait(int* st, bool for_full) {
  for (int it = 0; it < (1 << 22); ++it) {
    const int v = __builtin_amdgcn_readfirstlane(
        __hip_atomic_load(st, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP));
    if (for_full ? v != 0 : v == 0) return v;
    __builtin_amdgcn_s_sleep(2);
  }
  return -1;
}

__device__ __forceinline__ void pc_set(int* st, int v) {
  __hip_atomic_store(st, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <class T, bool GTAB = false>
__global__ void __launch_bounds__(PcShape<T>::THREADS, 1) fused_pc_kernel(FusedParams a) {
  using S = PcShape<T>;
  constexpr int NS = T::NSO, GS = T::GS, NC = S::NCONS, KPER = T::MINB - 1;
  static_assert(T::MINB >= 2, "a statistics wave per counting wave at least");
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  double* sedges = reinterpret_cast<double*>(smem);
  double* stab = sedges + PGLD_MAX_BINS + 8;
  const double* srt = GTAB ? a.rtab : stab;
  const int nb = a.n_bins;
  double* sstage = stab + 4 * a.rt_len;                      // [NC][nb][GS]
  unsigned long long* slimb =
      reinterpret_cast<unsigned long long*>(sstage + NC * nb * GS);  // [nb][NS+1][3]
  int* scnt = reinterpret_cast<int*>(slimb + nb * (NS + 1) * 3);     // [NC][nb]
  // mailboxes from the next 256-byte boundary (offsets from smem, so the
  // compiler keeps LDS addressing)
  const size_t mb_off =
      ((size_t)(reinterpret_cast<unsigned char*>(scnt + NC * nb) - smem) + 255) & ~size_t(255);
  int* mstate = reinterpret_cast<int*>(smem + mb_off);                // [NC] (256 B)
  PcDesc* mdesc = reinterpret_cast<PcDesc*>(smem + mb_off + 256);    // [NC]
  uint32_t* mbox = reinterpret_cast<uint32_t*>(mdesc + NC);         // [NC][MB_WORDS]
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  for (int e = tid; e <= nb; e += blockDim.x) sedges[e] = a.edges[e];
  if (!GTAB) {
    for (int p = 0; p < T::NP; ++p) {
      const int len = a.K[p] - a.rt_lo[p] + 1;
      for (int e = tid; e < 4 * len; e += blockDim.x)
        stab[4 * a.rt_off[p] + e] = a.rtab[4 * a.rt_lo[p] + e];
    }
  }
  for (int e = tid; e < NC * nb * GS; e += blockDim.x) sstage[e] = 0.0;
  for (int e = tid; e < nb * (NS + 1) * 3; e += blockDim.x) slimb[e] = 0ull;
  for (int e = tid; e < NC * nb; e += blockDim.x) scnt[e] = 0;
  for (int e = tid; e < NC; e += blockDim.x) mstate[e] = 0;
  __syncthreads();
  bool ok = true;
  if (wave < 4) {
    // ---- counting wave: tiles of this XCD's segment, as in fused_kernel
    const int nx = 8;
    const int xcd = blockIdx.x % nx;
    const int waves_per_xcd = (gridDim.x / nx) * 4;
    const int64_t seg_lo = a.n_tiles * xcd / nx, seg_hi = a.n_tiles * (xcd + 1) / nx;
    uint32_t* ctr = a.sched ? a.sched + xcd * 32 : nullptr;
    int64_t local = (blockIdx.x / nx) * 4 + wave;
    int32_t nrb = 0, ncb = 0;
    if (seg_lo + local < seg_hi) {
      nrb = a.tiles[2 * (seg_lo + local)];
      ncb = a.tiles[2 * (seg_lo + local) + 1];
    }
    TileCtx t;
    t.fp = &a;
    t.sedges = sedges;
    t.lane = lane;
    const double lane_edge = lane <= nb ? sedges[lane] : __builtin_inf();
    const double e64 = nb == 64 ? sedges[64] : 0.0;
    int turn = 0;
    while (seg_lo + local < seg_hi) {
      const int64_t rb = __builtin_amdgcn_readfirstlane(nrb), cb = __builtin_amdgcn_readfirstlane(ncb);
      uint32_t claim = 0;
      if (ctr && lane == 0) claim = atomicAdd(ctr, 1u);
      const int64_t i0 = rb * 32, j0 = cb * 32;
      const int64_t i1 = min(i0 + 31, a.V - 1), j1 = min(j0 + 31, a.V - 1);
      const double d_min = coord_diff(a, i1, j0), d_max = coord_diff(a, i0, j1);
      const bool in_cut = a.pos[j1] <= a.pos[i0] + a.max_dist;
      typename T::Counts cts;
      T::counts_pc(a, rb, cb, lane, cts);
      const int64_t next = ctr ? waves_per_xcd + (int64_t)__builtin_amdgcn_readfirstlane(__shfl(claim, 0, 64))
                               : local + waves_per_xcd;
      if (seg_lo + next < seg_hi) {
        nrb = a.tiles[2 * (seg_lo + next)];
        ncb = a.tiles[2 * (seg_lo + next) + 1];
      }
      const int bmin = pgld_bin_ballot(d_min, lane_edge, nb, e64);
      const int bmax = pgld_bin_ballot(d_max, lane_edge, nb, e64);
      const bool fast = (cb > rb) && (j1 == j0 + 31) && in_cut && bmin == bmax && bmin >= 0 &&
                        bmin < nb;
      const int blo = bmin < 0 ? 0 : bmin;
      const bool dual = !fast && blo < nb && (bmax >= nb ? nb - 1 : bmax) - blo <= 1;
      uint32_t bw[4] = {0u, 0u, 0u, 0u};
      if (!fast) {
        t.i0 = i0;
        t.j0 = j0;
        t.bmin = bmin;
        t.bmax = bmax;
        t.clean = (cb > rb) && (j1 == j0 + 31) && in_cut && bmin >= 0 && bmax < nb;
        tile_bins<clean_bins<T>()>(a, t, bw);
      }
      // the first empty mailbox of this wave's statistics waves
      int m = -1;
      for (int it = 0; it < (1 << 22) && m < 0; ++it) {
#pragma unroll
        for (int k = 0; k < KPER; ++k) {
          const int mm = wave + 4 * ((turn + k) % KPER);
          if (m < 0 && __builtin_amdgcn_readfirstlane(__hip_atomic_load(
                           mstate + mm, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP)) == 0)
            m = mm;
        }
        if (m < 0) __builtin_amdgcn_s_sleep(2);
      }
      if (m < 0) { ok = false; break; }
      turn = (turn + 1) % KPER;
      uint32_t* mb = mbox + (size_t)m * S::MB_WORDS;
      T::put(cts, mb, lane);
#pragma unroll
      for (int q = 0; q < 4; ++q) mb[T::NWORDS * 16 * 64 + q * 64 + lane] = bw[q];
      if (lane == 0) {
        PcDesc d;
        d.rb = (int32_t)rb; d.cb = (int32_t)cb; d.bmin = bmin; d.bmax = bmax;
        d.flags = (fast ? 1 : 0) | (dual ? 2 : 0); d.blo = blo; d.pad0 = d.pad1 = 0;
        mdesc[m] = d;
      }
      pc_set(mstate + m, 1);
      local = next;
    }
    // close this wave's mailboxes once they are empty
#pragma unroll
    for (int k = 0; k < KPER; ++k) {
      const int mm = wave + 4 * k;
      if (ok && pc_wait(mstate + mm, false) < 0) ok = false;
      pc_set(mstate + mm, 2);
    }
  } else {
    // ---- statistics wave m
    const int m = wave - 4;
    TileCtx t;
    t.fp = &a;
    t.sedges = sedges;
    t.srt = srt;
    t.stg = sstage + m * nb * GS;
    t.scw = scnt + m * nb;
    t.slimb = slimb;
    t.lane = lane;
    const uint32_t* mb = mbox + (size_t)m * S::MB_WORDS;
    for (;;) {
      const int v = pc_wait(mstate + m, true);
      if (v < 0) { ok = false; break; }
      if (v == 2) break;
      const int rb = __builtin_amdgcn_readfirstlane(mdesc[m].rb);
      const int cb = __builtin_amdgcn_readfirstlane(mdesc[m].cb);
      const int bmin = __builtin_amdgcn_readfirstlane(mdesc[m].bmin);
      const int bmax = __builtin_amdgcn_readfirstlane(mdesc[m].bmax);
      const int flags = __builtin_amdgcn_readfirstlane(mdesc[m].flags);
      const int blo = __builtin_amdgcn_readfirstlane(mdesc[m].blo);
      typename T::Counts cts;
      T::get(cts, mb, lane);
      uint32_t bw[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) bw[q] = mb[T::NWORDS * 16 * 64 + q * 64 + lane];
      pc_set(mstate + m, 0);
      t.i0 = (int64_t)rb * 32;
      t.j0 = (int64_t)cb * 32;
      t.bmin = bmin;
      t.bmax = bmax;
      t.fast = flags & 1;
      t.dual = (flags & 2) != 0;
      t.blo = blo;
      tile_groups<T, 0, true>(a, t, cts, bw);
    }
    stage_drain<T, 0, NS, T::template gsz<0>()>(a, srt, t.stg, t.scw, slimb, lane, 0, nb - 1);
  }
  if (!ok && lane == 0) atomicAdd(a.out + (int64_t)NS * 3, 1ull << 62);
  __syncthreads();
  flush_global<NS>(slimb, a.out, nb);
}

// The producer/consumer form where its LDS plan fits in one launch (all bins
// together); otherwise the per-wave form.
template <class T>
static int launch_fused_pc(const PanelLayout& L, const pgld_binned_args* g, hipStream_t s) {
  using S = PcShape<T>;
  constexpr size_t kLds = 160 * 1024;
  FusedParams a;
  size_t fixed = 0, per_bin = 0;
  bool gtab = false;
  int rc = fused_params<T>(L, g, S::NCONS, a, fixed, per_bin, gtab);
  if (rc) return rc;
  const size_t lds = fixed + (size_t)g->n_bins * per_bin + 256 + S::BYTES;
  if (lds > kLds) return launch_fused_waves<T>(L, g, s);
  int grid = g->grid > 0 ? g->grid : num_cus();
  grid = ((grid + 7) / 8) * 8;
  // at most one workgroup per 16 tiles (4 per counting wave), a multiple of 8
  const int64_t need = (((g->n_tiles + 15) / 16) + 7) / 8 * 8;
  if (grid > need) grid = (int)need;
  if (grid < 8) grid = 8;
  const void* fn = gtab ? reinterpret_cast<const void*>(fused_pc_kernel<T, true>)
                        : reinterpret_cast<const void*>(fused_pc_kernel<T, false>);
  PGLD_HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  if (gtab)
    hipLaunchKernelGGL((fused_pc_kernel<T, true>), dim3(grid), dim3(S::THREADS), lds, s, a);
  else
    hipLaunchKernelGGL((fused_pc_kernel<T, false>), dim3(grid), dim3(S::THREADS), lds, s, a);
  PGLD_LAUNCH_CHECK();
  return PGLD_OK;
}

// form 1 (producer/consumer) for the traits that run two or more waves per
// SIMD; form 0, or any other traits, the per-wave kernel.
template <class T>
static int launch_fused(const PanelLayout& L, const pgld_binned_args* g, hipStream_t s) {
  if constexpr (T::MINB >= 2 && T::PC_OK) {
    if (g->form == 1) return launch_fused_pc<T>(L, g, s);
  }
  return launch_fused_waves<T>(L, g, s);
}

// ---------------------------------------------------------------------------
// The counting stage by itself: the same traits, tiles and schedule as
// fused_kernel, with the epilogue replaced by a checksum of the tile's count
// words -- per population the sums over all band tiles and slots of the two
// 16-bit halves of both words (haplotype traits: n11, pA, pB, n with
// marginal slot words; n11, n10, n01, n00 with class counts).  It times the
// MFMA counting stage alone and pins its counts at full size (each sum is a
// sum of dot products of per-block sample tallies).
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(256, T::MINB) count_band_kernel(FusedParams a,
                                                                  unsigned long long* sums) {
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  constexpr int P = T::NP;
  unsigned long long acc[P][4];
#pragma unroll
  for (int p = 0; p < P; ++p)
#pragma unroll
    for (int k = 0; k < 4; ++k) acc[p][k] = 0ull;
  const int nx = 8;
  const int xcd = blockIdx.x % nx;
  const int waves_per_xcd = (gridDim.x / nx) * 4;
  const int64_t seg_lo = a.n_tiles * xcd / nx, seg_hi = a.n_tiles * (xcd + 1) / nx;
  uint32_t* ctr = a.sched ? a.sched + xcd * 32 : nullptr;
  int64_t local = (blockIdx.x / nx) * 4 + wave;
  int32_t nrb = 0, ncb = 0;
  if (seg_lo + local < seg_hi) {
    nrb = a.tiles[2 * (seg_lo + local)];
    ncb = a.tiles[2 * (seg_lo + local) + 1];
  }
  while (seg_lo + local < seg_hi) {
    const int64_t rb = __builtin_amdgcn_readfirstlane(nrb), cb = __builtin_amdgcn_readfirstlane(ncb);
    uint32_t claim = 0;
    if (ctr && lane == 0) claim = atomicAdd(ctr, 1u);
    typename T::Counts cts;
    T::counts(a, rb, cb, lane, cts);
    const int64_t next = ctr ? waves_per_xcd + (int64_t)__builtin_amdgcn_readfirstlane(__shfl(claim, 0, 64))
                             : local + waves_per_xcd;
    if (seg_lo + next < seg_hi) {
      nrb = a.tiles[2 * (seg_lo + next)];
      ncb = a.tiles[2 * (seg_lo + next) + 1];
    }
#pragma unroll
    for (int p = 0; p < P; ++p) {
      if (a.K[p] < 4096) {   // 16 slots of 12-bit fields: the packed sums cannot carry
        uint32_t s0 = 0, s1 = 0;
#pragma unroll
        for (int r = 0; r < 16; ++r) { s0 += cts.w0[p][r]; s1 += cts.w1[p][r]; }
        acc[p][0] += s0 & 0xffffu; acc[p][1] += s0 >> 16;
        acc[p][2] += s1 & 0xffffu; acc[p][3] += s1 >> 16;
      } else {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          acc[p][0] += cts.w0[p][r] & 0xffffu; acc[p][1] += cts.w0[p][r] >> 16;
          acc[p][2] += cts.w1[p][r] & 0xffffu; acc[p][3] += cts.w1[p][r] >> 16;
        }
      }
    }
    local = next;
  }
#pragma unroll
  for (int p = 0; p < P; ++p)
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      unsigned long long v = acc[p][k];
      for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
      if (lane == 0 && v) atomicAdd(sums + 4 * p + k, v);
    }
}

template <class T>
static int launch_count_band(const PanelLayout& L, const pgld_binned_args* g,
                             unsigned long long* sums, hipStream_t s) {
  FusedParams a = FusedParams{};
  a.panels = g->d_panels;
  a.site_stats = g->d_site_stats;
  a.tiles = g->d_tiles;
  a.n_tiles = g->n_tiles;
  a.sched = g->d_sched;
  a.V = L.n_sites;
  a.block_bytes = L.block_bytes;
  for (int p = 0; p < PGLD_MAX_POPS; ++p) {
    a.pop_off[p] = L.pop_off[p];
    a.ksteps[p] = L.ksteps[p];
    a.missing[p] = L.missing[p];
    a.K[p] = L.pop_size[p];
  }
  int grid = g->grid > 0 ? g->grid : T::MINB * num_cus();
  grid = ((grid + 7) / 8) * 8;
  const int64_t need = ((g->n_tiles + 31) / 32) * 8;
  if (grid > need) grid = (int)need;
  if (grid < 8) grid = 8;
  hipLaunchKernelGGL(count_band_kernel<T>, dim3(grid), dim3(256), 0, s, a, sums);
  PGLD_LAUNCH_CHECK();
  return PGLD_OK;
}

// ---------------------------------------------------------------------------
// Ragged windows: every i < j pair inside each window [lo_w, hi_w) of sites,
// summed per window.  The work list holds (window, row block, column block)
// triples covering each window's upper triangle; a tile that straddles a
// window edge appears once per window it touches, masked to that window.
// ---------------------------------------------------------------------------
struct WindowParams {
  FusedParams f;           // panels, site stats, reciprocal table, V, layout
  const int32_t* work;     // [n_work][3]
  int64_t n_work;
  const int64_t* win;      // [n_windows][2]
  unsigned long long* out; // [n_windows][NSO + 1][3]
  int32_t group;           // work items per wave turn
  int32_t slot_off;        // byte offset of the window slots in dynamic LDS
};

// A workgroup's turn touches a few consecutive windows: the first
// PGLD_WIN_SLOTS of them collect the four waves' limbs in LDS,
// [turn in batch][slot][q < 16][limb], and go to the output once per batch
// of PGLD_WIN_BATCH turns.
#define PGLD_WIN_SLOTS 4
#define PGLD_WIN_BATCH 8
#define PGLD_WIN_SET_WORDS (PGLD_WIN_SLOTS * 16 * 3)
#define PGLD_WIN_SLOT_WORDS (PGLD_WIN_BATCH * PGLD_WIN_SET_WORDS)

// Adds one item's wave total `tot` to the lane's running limbs: statistic
// lanes add its fixed-point limbs, the count lane the exact pair count.  One
// unconditional add per limb (a conditional target would put the limbs in
// scratch memory).
template <class T, int NV, int LOG>
__device__ __forceinline__ void window_add_limbs(const FusedParams& a, const double* srt,
                                                 double tot, int lane, long long& L0,
                                                 long long& L1, long long& L2) {
  constexpr int GS = T::template gsz<0>();
  const int q = pgld_bitrev(lane & (NV - 1), LOG);
  const bool stat = lane < NV && q < GS;
  const bool cnt = lane < NV && q == GS;
  long long x2 = 0, x1 = 0, x0 = 0;
  if (stat)
    pgld_to_limbs(tot * T::template stat_mult<0>(a, srt, q), x2, x1, x0,
                  pgld_limb_scale(T::template scale_exp<0>(q)));
  L0 += x0 + (long long)(cnt ? tot : 0.0);
  L1 += x1;
  L2 += x2;
}

// One work item's sums over its pairs inside the window, reduced across the
// wave and added as fixed-point limbs to the lane's running totals: lane l <
// NV holds statistic q = bitreverse(l) (q < GS) or, at q == GS, the pair
// count (in L0).  Integer limbs sum exactly, so the totals do not depend on
// how items are grouped.
template <class T>
__device__ __forceinline__ void window_item(const WindowParams& w, const double* srt,
                                            const typename T::Counts& cts, int lane, bool full,
                                            int64_t i0, int64_t j0, int64_t lo, int64_t hi,
                                            long long& L0, long long& L1, long long& L2) {
  constexpr int GS = T::template gsz<0>();
  double acc[GS];
#pragma unroll
  for (int q = 0; q < GS; ++q) acc[q] = 0.0;
  int count = 0;
  const int64_t j = j0 + (lane & 31);
  // row i = ib + o with o = 8 (r >> 2) + (r & 3) wave-uniform: lo <= i < j
  // becomes lo_o <= o < j_o on 32-bit bounds clamped to [-1, 32]
  const int64_t ib = i0 + 4 * (lane >> 5);
  const int lo_o = (int)(lo - ib < -1 ? -1 : lo - ib > 32 ? 32 : lo - ib);
  const int j_o = (int)(j - ib < -1 ? -1 : j - ib > 32 ? 32 : j - ib);
  const bool jin = j < hi;
  if (full) {
    // every pair of the item is in the window: plain sums, the slot loop
    // unrolled (static slot indices)
#pragma unroll
    for (int r = 0; r < 16; ++r) T::template accum<0>(w.f, srt, cts, r, acc, 1.0);
    count = 16;
  } else {
    // two slots per iteration: two pairs' FP64 chains interleave
#pragma unroll 2
    for (int r = 0; r < 16; ++r) {
      const int o = 8 * (r >> 2) + (r & 3);
      // every slot's statistics are finite (padded sites have zero counts), so
      // slots outside the window take weight 0 instead of a divergent branch
      const bool in = o >= lo_o && o < j_o && jin;
      T::template accum<0>(w.f, srt, cts, r, acc, in ? 1.0 : 0.0);
      count += in;
    }
  }
  // butterfly reduce-scatter (see reduce_to_limbs); the pair count rides in
  // the first padding slot (an exact small integer)
  constexpr int NV = GS <= 2 ? 2 : GS <= 4 ? 4 : GS <= 8 ? 8 : GS <= 16 ? 16 : GS <= 32 ? 32 : 64;
  constexpr int LOG = NV == 2 ? 1 : NV == 4 ? 2 : NV == 8 ? 3 : NV == 16 ? 4 : NV == 32 ? 5 : 6;
  static_assert(GS < NV, "count slot");
  double v[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = q < GS ? acc[q] : (q == GS ? (double)count : 0.0);
  window_add_limbs<T, NV, LOG>(w.f, srt, pgld_wave_sums<NV>(v, lane), lane, L0, L1, L2);
}

// Adds a lane's running totals of window `win` to the output.
template <class T>
__device__ __forceinline__ void window_flush(const WindowParams& w, int64_t win, int lane,
                                             long long L0, long long L1, long long L2) {
  constexpr int NS = T::NSO;
  constexpr int GS = T::template gsz<0>();
  constexpr int NV = GS <= 2 ? 2 : GS <= 4 ? 4 : GS <= 8 ? 8 : GS <= 16 ? 16 : GS <= 32 ? 32 : 64;
  constexpr int LOG = NV == 2 ? 1 : NV == 4 ? 2 : NV == 8 ? 3 : NV == 16 ? 4 : NV == 32 ? 5 : 6;
  const int q = pgld_bitrev(lane & (NV - 1), LOG);
  unsigned long long* dst_w = w.out + win * (NS + 1) * 3;
  if (lane < NV && q < GS) {
    unsigned long long* dst = dst_w + (int64_t)T::template col<0>(q) * 3;
    if (L0) atomicAdd(dst + 0, (unsigned long long)L0);
    if (L1) atomicAdd(dst + 1, (unsigned long long)L1);
    if (L2) atomicAdd(dst + 2, (unsigned long long)L2);
  } else if (lane < NV && q == GS) {
    if (L0) atomicAdd(dst_w + (int64_t)NS * 3, (unsigned long long)L0);
  }
}

// Window kernels that combine their waves' limbs in LDS slots: 32-site
// haplotype items without missing calls run as fast either way and write 8x
// less with slots; items with missing calls, genotype items and 16-site
// items run 1-4 % faster with per-wave atomics.
template <class T>
constexpr bool window_slots() {
  return std::is_same<T, HapNMRawTraits<true>>::value;
}

// Adds a wave's limbs of window `win` to LDS slot `slot` of its workgroup
// turn's slot set, or straight to the output when the window has no slot.
// `slot` is wave-uniform.
template <class T, bool SLOTS>
__device__ __forceinline__ void window_put(const WindowParams& w, unsigned long long* slots,
                                           int64_t win, int64_t slot, int lane, long long L0,
                                           long long L1, long long L2) {
  constexpr int GS = T::template gsz<0>();
  constexpr int NV = GS <= 2 ? 2 : GS <= 4 ? 4 : GS <= 8 ? 8 : GS <= 16 ? 16 : GS <= 32 ? 32 : 64;
  constexpr int LOG = NV == 2 ? 1 : NV == 4 ? 2 : NV == 8 ? 3 : NV == 16 ? 4 : NV == 32 ? 5 : 6;
  static_assert(GS < 16, "a window slot holds 16 columns");
  if (!SLOTS || slot < 0 || slot >= PGLD_WIN_SLOTS) {
    window_flush<T>(w, win, lane, L0, L1, L2);
    return;
  }
  const int q = pgld_bitrev(lane & (NV - 1), LOG);
  if (lane < NV && q <= GS) {
    unsigned long long* d = slots + ((int)slot * 16 + q) * 3;
    if (L0) atomicAdd(d + 0, (unsigned long long)L0);
    if (L1) atomicAdd(d + 1, (unsigned long long)L1);
    if (L2) atomicAdd(d + 2, (unsigned long long)L2);
  }
}

// Work items of one XCD (workgroup b runs on XCD b % 8) are a contiguous
// eighth of the work list, cut into R * m turns of equal length to within one
// item (R waves on the XCD; m = the XCD's items / (R * group), rounded, at
// least 1); wave r takes turns r, r + R, ..., r + (m - 1) R, so every wave
// holds the same number of items to within m.  At any time the XCD's waves
// work on neighbouring items (the same windows' panels, held in that XCD's
// L2), and a wave keeps its window's totals in registers while its items
// stay in that window: global atomics once per (wave, turn, window) instead
// of per item.  Every bound is wave-uniform.
struct WinSched {
  int64_t lo, n, parts, R, r;
  int m;
  __device__ WinSched(int64_t n_work, int wave, int group) {
    constexpr int nx = 8;  // the grid is a multiple of 8
    const int64_t xcd = blockIdx.x % nx;
    R = (int64_t)(gridDim.x / nx) * 4;
    r = (int64_t)(blockIdx.x / nx) * 4 + wave;
    lo = n_work * xcd / nx;
    n = n_work * (xcd + 1) / nx - lo;
    const int64_t per = R * group;
    const int64_t mm = (n + per / 2) / per;
    m = (int)(mm < 1 ? 1 : mm);
    parts = R * m;
  }
  // items [k0, k1) of the wave's turn t
  __device__ void turn(int t, int64_t& k0, int64_t& k1) const {
    const int64_t p = (int64_t)t * R + r;
    k0 = lo + n * p / parts;
    k1 = lo + n * (p + 1) / parts;
  }
  // the first item of the workgroup's turn t (its four waves take
  // neighbouring turns r - wave, ..., r - wave + 3)
  __device__ int64_t block_first(int t, int wave) const {
    return lo + n * ((int64_t)t * R + r - wave) / parts;
  }
};

// Runs `item(k, win, L0, L1, L2)` over the wave's work items turn by turn.
// With SLOTS the limbs of each window go to the LDS slot set of the
// workgroup's turn (slot = window offset from the turn's first window), and
// the sets are added to the output and cleared after every PGLD_WIN_BATCH
// turns: global atomics once per (workgroup, turn, window), two barriers
// per batch.  The turn count is the same for the four waves (one XCD), so
// every wave reaches every barrier.  Without SLOTS each wave adds its limbs
// to the output once per (wave, turn, window).
template <class T, bool SLOTS, class Item>
__device__ __forceinline__ void window_turns(const WindowParams& w, unsigned long long* slots,
                                             int tid, int wave, Item item) {
  constexpr int NS = T::NSO;
  constexpr int GS = T::template gsz<0>();
  const int lane = tid & 63;
  const WinSched ws(w.n_work, wave, w.group);
  if constexpr (SLOTS) {
    for (int e = tid; e < PGLD_WIN_SLOT_WORDS; e += blockDim.x) slots[e] = 0;
    __syncthreads();
  }
  for (int t = 0; t < ws.m; ++t) {
    int64_t k0, k1;
    ws.turn(t, k0, k1);
    const int64_t b0 = ws.block_first(t, wave);
    const int64_t w0 = b0 < w.n_work ? (int64_t)__builtin_amdgcn_readfirstlane(w.work[3 * b0]) : 0;
    unsigned long long* set = slots + (t % PGLD_WIN_BATCH) * PGLD_WIN_SET_WORDS;
    int64_t cur = -1;
    long long L0 = 0, L1 = 0, L2 = 0;
    for (int64_t k = k0; k < k1; ++k) {
      const int64_t win = __builtin_amdgcn_readfirstlane(w.work[3 * k]);
      if (win != cur) {
        if (cur >= 0) window_put<T, SLOTS>(w, set, cur, cur - w0, lane, L0, L1, L2);
        cur = win;
        L0 = L1 = L2 = 0;
      }
      item(k, win, L0, L1, L2);
    }
    if (cur >= 0) window_put<T, SLOTS>(w, set, cur, cur - w0, lane, L0, L1, L2);
    if (SLOTS && (t % PGLD_WIN_BATCH == PGLD_WIN_BATCH - 1 || t == ws.m - 1)) {
      __syncthreads();
      const int t0 = t - t % PGLD_WIN_BATCH;
      for (int e = tid; e < (t - t0 + 1) * PGLD_WIN_SET_WORDS; e += blockDim.x) {
        const unsigned long long v = slots[e];
        if (v) {
          slots[e] = 0;
          const int j = e / PGLD_WIN_SET_WORDS, f = e % PGLD_WIN_SET_WORDS;
          const int slot = f / 48, q = (f / 3) % 16, limb = f % 3;
          // a set with a nonzero word belongs to a turn with items, whose
          // first item lies inside the work list
          const int64_t first = w.work[3 * ws.block_first(t0 + j, wave)];
          const int c = q < GS ? T::template col<0>(q) : NS;
          atomicAdd(w.out + ((first + slot) * (NS + 1) + c) * 3 + limb, v);
        }
      }
      __syncthreads();
    }
  }
}

template <class T, bool GTAB = false>
__global__ void __launch_bounds__(256, T::MINB) window_kernel(WindowParams w) {
  static_assert(T::NGT == 1, "window sums use one statistic pass");
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  double* stab = reinterpret_cast<double*>(smem);
  const double* srt = GTAB ? w.f.rtab : stab;
  const FusedParams& a = w.f;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  if (!GTAB) {
    for (int p = 0; p < T::NP; ++p) {
      const int len = a.K[p] - a.rt_lo[p] + 1;
      for (int e = tid; e < 4 * len; e += blockDim.x)
        stab[4 * a.rt_off[p] + e] = a.rtab[4 * a.rt_lo[p] + e];
    }
    __syncthreads();
  }
  unsigned long long* slots = reinterpret_cast<unsigned long long*>(smem + w.slot_off);
  window_turns<T, window_slots<T>()>(w, slots, tid, wave,
                  [&](int64_t k, int64_t win, long long& L0, long long& L1, long long& L2) {
    const int64_t rb = __builtin_amdgcn_readfirstlane(w.work[3 * k + 1]);
    const int64_t cb = __builtin_amdgcn_readfirstlane(w.work[3 * k + 2]);
    const int64_t lo = w.win[2 * win], hi = w.win[2 * win + 1];
    typename T::Counts cts;
    T::counts(a, rb, cb, lane, cts);
    const int64_t i0 = rb * 32, j0 = cb * 32;
    const bool full = cb > rb && i0 >= lo && j0 + 32 <= hi;
    window_item<T>(w, srt, cts, lane, full, i0, j0, lo, hi, L0, L1, L2);
  });
}

// ---------------------------------------------------------------------------
// 16 x 16 site tiles for the window kernel (haplotypes, one population):
// v_mfma_scale_f32_16x16x128_f8f6f4 reads the same packed panels.  Its
// operand for 16 sites (the half h of a 32-site block) and 128 samples (two
// 64-sample k-steps) gives lane l site 16 h + (l & 15) and samples
// 128 t + 32 (l >> 4) + [0, 32): the 16 bytes that lane
// 16 h + (l & 15) + 32 ((l >> 4) & 1) of k-step 2 t + (l >> 5) holds.  Its
// 16 x 16 result gives lane l column l & 15 and rows 4 (l >> 4) + r, r < 4.
// Quarter tiles cover a window's triangle with less padding where windows
// are short; which side runs faster on a job is measured (autotune).
// ---------------------------------------------------------------------------
typedef float v4f __attribute__((ext_vector_type(4)));
typedef uint32_t v4u __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4f pgld_mfma16_fp4(v4i a, v4i b, v4f c) {
  const v8i a8 = {a[0], a[1], a[2], a[3], 0, 0, 0, 0};
  const v8i b8 = {b[0], b[1], b[2], b[3], 0, 0, 0, 0};
  return __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a8, b8, c, 4, 4, 0, 127, 0, 127);
}

// Lane's operand bytes of 128-sample step t for the 16 sites of half h; f
// points at the population's feature start in the 32-site block (k-step 0);
// k-steps past the population's last are zero.
__device__ __forceinline__ v4i frag16(const int8_t* f, int t, int h, int ks, int lane) {
  const int s = 2 * t + (lane >> 5);
  const int lp = 16 * h + (lane & 15) + 32 * ((lane >> 4) & 1);
  v4i v = {0, 0, 0, 0};
  if (s < ks) v = *reinterpret_cast<const v4i*>(f + (int64_t)s * PGLD_FRAG_BYTES + lp * 16);
  return v;
}

// Class counts of the population for the 16-site blocks (rb, cb), packed per
// slot as w0 = n11 | n10 << 16, w1 = n01 | n00 << 16.
__device__ __forceinline__ void tile16_counts(const FusedParams& a, int64_t rb, int64_t cb,
                                              int lane, v4u& w0, v4u& w1) {
  const int ks = a.ksteps[0];
  const int nt = (ks + 1) / 2;
  const int8_t* A = a.panels + (rb >> 1) * a.block_bytes + a.pop_off[0];
  const int8_t* B = a.panels + (cb >> 1) * a.block_bytes + a.pop_off[0];
  const int ha = (int)(rb & 1), hb = (int)(cb & 1);
  const int64_t fs = (int64_t)ks * PGLD_FRAG_BYTES;   // feature stride
  v4f hh = {0, 0, 0, 0};
  if (a.missing[0]) {
    v4f hv = hh, vh = hh, vv = hh;
    for (int t = 0; t < nt; ++t) {
      const v4i ah = frag16(A, t, ha, ks, lane), av = frag16(A + fs, t, ha, ks, lane);
      const v4i bh = frag16(B, t, hb, ks, lane), bv = frag16(B + fs, t, hb, ks, lane);
      hh = pgld_mfma16_fp4(ah, bh, hh);
      hv = pgld_mfma16_fp4(ah, bv, hv);
      vh = pgld_mfma16_fp4(av, bh, vh);
      vv = pgld_mfma16_fp4(av, bv, vv);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int n11 = (int)hh[r], h_v = (int)hv[r], v_h = (int)vh[r], v_v = (int)vv[r];
      w0[r] = (uint32_t)n11 | ((uint32_t)(h_v - n11) << 16);
      w1[r] = (uint32_t)(v_h - n11) | ((uint32_t)(v_v - h_v - v_h + n11) << 16);
    }
  } else {
    for (int t = 0; t < nt; ++t)
      hh = pgld_mfma16_fp4(frag16(A, t, ha, ks, lane), frag16(B, t, hb, ks, lane), hh);
    const int32_t* cnt1 = a.site_stats;
    const int64_t j = cb * 16 + (lane & 15);
    const int kj = j < a.V ? cnt1[j] : 0;
    const int K = a.K[0];
    const int64_t ib = rb * 16 + 4 * (lane >> 4);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int ki = ib + r < a.V ? cnt1[ib + r] : 0;
      const int n11 = (int)hh[r];
      w0[r] = (uint32_t)n11 | ((uint32_t)(ki - n11) << 16);
      w1[r] = (uint32_t)(kj - n11) | ((uint32_t)(K - ki - kj + n11) << 16);
    }
  }
}

// One population's haplotype statistics and r^2 per slot of a quarter tile
// (HapTraits<1, true>'s statistic pass over 4 slots).
struct Hap1Tile16 : HapTraits<1, true> {
  struct Counts { v4u w0, w1; };
  __device__ __forceinline__ static void accum(const FusedParams& a, const double* srt,
                                               const Counts& c, int r, double* acc, double wt) {
    const uint32_t w0 = c.w0[r], w1 = c.w1[r];
    const int n11 = w0 & 0xffff, n10 = w0 >> 16, n01 = w1 & 0xffff, n00 = w1 >> 16;
    HapPop pp[1];
    pp[0] = hap_pop(n11, n10, n01, n00, 0, rt_entry(a, srt, 0, n11 + n10 + n01 + n00));
    hap_acc_group<1, 0>(pp, acc, wt);
    double x[2];
    hap_r2(pp[0], x);
    acc[GS - 2] = fma(x[0], wt, acc[GS - 2]);
    acc[GS - 1] = fma(x[1], wt, acc[GS - 1]);
  }
};

__device__ __forceinline__ void window_item16(const WindowParams& w, const double* srt,
                                              const Hap1Tile16::Counts& cts, int lane, bool full,
                                              int64_t i0, int64_t j0, int64_t lo, int64_t hi,
                                              long long& L0, long long& L1, long long& L2) {
  using T = Hap1Tile16;
  constexpr int GS = T::template gsz<0>();
  double acc[GS];
#pragma unroll
  for (int q = 0; q < GS; ++q) acc[q] = 0.0;
  int count = 0;
  const int64_t j = j0 + (lane & 15);
  const int64_t ib = i0 + 4 * (lane >> 4);
  const bool jin = j < hi;
  if (full) {
#pragma unroll
    for (int r = 0; r < 4; ++r) T::accum(w.f, srt, cts, r, acc, 1.0);
    count = 4;
  } else {
#pragma unroll 2
    for (int r = 0; r < 4; ++r) {
      const int64_t i = ib + r;
      const bool in = i >= lo && i < j && jin;
      T::accum(w.f, srt, cts, r, acc, in ? 1.0 : 0.0);
      count += in;
    }
  }
  constexpr int NV = 8;
  static_assert(GS < NV, "count slot");
  double v[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = q < GS ? acc[q] : (q == GS ? (double)count : 0.0);
  window_add_limbs<T, NV, 3>(w.f, srt, pgld_wave_sums<NV>(v, lane), lane, L0, L1, L2);
}

// The window kernel over 16-site work items (blocks of 16 sites in d_work).
template <bool GTAB>
__global__ void __launch_bounds__(256, 3) window16_kernel(WindowParams w) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  double* stab = reinterpret_cast<double*>(smem);
  const double* srt = GTAB ? w.f.rtab : stab;
  const FusedParams& a = w.f;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  if (!GTAB) {
    const int len = a.K[0] - a.rt_lo[0] + 1;
    for (int e = tid; e < 4 * len; e += blockDim.x) stab[e] = a.rtab[4 * a.rt_lo[0] + e];
    __syncthreads();
  }
  unsigned long long* slots = reinterpret_cast<unsigned long long*>(smem + w.slot_off);
  window_turns<Hap1Tile16, false>(w, slots, tid, wave,
                           [&](int64_t k, int64_t win, long long& L0, long long& L1,
                               long long& L2) {
    const int64_t rb = __builtin_amdgcn_readfirstlane(w.work[3 * k + 1]);
    const int64_t cb = __builtin_amdgcn_readfirstlane(w.work[3 * k + 2]);
    const int64_t lo = w.win[2 * win], hi = w.win[2 * win + 1];
    Hap1Tile16::Counts cts;
    tile16_counts(a, rb, cb, lane, cts.w0, cts.w1);
    const int64_t i0 = rb * 16, j0 = cb * 16;
    const bool full = cb > rb && i0 >= lo && j0 + 16 <= hi;
    window_item16(w, srt, cts, lane, full, i0, j0, lo, hi, L0, L1, L2);
  });
}

template <class T>
static int launch_windows(const PanelLayout& L, const pgld_window_args* g, hipStream_t s) {
  WindowParams w;
  FusedParams& a = w.f;
  a = FusedParams{};
  a.panels = g->d_panels;
  a.site_stats = g->d_site_stats;
  a.rtab = g->d_rtab;
  a.V = L.n_sites;
  a.block_bytes = L.block_bytes;
  a.pop_off[0] = L.pop_off[0];
  a.ksteps[0] = L.ksteps[0];
  a.missing[0] = L.missing[0];
  a.K[0] = L.pop_size[0];
  const int lo = L.missing[0] ? g->n_lo[0] : L.pop_size[0];
  if (lo < 0 || lo > L.pop_size[0]) return fail(PGLD_EINVAL, "n_lo out of range");
  a.rt_lo[0] = lo;
  a.rt_off[0] = 0;
  a.rt_len = L.pop_size[0] - lo + 1;
  a.sched = g->d_sched;
  w.work = g->d_work;
  w.n_work = g->n_work;
  w.win = g->d_windows;
  w.out = reinterpret_cast<unsigned long long*>(g->d_out);
  // items per wave turn: an XCD's R waves work on about R * group
  // neighbouring items at a time, whose panels stay in that XCD's L2 at 16
  // items per turn (longer turns flush fewer atomics but re-read panels)
  w.group = g->group > 0 ? g->group : (g->tile == 16 ? 32 : 16);
  const size_t slot_bytes =
      window_slots<T>() && g->tile != 16 ? PGLD_WIN_SLOT_WORDS * sizeof(unsigned long long) : 0;
  size_t tab = ((size_t)a.rt_len * 4 * sizeof(double) + 15) / 16 * 16;
  // a table that leaves no room for the window slots in the LDS (a large
  // population with heavily missing sites) is read from global memory,
  // indexed by n
  const bool gtab = tab + slot_bytes > 160 * 1024;
  if (gtab) {
    a.rt_lo[0] = 0;
    a.rt_len = 0;
    tab = 0;
  }
  w.slot_off = (int32_t)tab;
  const size_t lds = tab + slot_bytes;
  if (gtab)
    PGLD_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(window_kernel<T, true>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  else
    PGLD_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(window_kernel<T, false>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  int grid = g->grid > 0 ? g->grid : T::MINB * num_cus();
  grid = ((grid + 7) / 8) * 8;
  const int64_t need = ((g->n_work + 31) / 32) * 8;
  if (grid > need) grid = (int)need;
  if (grid < 8) grid = 8;
  if (g->tile == 16) {
    // the 16-site kernel covers haplotypes with and without missing calls
    if constexpr (std::is_same<T, HapTraits<1, true>>::value ||
                  std::is_same<T, HapNMTraits<true>>::value ||
                  std::is_same<T, HapNMRawTraits<true>>::value) {
      PGLD_HIP_CHECK(hipFuncSetAttribute(
          gtab ? reinterpret_cast<const void*>(window16_kernel<true>)
               : reinterpret_cast<const void*>(window16_kernel<false>),
          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      if (gtab)
        hipLaunchKernelGGL((window16_kernel<true>), dim3(grid), dim3(256), lds, s, w);
      else
        hipLaunchKernelGGL((window16_kernel<false>), dim3(grid), dim3(256), lds, s, w);
      PGLD_LAUNCH_CHECK();
      return PGLD_OK;
    } else {
      return fail(PGLD_EUNSUPPORTED, "16-site window tiles are built for haplotypes");
    }
  }
  if (gtab)
    hipLaunchKernelGGL((window_kernel<T, true>), dim3(grid), dim3(256), lds, s, w);
  else
    hipLaunchKernelGGL((window_kernel<T, false>), dim3(grid), dim3(256), lds, s, w);
  PGLD_LAUNCH_CHECK();
  return PGLD_OK;
}

}  // namespace pgld
