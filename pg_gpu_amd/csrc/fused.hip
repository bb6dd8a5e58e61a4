// Entry points of the fused binned two-locus LD kernels.  The kernels
// (fused_kernels.h, fused_tiles.h; instantiated in fused_hap.hip and
// fused_geno.hip) work as follows.
//
// Every wave owns whole 32 x 32 site tiles of the distance band (row block rb,
// column block cb >= rb) and processes each one end to end, with no
// workgroup barrier inside the tile loop:
//   1. counting -- v_mfma_scale_f32_32x32x64_f8f6f4 multiplies the packed
//      4-bit (E2M1) indicator panels, 64 samples per step, with FP32
//      accumulation that is exact for counts below 2^24: H.H, H.V, V.H, V.V give
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
// With two waves per SIMD (haplotypes, and genotypes of one population) one
// wave's MFMA/loads overlap the other's FP64 epilogue; two-population
// genotype sets need the full register file of a one-wave-per-SIMD launch.  No pair list and no per-pair statistic ever reaches HBM.
#include <mutex>
#include "fused_tiles.h"

namespace pgld {

int validate_binned(const pgld_panel_desc* desc, const pgld_binned_args* g, PanelLayout* L) {
  int rc = make_layout(desc, L);
  if (rc) return rc;
  if (!g) return fail(PGLD_EINVAL, "null binned args");
  if (g->n_bins < 1 || g->n_bins > PGLD_MAX_BINS)
    return fail(PGLD_EINVAL, "n_bins must be in [1, 64]");
  if (!g->d_pos || !g->d_edges || !g->d_rtab || !g->d_out)
    return fail(PGLD_EINVAL, "null device pointer in binned args");
  if (g->with_r2 != 0 && g->with_r2 != 1) return fail(PGLD_EINVAL, "with_r2 must be 0 or 1");
  if (g->n_tiles < 0 || g->n_tiles > 0x7fffffffLL)
    return fail(PGLD_EINVAL, "n_tiles must be in [0, 2^31)");
  if (g->with_r2 && desc->n_pops != 1)
    return fail(PGLD_EUNSUPPORTED, "with_r2 is built for one population");
  return PGLD_OK;
}

}  // namespace pgld

using namespace pgld;

extern "C" int pgld_n_stats(int32_t mode, int32_t n_pops) {
  if (n_pops < 1 || n_pops > 4) return -1;
  static const int hap[5] = {0, HAP_NSTATS_P1, HAP_NSTATS_P2, HAP_NSTATS_P3, HAP_NSTATS_P4};
  static const int geno[5] = {0, GENO_NSTATS_P1, GENO_NSTATS_P2, GENO_NSTATS_P3, GENO_NSTATS_P4};
  if (mode == PGLD_MODE_HAP) return hap[n_pops];
  if (mode == PGLD_MODE_GENO) return geno[n_pops];
  return -1;
}

static int binned_fused_dispatch(const pgld_panel_desc* desc, const pgld_binned_args* args,
                                 const PanelLayout& L, hipStream_t s) {
  return desc->mode == PGLD_MODE_HAP ? launch_binned_hap(L, args, s)
                                     : launch_binned_geno(L, args, s);
}

extern "C" int pgld_binned_fused(const pgld_panel_desc* desc, const pgld_binned_args* args,
                                 void* stream) {
  PanelLayout L;
  int rc = validate_binned(desc, args, &L);
  if (rc) return rc;
  if (!args->d_panels || !args->d_site_stats || (!args->d_tiles && args->n_tiles))
    return fail(PGLD_EINVAL, "fused kernel needs panels, site stats and a tile list");
  if (args->n_tiles == 0) return PGLD_OK;
  return binned_fused_dispatch(desc, args, L, as_stream(stream));
}

// ---------------------------------------------------------------------------
// pgld_binned_fused_noalloc: the same launch with the tile-scheduler words in
// a file-scope device array instead of a caller buffer, and the output and
// scheduler cleared by a fill kernel, so a caller that cannot allocate (code
// under graph capture, serving loops) needs no device memory of its own
// beyond inputs and outputs.  Each device holds its own copy of `scratch`
// for the life of the process; its address is looked up once per device.
// Calls that overlap in time on one device share it and race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) uint32_t scratch[PGLD_SCHED_WORDS];
constexpr int kScratchMaxDevices = 64;
static uint32_t* g_scratch_ptr[kScratchMaxDevices];
static std::mutex g_scratch_mu;

static int scratch_words(uint32_t** out) {
  int dev = 0;
  PGLD_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kScratchMaxDevices) return fail(PGLD_EDEVICE, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    PGLD_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<uint32_t*>(p);
  }
  *out = g_scratch_ptr[dev];
  return PGLD_OK;
}

__global__ void __launch_bounds__(256) clear_words_kernel(uint32_t* __restrict__ a, int64_t n) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    a[e] = 0u;
}

static int clear_words(void* p, int64_t n_words, hipStream_t s) {
  if (n_words <= 0) return PGLD_OK;
  const int64_t blocks = std::min<int64_t>((n_words + 255) / 256, 1024);
  hipLaunchKernelGGL(clear_words_kernel, dim3((unsigned)blocks), dim3(256), 0, s,
                     static_cast<uint32_t*>(p), n_words);
  PGLD_LAUNCH_CHECK();
  return PGLD_OK;
}

extern "C" int pgld_binned_fused_noalloc(const pgld_panel_desc* desc,
                                         const pgld_binned_args* args, void* stream) {
  static_assert(sizeof(scratch) == PGLD_SCHED_WORDS * sizeof(uint32_t), "scheduler words");
  PanelLayout L;
  int rc = validate_binned(desc, args, &L);
  if (rc) return rc;
  if (!args->d_panels || !args->d_site_stats || (!args->d_tiles && args->n_tiles))
    return fail(PGLD_EINVAL, "fused kernel needs panels, site stats and a tile list");
  if (args->d_sched)
    return fail(PGLD_EINVAL, "pgld_binned_fused_noalloc uses its own scheduler words: "
                             "d_sched must be NULL");
  const int n_out = pgld_n_stats(desc->mode, desc->n_pops) + (args->with_r2 ? 2 : 0);
  if (n_out < 0) return fail(PGLD_EUNSUPPORTED, "statistic set not available");
  hipStream_t s = as_stream(stream);
  // the output: n_bins x (n_out + 1) x 3 int64 limbs = 6 words each
  rc = clear_words(args->d_out, (int64_t)args->n_bins * (n_out + 1) * 6, s);
  if (rc) return rc;
  if (args->n_tiles == 0) return PGLD_OK;
  uint32_t* sched = nullptr;
  rc = scratch_words(&sched);
  if (rc) return rc;
  rc = clear_words(sched, PGLD_SCHED_WORDS, s);
  if (rc) return rc;
  pgld_binned_args a = *args;
  a.d_sched = sched;
  return binned_fused_dispatch(desc, &a, L, s);
}

extern "C" int pgld_count_band(const pgld_panel_desc* desc, const pgld_binned_args* args,
                               unsigned long long* d_sums, void* stream) {
  PanelLayout L;
  int rc = make_layout(desc, &L);
  if (rc) return rc;
  if (!args) return fail(PGLD_EINVAL, "null binned args");
  if (desc->mode != PGLD_MODE_HAP)
    return fail(PGLD_EUNSUPPORTED, "the counting-stage kernel is built for haplotypes");
  if (!d_sums || !args->d_panels || !args->d_site_stats || (!args->d_tiles && args->n_tiles))
    return fail(PGLD_EINVAL, "count band needs panels, site stats, a tile list and d_sums");
  if (args->n_tiles < 0 || args->n_tiles > 0x7fffffffLL)
    return fail(PGLD_EINVAL, "n_tiles must be in [0, 2^31)");
  if (args->n_tiles == 0) return PGLD_OK;
  hipStream_t s = as_stream(stream);
  return launch_count_band_hap(L, args, d_sums, s);
}

extern "C" int pgld_windows_fused(const pgld_panel_desc* desc, const pgld_window_args* args,
                                  void* stream) {
  PanelLayout L;
  int rc = make_layout(desc, &L);
  if (rc) return rc;
  if (!args) return fail(PGLD_EINVAL, "null window args");
  if (desc->n_pops != 1) return fail(PGLD_EUNSUPPORTED, "window sums are built for one population");
  if (args->n_work < 0 || args->n_windows < 0) return fail(PGLD_EINVAL, "negative sizes");
  if (args->tile != 0 && args->tile != 16 && args->tile != 32)
    return fail(PGLD_EINVAL, "tile must be 16 or 32");
  if (args->group < 0 || args->group > 4096)
    return fail(PGLD_EINVAL, "group must be in [0, 4096]");
  if (args->n_work == 0) return PGLD_OK;
  if (!args->d_panels || !args->d_site_stats || !args->d_rtab || !args->d_work ||
      !args->d_windows || !args->d_out)
    return fail(PGLD_EINVAL, "null device pointer in window args");
  hipStream_t s = as_stream(stream);
  if (desc->mode == PGLD_MODE_HAP) return launch_windows_hap(L, args, s);
  return launch_windows_geno(L, args, s);
}

