// Site tallies and packing of sample-major int8 matrices into MFMA panels.
//
// Panel format: 0/1 indicator features as 4-bit E2M1 values (1.0 = 0x2, 0).
// For 32-site block b, population p, feature f and 64-sample k-step s there
// is one 1 KiB fragment at
//   b * block_bytes + pop_off[p] + (f * ksteps[p] + s) * 1024
// holding, for lane l (0..63), byte e (0..15) and nibble h (0 low, 1 high),
//   site  = 32 b + (l & 31),   sample = 64 s + 32 (l >> 5) + 2 e + h
// which is the operand layout a wave feeds to v_mfma_scale_f32_32x32x64_f8f6f4
// (FP4 operands) both as the A (row sites) and the B (column sites) matrix.
// Samples past the population size and sites past n_sites are zero in every
// feature.
#include <string>
#include <thread>
#include "pgld_internal.h"

namespace pgld {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
const char* last_error() { return g_err.c_str(); }

int make_layout(const pgld_panel_desc* d, PanelLayout* L) {
  if (!d) return fail(PGLD_EINVAL, "null panel descriptor");
  if (d->mode != PGLD_MODE_HAP && d->mode != PGLD_MODE_GENO)
    return fail(PGLD_EINVAL, "mode must be PGLD_MODE_HAP or PGLD_MODE_GENO");
  if (d->n_pops < 1 || d->n_pops > PGLD_MAX_POPS)
    return fail(PGLD_EINVAL, "1-4 populations supported");
  if (d->n_sites < 0) return fail(PGLD_EINVAL, "negative n_sites");
  L->n_pops = d->n_pops;
  L->mode = d->mode;
  L->n_sites = d->n_sites;
  L->n_blocks = (d->n_sites + PGLD_SITE_BLOCK - 1) / PGLD_SITE_BLOCK;
  int64_t off = 0;
  for (int p = 0; p < PGLD_MAX_POPS; ++p) {
    L->n_feat[p] = L->ksteps[p] = L->pop_size[p] = L->missing[p] = 0;
    L->pop_off[p] = 0;
  }
  for (int p = 0; p < d->n_pops; ++p) {
    if (d->pop_size[p] < 0 || d->pop_size[p] > 32767)
      return fail(PGLD_EINVAL, "population size must be in [0, 32767]");
    L->pop_size[p] = d->pop_size[p];
    L->missing[p] = d->has_missing[p] ? 1 : 0;
    L->n_feat[p] = (d->mode == PGLD_MODE_HAP ? 1 : 2) + L->missing[p];
    L->ksteps[p] = (d->pop_size[p] + PGLD_KSTEP - 1) / PGLD_KSTEP;
    L->pop_off[p] = off;
    off += (int64_t)L->n_feat[p] * L->ksteps[p] * PGLD_FRAG_BYTES;
  }
  L->block_bytes = off;
  return PGLD_OK;
}

// ---------------------------------------------------------------------------
// Site tallies: one thread per (site, pop); rows of the pop streamed.
// Consecutive threads read consecutive sites of the same row (coalesced).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) scan_sites_kernel(
    const int8_t* __restrict__ mat, int64_t n_sites, int64_t ld,
    const int32_t* __restrict__ rows, int32_t row_begin, int32_t row_end, int32_t mode,
    int32_t* __restrict__ stats /* [3][n_sites] for this pop */, int32_t* __restrict__ flags) {
  const int64_t site = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (site >= n_sites) return;
  int c1 = 0, c2 = 0, cm = 0, bad = 0;
  const int maxv = mode == PGLD_MODE_HAP ? 1 : 2;
  for (int r = row_begin; r < row_end; ++r) {
    const int v = mat[(int64_t)rows[r] * ld + site];
    c1 += (v == 1);
    c2 += (v == 2);
    cm += (v == -1);
    bad |= (v < -1) | (v > maxv);
  }
  stats[site] = c1;
  stats[n_sites + site] = c2;
  stats[2 * n_sites + site] = cm;
  if (bad) atomicOr(flags, 1);
}

// ---------------------------------------------------------------------------
// Pack: one workgroup per (4 site blocks = 128 sites, pop, 64-sample k-step).
// Phase 1 stages 64 sample rows x 128 sites through LDS (4 threads read one
// row's 128 contiguous bytes); phase 2 writes 4 x n_feat fragments, 16 bytes
// per lane, fully coalesced.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) pack_kernel(
    const int8_t* __restrict__ mat, int64_t n_sites, int64_t ld,
    const int32_t* __restrict__ rows, int32_t row_begin, int32_t pop_size, int32_t mode,
    int32_t n_feat, int32_t ksteps, int64_t pop_off, int64_t block_bytes, int64_t n_blocks,
    int8_t* __restrict__ panels) {
  __shared__ int8_t tile[64][128 + 4];  // [sample][site]
  const int64_t blk0 = (int64_t)blockIdx.x * 4;
  const int ks = blockIdx.y;
  const int t = threadIdx.x;
  {  // Phase 1: thread t loads 32 sites of sample row (t >> 2).
    const int s = t >> 2;
    const int k = ks * 64 + s;
    const int sx = (t & 3) * 32;
    const bool row_ok = k < pop_size;
    const int8_t* src = row_ok ? mat + (int64_t)rows[row_begin + k] * ld : mat;
#pragma unroll 8
    for (int e = 0; e < 32; ++e) {
      const int64_t site = blk0 * 32 + sx + e;
      const bool have = row_ok && site < n_sites;
      tile[s][sx + e] = have ? src[site] : (int8_t)-2;  // -2: padding (every feature 0)
    }
  }
  __syncthreads();
  // Phase 2: (block, feature, lane) -> 16 bytes = 32 samples of one site.
  for (int w = t; w < 4 * n_feat * 64; w += 256) {
    const int lane = w & 63, f = (w >> 6) % n_feat, bl = (w >> 6) / n_feat;
    if (blk0 + bl >= n_blocks) continue;
    const int site = bl * 32 + (lane & 31);
    const int s0 = (lane >> 5) * 32;
    // feature code: hap {0: H, 1: V}; geno {0: I1, 1: I2, 2: V}
    uint32_t word[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      uint32_t acc = 0;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int v = tile[s0 + q * 8 + e][site];
        int bit;
        if (mode == PGLD_MODE_HAP)
          bit = f == 0 ? (v == 1) : (v >= 0);
        else
          bit = f == 0 ? (v == 1) : (f == 1 ? (v == 2) : (v >= 0));
        acc |= (bit ? PGLD_FP4_ONE : 0u) << (4 * e);
      }
      word[q] = acc;
    }
    int8_t* dst = panels + (blk0 + bl) * block_bytes + pop_off +
                  ((int64_t)f * ksteps + ks) * PGLD_FRAG_BYTES + lane * 16;
    *reinterpret_cast<uint4*>(dst) = make_uint4(word[0], word[1], word[2], word[3]);
  }
}

}  // namespace pgld

using namespace pgld;

extern "C" int pgld_version(void) { return 100; }
extern "C" const char* pgld_last_error(void) { return pgld::last_error(); }

extern "C" int64_t pgld_panel_bytes(const pgld_panel_desc* desc) {
  PanelLayout L;
  if (make_layout(desc, &L) != PGLD_OK) return -1;
  return L.block_bytes * L.n_blocks;
}

static int check_rows(const pgld_panel_desc* desc, const int32_t* pop_offsets) {
  if (!pop_offsets) return fail(PGLD_EINVAL, "null pop_offsets");
  for (int p = 0; p < desc->n_pops; ++p) {
    if (pop_offsets[p + 1] - pop_offsets[p] != desc->pop_size[p])
      return fail(PGLD_EINVAL, "pop_offsets disagree with pop_size");
  }
  return PGLD_OK;
}

extern "C" int pgld_scan_sites(const int8_t* d_mat, int64_t n_rows, int64_t n_sites, int64_t ld,
                               const int32_t* d_rows, const int32_t* pop_offsets,
                               const pgld_panel_desc* desc, int32_t* d_site_stats,
                               int32_t* d_flags, void* stream) {
  PanelLayout L;
  int rc = make_layout(desc, &L);
  if (rc) return rc;
  if ((rc = check_rows(desc, pop_offsets))) return rc;
  if (n_sites != desc->n_sites || ld < n_sites || n_rows < 0)
    return fail(PGLD_EINVAL, "matrix shape disagrees with descriptor");
  if (n_sites == 0) return PGLD_OK;
  hipStream_t s = as_stream(stream);
  const unsigned grid = (unsigned)((n_sites + 255) / 256);
  for (int p = 0; p < desc->n_pops; ++p) {
    hipLaunchKernelGGL(scan_sites_kernel, dim3(grid), dim3(256), 0, s, d_mat, n_sites, ld, d_rows,
                       pop_offsets[p], pop_offsets[p + 1], desc->mode,
                       d_site_stats + (int64_t)p * 3 * n_sites, d_flags);
    PGLD_LAUNCH_CHECK();
  }
  return PGLD_OK;
}

extern "C" int pgld_pack(const int8_t* d_mat, int64_t n_rows, int64_t n_sites, int64_t ld,
                         const int32_t* d_rows, const int32_t* pop_offsets,
                         const pgld_panel_desc* desc, int8_t* d_panels, void* stream) {
  PanelLayout L;
  int rc = make_layout(desc, &L);
  if (rc) return rc;
  if ((rc = check_rows(desc, pop_offsets))) return rc;
  if (n_sites != desc->n_sites || ld < n_sites || n_rows < 0)
    return fail(PGLD_EINVAL, "matrix shape disagrees with descriptor");
  if (L.n_blocks == 0) return PGLD_OK;
  hipStream_t s = as_stream(stream);
  for (int p = 0; p < desc->n_pops; ++p) {
    if (L.ksteps[p] == 0) continue;
    hipLaunchKernelGGL(pack_kernel, dim3((unsigned)((L.n_blocks + 3) / 4), (unsigned)L.ksteps[p]),
                       dim3(256), 0, s, d_mat, n_sites, ld, d_rows, pop_offsets[p],
                       L.pop_size[p], desc->mode, L.n_feat[p], L.ksteps[p], L.pop_off[p],
                       L.block_bytes, L.n_blocks, d_panels);
    PGLD_LAUNCH_CHECK();
  }
  return PGLD_OK;
}
