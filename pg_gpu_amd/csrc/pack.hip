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
// Pack: one workgroup per (PACK_SB site blocks, pop, 64-sample k-step).
// Phase 1: thread (row quad j, 16-site group g) reads 16 bytes of each of the
// four sample rows 4j..4j+3 (rows not aligned to 16 bytes: two aligned loads
// and a byte funnel shift), transposes the 4 x 4 byte blocks in registers and
// stores one dword (four samples) per site into the LDS tile [site][sample].
// Phase 2 gives each lane the 32 samples of one site, turns them into 4-bit
// indicators with bytewise (SWAR) tests, and writes 16 bytes per lane, fully
// coalesced.
// ---------------------------------------------------------------------------
// Bit 0 of byte b set where byte b of x equals c (exact per byte: the 7-bit
// sums cannot carry across bytes).
__device__ __forceinline__ uint32_t bytes_eq(uint32_t x, uint32_t c) {
  const uint32_t u = x ^ c;
  return (~(((u & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | u | 0x7F7F7F7Fu)) >> 7;
}
// The four byte flags (bits 0, 8, 16, 24) moved to nibbles 0..3 (bits 0, 4, 8, 12).
__device__ __forceinline__ uint32_t flags_to_nibbles(uint32_t m) {
  const uint32_t y = m | (m >> 4);
  return (y & 0x11u) | ((y >> 8) & 0x1100u);
}

constexpr int PACK_SB = 8;  // site blocks per workgroup (256 sites, 16 groups of 16)
// Dwords per site row of the LDS tile: 64 samples + 3 dwords of padding.  The
// odd pitch puts the 32 sites one lane pair reads (phase 2) and the two
// halves of the groups one wave writes (phase 1) on distinct banks.
constexpr int PACK_PITCH = 19;

// The 16 bytes at src from the two aligned 16-byte words around them (a
// funnel shift; m = src mod 16, 0 included).  The 32 bytes from src - m must
// lie inside the matrix.
__device__ __forceinline__ uint4 load16_shifted(const int8_t* src) {
  const int m = (int)(reinterpret_cast<uintptr_t>(src) & 15u);
  // named values, not arrays: a select between neighbouring array elements
  // would become an indexed load through scratch memory
  const uint4* base = reinterpret_cast<const uint4*>(src - m);
  const uint4 q0 = base[0], q1 = base[1];
  const bool s4 = (m & 4) != 0, s8 = (m & 8) != 0;
  const uint32_t t0 = s4 ? q0.y : q0.x, t1 = s4 ? q0.z : q0.y, t2 = s4 ? q0.w : q0.z,
                 t3 = s4 ? q1.x : q0.w, t4 = s4 ? q1.y : q1.x, t5 = s4 ? q1.z : q1.y,
                 t6 = s4 ? q1.w : q1.z;
  const uint32_t u0 = s8 ? t2 : t0, u1 = s8 ? t3 : t1, u2 = s8 ? t4 : t2, u3 = s8 ? t5 : t3,
                 u4 = s8 ? t6 : t4;
  const uint32_t sh = (uint32_t)(m & 3);
  return make_uint4(__builtin_amdgcn_alignbyte(u1, u0, sh), __builtin_amdgcn_alignbyte(u2, u1, sh),
                    __builtin_amdgcn_alignbyte(u3, u2, sh), __builtin_amdgcn_alignbyte(u4, u3, sh));
}

// Sites site0 .. site0+15 of the row at src, as four little-endian dwords;
// sites at or past n_sites read as -2 (no feature set).
__device__ __forceinline__ uint4 pack_load16(const int8_t* mat, const int8_t* mat_end,
                                             const int8_t* src, int64_t site0, int64_t n_sites) {
  const int m = (int)(reinterpret_cast<uintptr_t>(src) & 15u);  // row misalignment
  uint32_t w[4];
  if (m == 0 && site0 + 16 <= n_sites) {
    return *reinterpret_cast<const uint4*>(src);
  } else if (site0 + 16 <= n_sites && src - m >= mat && src - m + 32 <= mat_end) {
    return load16_shifted(src);
  } else {
#pragma unroll
    for (int e = 0; e < 4; ++e) w[e] = 0xFEFEFEFEu;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const uint32_t v = site0 + e < n_sites ? (uint8_t)src[e] : 0xFEu;
      w[e >> 2] = (w[e >> 2] & ~(0xFFu << (8 * (e & 3)))) | (v << (8 * (e & 3)));
    }
  }
  return make_uint4(w[0], w[1], w[2], w[3]);
}

// ---------------------------------------------------------------------------
// Site tallies.  A workgroup covers 256 consecutive sites (16 groups of 16)
// of one population and splits the population's rows into 32 interleaved
// slices: thread (slice, group) reads 16 bytes of each of its rows (a wave
// reads four rows' 256-byte segments per load) and counts the bytes equal to
// 1, 2 and -1 with bytewise (SWAR) tests into per-byte counters, flushed to
// integers every 255 rows.  The four slices of a wave are summed with two
// lane exchanges, the eight waves through LDS; the tallies are written once,
// coalesced.  (Many slices and small site blocks keep enough loads in flight:
// one workgroup per 512 sites with 8 slices ran below one wave per SIMD on
// 100k sites.)  A byte of a real site outside the mode's alphabet sets
// flags[0].
// ---------------------------------------------------------------------------
#ifndef PGLD_SCAN_SLICES
#define PGLD_SCAN_SLICES 32
#endif
#ifndef PGLD_SCAN_GROUPS
#define PGLD_SCAN_GROUPS 16
#endif
constexpr int SCAN_SLICES = PGLD_SCAN_SLICES;
constexpr int SCAN_GROUPS = PGLD_SCAN_GROUPS;
constexpr int SCAN_SITES = SCAN_GROUPS * 16;
constexpr int SCAN_THREADS = SCAN_SLICES * SCAN_GROUPS;
constexpr int SCAN_WAVES = SCAN_THREADS / 64;

struct ScanParams {
  const int8_t* mat;
  const int8_t* mat_end;
  int64_t n_sites, ld;
  const int32_t* rows;
  int32_t offs[PGLD_MAX_POPS + 1];
  int32_t mode;
  int32_t aligned;  // matrix and leading dimension 16-byte aligned
  int32_t* stats;   // [P][3][n_sites]
  int32_t* flags;
};

__global__ void __launch_bounds__(SCAN_THREADS) scan_sites_kernel(ScanParams a) {
  __shared__ int part[SCAN_WAVES][3 * SCAN_SITES];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int p = blockIdx.y;
  const int g = tid % SCAN_GROUPS, slice = tid / SCAN_GROUPS;
  const int64_t base = (int64_t)blockIdx.x * SCAN_SITES;
  const int64_t site0 = base + 16 * g;
  const bool geno = a.mode == PGLD_MODE_GENO;
  int c[16][3];
#pragma unroll
  for (int k = 0; k < 16; ++k) c[k][0] = c[k][1] = c[k][2] = 0;
  if (site0 < a.n_sites) {
    // bit 8b of vm[e] set where site site0 + 4e + b is a real site
    uint32_t vm[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      uint32_t m = 0;
#pragma unroll
      for (int b = 0; b < 4; ++b) m |= (site0 + 4 * e + b < a.n_sites ? 1u : 0u) << (8 * b);
      vm[e] = m;
    }
    uint32_t a1[4] = {0, 0, 0, 0}, a2[4] = {0, 0, 0, 0}, am[4] = {0, 0, 0, 0};
    uint32_t bad = 0;
    int pending = 0;  // rows in the byte counters (flushed before 255)
    auto tally = [&](const uint4 q) {
      const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const uint32_t e1 = bytes_eq(w[e], 0x01010101u), e2 = bytes_eq(w[e], 0x02020202u);
        const uint32_t em = bytes_eq(w[e], 0xFFFFFFFFu), e0 = bytes_eq(w[e], 0u);
        a1[e] += e1;
        a2[e] += e2;
        am[e] += em;
        bad |= ~(e0 | e1 | em | (geno ? e2 : 0u)) & vm[e];
      }
    };
    auto flush = [&]() {
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const int e = k >> 2, sh = 8 * (k & 3);
        c[k][0] += (a1[e] >> sh) & 0xFFu;
        c[k][1] += (a2[e] >> sh) & 0xFFu;
        c[k][2] += (am[e] >> sh) & 0xFFu;
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) a1[e] = a2[e] = am[e] = 0u;
      pending = 0;
    };
    const int r0 = a.offs[p], r1 = a.offs[p + 1];
    const int last = r1 - 1;
    int r = r0 + slice;
    if (a.aligned && site0 + 16 <= a.n_sites) {
      // rows 16-byte aligned and this thread's 16 sites all real: four rows'
      // plain loads in flight, the next four row ids loaded one step ahead
      // (a separate loop, so no alignment branch sits between the loads)
      int id[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) id[k] = r0 < r1 ? a.rows[min(r + k * SCAN_SLICES, last)] : 0;
      // (the last step's ids past the population repeat its last row and are
      // not tallied)
      for (; r < r1; r += 4 * SCAN_SLICES) {
        uint4 q[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
          q[k] = *reinterpret_cast<const uint4*>(a.mat + (int64_t)id[k] * a.ld + site0);
#pragma unroll
        for (int k = 0; k < 4; ++k) id[k] = a.rows[min(r + (4 + k) * SCAN_SLICES, last)];
#pragma unroll
        for (int k = 0; k < 4; ++k)
          if (r + k * SCAN_SLICES < r1) tally(q[k]);
        pending += 4;
        if (pending > 250) flush();
      }
    } else if (site0 + 16 <= a.n_sites) {
      // rows not 16-byte aligned: every row through the branch-free shifted
      // load (a shift of 0 for a row that happens to be aligned), so lanes of
      // different rows do not diverge; the generic load only where the two
      // words would pass the end of the matrix
      int id[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) id[k] = r0 < r1 ? a.rows[min(r + k * SCAN_SLICES, last)] : 0;
      for (; r < r1; r += 4 * SCAN_SLICES) {
        const int8_t* src[4];
        bool inside = true;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          src[k] = a.mat + (int64_t)id[k] * a.ld + site0;
          const int8_t* b0 = src[k] - (reinterpret_cast<uintptr_t>(src[k]) & 15u);
          inside = inside && b0 >= a.mat && b0 + 32 <= a.mat_end;
        }
        uint4 q[4];
        if (inside) {
#pragma unroll
          for (int k = 0; k < 4; ++k) q[k] = load16_shifted(src[k]);
        } else {
#pragma unroll
          for (int k = 0; k < 4; ++k) q[k] = pack_load16(a.mat, a.mat_end, src[k], site0, a.n_sites);
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) id[k] = a.rows[min(r + (4 + k) * SCAN_SLICES, last)];
#pragma unroll
        for (int k = 0; k < 4; ++k)
          if (r + k * SCAN_SLICES < r1) tally(q[k]);
        pending += 4;
        if (pending > 250) flush();
      }
    }
    for (; r < r1; r += SCAN_SLICES) {
      tally(pack_load16(a.mat, a.mat_end, a.mat + (int64_t)a.rows[r] * a.ld + site0, site0,
                        a.n_sites));
      if (++pending == 255) flush();
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int e = k >> 2, sh = 8 * (k & 3);
      c[k][0] += (a1[e] >> sh) & 0xFFu;
      c[k][1] += (a2[e] >> sh) & 0xFFu;
      c[k][2] += (am[e] >> sh) & 0xFFu;
    }
    if (bad) atomicOr(a.flags, 1);
  }
  // lanes g, g + SCAN_GROUPS, ... of a wave hold the same sites
#pragma unroll
  for (int k = 0; k < 16; ++k)
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      int v = c[k][q];
#pragma unroll
      for (int m = SCAN_GROUPS; m < 64; m <<= 1) v += __shfl_xor(v, m, 64);
      if (lane < SCAN_GROUPS) part[wave][q * SCAN_SITES + 16 * g + k] = v;
    }
  __syncthreads();
  int32_t* st = a.stats + (int64_t)p * 3 * a.n_sites;
  for (int e = tid; e < 3 * SCAN_SITES; e += SCAN_THREADS) {
    const int q = e / SCAN_SITES, k = e - q * SCAN_SITES;
    int v = 0;
#pragma unroll
    for (int w = 0; w < SCAN_WAVES; ++w) v += part[w][e];
    if (base + k < a.n_sites) st[(int64_t)q * a.n_sites + base + k] = v;
  }
}

__global__ void __launch_bounds__(256) pack_kernel(
    const int8_t* __restrict__ mat, int64_t n_sites, int64_t ld,
    const int32_t* __restrict__ rows, int32_t row_begin, int32_t pop_size, int32_t mode,
    int32_t n_feat, int32_t ksteps, int64_t pop_off, int64_t block_bytes, int64_t n_blocks,
    const int8_t* __restrict__ mat_end, int8_t* __restrict__ panels) {
  constexpr int SITES = 32 * PACK_SB;
  __shared__ uint32_t tile[SITES * PACK_PITCH];  // [site][4-sample dword]
  const int64_t blk0 = (int64_t)blockIdx.x * PACK_SB;
  const int ks = blockIdx.y;
  const int t = threadIdx.x;
  {
    // Phase 1: 16 row quads x 16 site groups = 256 threads
    const int j = t & 15, g = t >> 4;
    const int64_t site0 = blk0 * 32 + g * 16;
    uint32_t w[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int k = ks * 64 + 4 * j + i;
      const uint4 r = k < pop_size ? pack_load16(mat, mat_end,
                                                 mat + (int64_t)rows[row_begin + k] * ld + site0,
                                                 site0, n_sites)
                                   : make_uint4(0xFEFEFEFEu, 0xFEFEFEFEu, 0xFEFEFEFEu,
                                                0xFEFEFEFEu);  // -2: padding samples
      w[i][0] = r.x; w[i][1] = r.y; w[i][2] = r.z; w[i][3] = r.w;
    }
    // byte b of dword q of row i -> byte i of the dword of site 4q + b
#pragma unroll
    for (int q = 0; q < 4; ++q) {
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const uint32_t v = ((w[0][q] >> (8 * b)) & 0xFFu) | (((w[1][q] >> (8 * b)) & 0xFFu) << 8) |
                           (((w[2][q] >> (8 * b)) & 0xFFu) << 16) |
                           (((w[3][q] >> (8 * b)) & 0xFFu) << 24);
        tile[(g * 16 + 4 * q + b) * PACK_PITCH + j] = v;
      }
    }
  }
  __syncthreads();
  // Phase 2: (block, feature, lane) -> 16 bytes = 32 samples of one site.
  for (int w = t; w < PACK_SB * n_feat * 64; w += 256) {
    const int lane = w & 63, f = (w >> 6) % n_feat, bl = (w >> 6) / n_feat;
    if (blk0 + bl >= n_blocks) continue;
    const int site = bl * 32 + (lane & 31);
    const uint32_t* src = tile + site * PACK_PITCH + (lane >> 5) * 8;
    uint32_t x[8];  // samples s0 + 4j .. s0 + 4j + 3, s0 = 32 (lane >> 5)
#pragma unroll
    for (int e = 0; e < 8; ++e) x[e] = src[e];
    // feature code: hap {0: [v==1], 1: [v>=0]}; geno (centred dosages)
    // {0: u = v - 1 in {-1, 0, 1}, 1: q = u^2, 2: [v>=0]}, u and q 0 where
    // missing, so the products of the features are the centred dosage moments
    const bool valid_f = f == n_feat - 1 && f > 0 && (mode == PGLD_MODE_HAP ? f == 1 : f == 2);
    const bool geno = mode == PGLD_MODE_GENO;
    uint32_t word[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      // nibble i of word q <-> sample s0 + 8q + i; 1.0 in E2M1 is 0x2, -1.0 0xA
      if (valid_f) {
        const uint32_t m0 = (~x[2 * q] & 0x80808080u) >> 7;
        const uint32_t m1 = (~x[2 * q + 1] & 0x80808080u) >> 7;
        word[q] = (flags_to_nibbles(m0) | (flags_to_nibbles(m1) << 16)) * PGLD_FP4_ONE;
      } else if (geno) {
        const uint32_t z = flags_to_nibbles(bytes_eq(x[2 * q], 0u)) |
                           (flags_to_nibbles(bytes_eq(x[2 * q + 1], 0u)) << 16);
        const uint32_t t = flags_to_nibbles(bytes_eq(x[2 * q], 0x02020202u)) |
                           (flags_to_nibbles(bytes_eq(x[2 * q + 1], 0x02020202u)) << 16);
        // per nibble at most one of z, t is set: no carries
        word[q] = f == 0 ? t * PGLD_FP4_ONE + z * PGLD_FP4_NEG_ONE : (z | t) * PGLD_FP4_ONE;
      } else {
        const uint32_t m0 = bytes_eq(x[2 * q], 0x01010101u);
        const uint32_t m1 = bytes_eq(x[2 * q + 1], 0x01010101u);
        word[q] = (flags_to_nibbles(m0) | (flags_to_nibbles(m1) << 16)) * PGLD_FP4_ONE;
      }
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
  ScanParams a;
  a.mat = d_mat;
  a.mat_end = d_mat + (n_rows > 0 ? (n_rows - 1) * ld + n_sites : 0);
  a.n_sites = n_sites;
  a.ld = ld;
  a.rows = d_rows;
  for (int p = 0; p <= PGLD_MAX_POPS; ++p) a.offs[p] = p <= desc->n_pops ? pop_offsets[p] : 0;
  a.mode = desc->mode;
  a.aligned = (reinterpret_cast<uintptr_t>(d_mat) % 16 == 0) && (ld % 16 == 0);
  a.stats = d_site_stats;
  a.flags = d_flags;
  const unsigned grid = (unsigned)((n_sites + SCAN_SITES - 1) / SCAN_SITES);
  hipLaunchKernelGGL(scan_sites_kernel, dim3(grid, (unsigned)desc->n_pops), dim3(SCAN_THREADS), 0,
                     as_stream(stream), a);
  PGLD_LAUNCH_CHECK();
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
  const int8_t* mat_end = d_mat + (n_rows > 0 ? (n_rows - 1) * ld + n_sites : 0);
  for (int p = 0; p < desc->n_pops; ++p) {
    if (L.ksteps[p] == 0) continue;
    hipLaunchKernelGGL(pack_kernel,
                       dim3((unsigned)((L.n_blocks + PACK_SB - 1) / PACK_SB), (unsigned)L.ksteps[p]),
                       dim3(256), 0, s, d_mat, n_sites, ld, d_rows, pop_offsets[p],
                       L.pop_size[p], desc->mode, L.n_feat[p], L.ksteps[p], L.pop_off[p],
                       L.block_bytes, L.n_blocks, mat_end, d_panels);
    PGLD_LAUNCH_CHECK();
  }
  return PGLD_OK;
}
