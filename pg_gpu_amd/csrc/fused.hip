// Fused binned two-locus LD for haplotype and genotype data on gfx950.
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
#include <algorithm>
#include <mutex>
#include "pgld_internal.h"
#include "stats_hap.inc"
#include "stats_geno.inc"

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
  uint32_t* sched;                // per-XCD tile counters (null: static schedule)
  int32_t use_ring;               // stream operands through the LDS ring (traits with RING)
};

// 16 packed count words per lane (one per MFMA output slot).  Kept as a
// vector value so that the epilogue's wave-uniform slot index r selects a
// register (index mode) instead of spilling the tile's counts to scratch.
typedef uint32_t v16u __attribute__((ext_vector_type(16)));

// Row (within a 32x32 MFMA output block) of accumulator register r of `lane`;
// the column is lane & 31.
__device__ __forceinline__ int mfma_row(int lane, int r) {
  return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
}

// acc1 += R.C1 and acc2 += R.C2 over the ks k-steps of one row operand R and
// two column operands; operand loads run PF steps ahead in a register ring,
// so the L2 latency of one step overlaps the multiplies of the previous ones.
__device__ __forceinline__ void mfma_row_pass(const v4i* R, const v4i* C1, const v4i* C2, int ks,
                                              v16f& acc1, v16f& acc2) {
  constexpr int PF = 2;
  v4i r[PF], c1[PF], c2[PF];
#pragma unroll
  for (int d = 0; d < PF; ++d)
    if (d < ks) { r[d] = R[d * 64]; c1[d] = C1[d * 64]; c2[d] = C2[d * 64]; }
  for (int s0 = 0; s0 < ks; s0 += PF) {
#pragma unroll
    for (int d = 0; d < PF; ++d) {
      const int s = s0 + d;
      if (s < ks) {
        acc1 = pgld_mfma_fp4(r[d], c1[d], acc1);
        acc2 = pgld_mfma_fp4(r[d], c2[d], acc2);
        if (s + PF < ks) {
          const int n = (s + PF) * 64;
          r[d] = R[n]; c1[d] = C1[n]; c2[d] = C2[n];
        }
      }
    }
  }
}

// Operand ring in LDS: each wave streams its tile's fragments through a
// private ring of RING k-steps x 4 fragments (1 KiB each) with
// global_load_lds (lane l's 16 bytes land at byte 16 l of the slot), so loads
// run RING k-steps ahead without holding registers.  The LDS writes of these
// loads are tracked by vmcnt only, hence the explicit waits.
typedef __attribute__((address_space(1))) void gmem_void;
typedef __attribute__((address_space(3))) void lmem_void;

__device__ __forceinline__ void lds_dma16(const v4i* g_lane, v4i* lds_slot) {
  __builtin_amdgcn_global_load_lds((gmem_void*)g_lane, (lmem_void*)lds_slot, 16, 0, 0);
}
// Wait until at most n (0, 4, 8, 12 or 16) vector-memory operations of this
// wave are outstanding; also a compiler barrier for memory accesses.
__device__ __forceinline__ void wait_vm(int n) {
  if (n >= 16) asm volatile("s_waitcnt vmcnt(16)" ::: "memory");
  else if (n >= 12) asm volatile("s_waitcnt vmcnt(12)" ::: "memory");
  else if (n >= 8) asm volatile("s_waitcnt vmcnt(8)" ::: "memory");
  else if (n >= 4) asm volatile("s_waitcnt vmcnt(4)" ::: "memory");
  else asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

// The four fragments of one ring slot (1 KiB apart).  Read in one asm block
// that waits for its own results: the compiler cannot tell these reads from
// the in-flight ring loads of later k-steps and would otherwise wait for all
// of them (vmcnt(0)) before every read.
__device__ __forceinline__ void ring_read4(const v4i* slot_lane, v4i& f0, v4i& f1, v4i& f2,
                                           v4i& f3) {
  typedef const __attribute__((address_space(3))) v4i lds_v4i;
  const uint32_t addr = (uint32_t)(uintptr_t)(lds_v4i*)slot_lane;  // LDS byte address
  asm volatile(
      "ds_read_b128 %0, %4\n\t"
      "ds_read_b128 %1, %4 offset:1024\n\t"
      "ds_read_b128 %2, %4 offset:2048\n\t"
      "ds_read_b128 %3, %4 offset:3072\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(f0), "=&v"(f1), "=&v"(f2), "=&v"(f3)
      : "v"(addr)
      : "memory");
}

// hh, hv, vh, vv over the ks k-steps, operands streamed through the ring
template <int RING>
__device__ __forceinline__ void ring_products(const v4i* A, const v4i* Av, const v4i* B,
                                              const v4i* Bv, int ks, int lane, v4i* ring,
                                              v16f& hh, v16f& hv, v16f& vh, v16f& vv) {
  const v4i* src[4] = {A, Av, B, Bv};
  const int pre = ks < RING ? ks : RING;
  for (int d = 0; d < pre; ++d) {
#pragma unroll
    for (int f = 0; f < 4; ++f) lds_dma16(src[f] + d * 64, ring + (d * 4 + f) * 64);
  }
  for (int s = 0; s < ks; ++s) {
    const int ahead = ks - 1 - s < RING - 1 ? ks - 1 - s : RING - 1;  // later k-steps in flight
    wait_vm(4 * ahead);
    v4i ah, av, bh, bv;
    ring_read4(ring + (s % RING) * 4 * 64 + lane, ah, av, bh, bv);
    hh = pgld_mfma_fp4(ah, bh, hh);
    hv = pgld_mfma_fp4(ah, bv, hv);
    vh = pgld_mfma_fp4(av, bh, vh);
    vv = pgld_mfma_fp4(av, bv, vv);
    if (s + RING < ks) {
      asm volatile("" ::: "memory");
      v4i* dst = ring + (s % RING) * 4 * 64;
#pragma unroll
      for (int f = 0; f < 4; ++f) lds_dma16(src[f] + (s + RING) * 64, dst + f * 64);
    }
  }
}

// Counts of population p for tile (rb, cb), packed per slot as
// w0 = n11 | n10 << 16, w1 = n01 | n00 << 16.  With RING > 0 (and missing
// calls) the operands stream through the wave's LDS ring.
template <bool TWO_PASS, int RING = 0>
__device__ __forceinline__ void tile_counts(const FusedParams& a, int p, int64_t rb, int64_t cb,
                                            int lane, v16u& w0, v16u& w1, v4i* ring = nullptr) {
  // MARG: slot words hold the products' marginals (n11 | pA << 16, pB | n << 16
  // with pA = n11 + n10, pB = n11 + n01, n the valid pairs); otherwise the
  // four class counts (n11 | n10 << 16, n01 | n00 << 16).  Measured per
  // traits: marginals for two to four populations (2 x 500 haplotypes over
  // 200k sites: 3.36 -> 3.19 ms; 4 populations of 40-800 over 500k sites:
  // 169.3 -> 166.9 ms); class counts for one (300 haplotypes in 5,000 ragged
  // windows: 3.17 -> 3.29 ms with marginals).
  constexpr bool MARG = !TWO_PASS;
  const int ks = a.ksteps[p];
  const v4i* A = reinterpret_cast<const v4i*>(a.panels + rb * a.block_bytes + a.pop_off[p]) + lane;
  const v4i* B = reinterpret_cast<const v4i*>(a.panels + cb * a.block_bytes + a.pop_off[p]) + lane;
  if (RING > 0 && ring != nullptr && a.missing[p] && !TWO_PASS) {
    v16f hh = zero16f(), hv = zero16f(), vh = zero16f(), vv = zero16f();
    ring_products<(RING > 0 ? RING : 1)>(A, A + ks * 64, B, B + ks * 64, ks, lane, ring, hh, hv,
                                         vh, vv);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int n11 = (int)hh[r], h_v = (int)hv[r], v_h = (int)vh[r], v_v = (int)vv[r];
      w0[r] = (uint32_t)n11 | ((uint32_t)h_v << 16);
      w1[r] = (uint32_t)v_h | ((uint32_t)v_v << 16);
    }
  } else if (a.missing[p] && !TWO_PASS) {
    v16f hh, hv, vh, vv;
    const v4i* Av = A + ks * 64;
    const v4i* Bv = B + ks * 64;
    // one k-step ahead: the registers a second step would hold cost more
    // (spills, occupancy of the epilogue) than its latency cover gains.  The
    // next k-step's loads go to the operand registers as soon as the MFMAs of
    // this step have read them.
    if constexpr (RING == 0) {
      // the first k-step starts from zero accumulators (an inline constant
      // of the MFMA), so the 64 accumulator registers need no clearing
      if (ks > 0) {
        v4i ah = A[0], av = Av[0], bh = B[0], bv = Bv[0];
        hh = pgld_mfma_fp4(ah, bh, zero16f());
        hv = pgld_mfma_fp4(ah, bv, zero16f());
        vh = pgld_mfma_fp4(av, bh, zero16f());
        vv = pgld_mfma_fp4(av, bv, zero16f());
        if (ks > 1) { ah = A[64]; av = Av[64]; bh = B[64]; bv = Bv[64]; }
        for (int s = 1; s < ks; ++s) {
          hh = pgld_mfma_fp4(ah, bh, hh);
          hv = pgld_mfma_fp4(ah, bv, hv);
          vh = pgld_mfma_fp4(av, bh, vh);
          vv = pgld_mfma_fp4(av, bv, vv);
          if (s + 1 < ks) {
            const int n = (s + 1) * 64;
            ah = A[n]; av = Av[n]; bh = B[n]; bv = Bv[n];
          }
        }
      } else {
        hh = hv = vh = vv = zero16f();
      }
    } else {
      // fallback of the LDS ring traits: the form their register allocation
      // was measured with
      hh = hv = vh = vv = zero16f();
      v4i ah, av, bh, bv;
      if (ks > 0) { ah = A[0]; av = Av[0]; bh = B[0]; bv = Bv[0]; }
      for (int s = 0; s < ks; ++s) {
        hh = pgld_mfma_fp4(ah, bh, hh);
        hv = pgld_mfma_fp4(ah, bv, hv);
        vh = pgld_mfma_fp4(av, bh, vh);
        vv = pgld_mfma_fp4(av, bv, vv);
        if (s + 1 < ks) {
          const int n = (s + 1) * 64;
          ah = A[n]; av = Av[n]; bh = B[n]; bv = Bv[n];
        }
      }
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int n11 = (int)hh[r], h_v = (int)hv[r], v_h = (int)vh[r], v_v = (int)vv[r];
      if constexpr (MARG) {
        w0[r] = (uint32_t)n11 | ((uint32_t)h_v << 16);
        w1[r] = (uint32_t)v_h | ((uint32_t)v_v << 16);
      } else {
        w0[r] = (uint32_t)n11 | ((uint32_t)(h_v - n11) << 16);
        w1[r] = (uint32_t)(v_h - n11) | ((uint32_t)(v_v - h_v - v_h + n11) << 16);
      }
    }
  } else if (a.missing[p]) {
    // Two passes over K with one row operand each (H, then V) keep two
    // accumulators live instead of four (for three waves per SIMD); the
    // column operands are re-read from L1/L2 in the second pass.
    v16f hh = zero16f(), hv = zero16f();
    mfma_row_pass(A, B, B + ks * 64, ks, hh, hv);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int n11 = (int)hh[r], h_v = (int)hv[r];
      w0[r] = (uint32_t)n11 | ((uint32_t)(MARG ? h_v : h_v - n11) << 16);
    }
    v16f vh = zero16f(), vv = zero16f();
    mfma_row_pass(A + ks * 64, B, B + ks * 64, ks, vh, vv);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int v_h = (int)vh[r], v_v = (int)vv[r];
      if constexpr (MARG) {
        w1[r] = (uint32_t)v_h | ((uint32_t)v_v << 16);
      } else {
        const int n11 = (int)(w0[r] & 0xffffu), h_v = n11 + (int)(w0[r] >> 16);
        w1[r] = (uint32_t)(v_h - n11) | ((uint32_t)(v_v - h_v - v_h + n11) << 16);
      }
    }
  } else {
    v16f hh = zero16f();
    constexpr int PF = 2;
    v4i ah[PF], bh[PF];
#pragma unroll
    for (int d = 0; d < PF; ++d)
      if (d < ks) { ah[d] = A[d * 64]; bh[d] = B[d * 64]; }
    for (int s0 = 0; s0 < ks; s0 += PF) {
#pragma unroll
      for (int d = 0; d < PF; ++d) {
        const int s = s0 + d;
        if (s < ks) {
          hh = pgld_mfma_fp4(ah[d], bh[d], hh);
          if (s + PF < ks) { ah[d] = A[(s + PF) * 64]; bh[d] = B[(s + PF) * 64]; }
        }
      }
    }
    const int32_t* cnt1 = a.site_stats + (int64_t)p * 3 * a.V;
    const int64_t j = cb * 32 + (lane & 31);
    const int kj = j < a.V ? cnt1[j] : 0;
    const int K = a.K[p];
    // row i = ib + o, o = mfma_row(lane, r) - 4 (lane >> 5) a constant per r:
    // immediate load offsets from one base, o < lim <=> i < V
    const int64_t ib = rb * 32 + 4 * (lane >> 5);
    const int32_t* ci = cnt1 + ib;
    const int lim = (int)(a.V - ib < 32 ? a.V - ib : 32);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int o = 8 * (r >> 2) + (r & 3);
      const int ki = o < lim ? ci[o] : 0;
      const int n11 = (int)hh[r];
      if constexpr (MARG) {
        w0[r] = (uint32_t)n11 | ((uint32_t)ki << 16);
        w1[r] = (uint32_t)kj | ((uint32_t)K << 16);
      } else {
        w0[r] = (uint32_t)n11 | ((uint32_t)(ki - n11) << 16);
        w1[r] = (uint32_t)(kj - n11) | ((uint32_t)(K - ki - kj + n11) << 16);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Genotype tile counts.  Features per sample: b = [g==1], c = [g==2] and, with
// missing calls, v = [g>=0].  The 9 products give the 9 class counts:
//   n11 = b.b, n12 = b.c, n21 = c.b, n22 = c.c, n10 = b.v - n11 - n12,
//   n20 = c.v - n21 - n22, n01 = v.b - n11 - n21, n02 = v.c - n12 - n22,
//   n00 = v.v - (b.v + c.v) - n01 - n02.
// Packed per slot in moments order g1..g9 (n22, n21, n20, n12, n11, n10, n02,
// n01, n00) as five 16-bit pairs.
// ---------------------------------------------------------------------------
// Pack the nine class counts of slot r (moments order) into w[.][r].
template <int W>
__device__ __forceinline__ void geno_pack_slot(v16u (&w)[W], int r, const int (&c)[9]) {
  if constexpr (W == 3) {  // 10-bit fields: every count <= population size <= 1023
#pragma unroll
    for (int q = 0; q < 3; ++q)
      w[q][r] = (uint32_t)c[3 * q] | ((uint32_t)c[3 * q + 1] << 10) |
                ((uint32_t)c[3 * q + 2] << 20);
  } else {  // 16-bit fields
#pragma unroll
    for (int q = 0; q < 4; ++q) w[q][r] = (uint32_t)c[2 * q] | ((uint32_t)c[2 * q + 1] << 16);
    w[4][r] = (uint32_t)c[8];
  }
}

template <int W>
__device__ __forceinline__ void geno_tile_counts(const FusedParams& a, int p, int64_t rb,
                                                 int64_t cb, int lane, v16u (&w)[W]) {
  const int ks = a.ksteps[p];
  const v4i* A = reinterpret_cast<const v4i*>(a.panels + rb * a.block_bytes + a.pop_off[p]) + lane;
  const v4i* B = reinterpret_cast<const v4i*>(a.panels + cb * a.block_bytes + a.pop_off[p]) + lane;
  const int fs = ks * 64;  // feature stride in v4i
  v16f bb, bc, cbm, cc;
  if (a.missing[p]) {
    v16f bv, cv, vb, vc, vv;
    // one k-step ahead: nine accumulators leave no registers for more (two
    // steps ahead spilled and was no faster).  The first k-step starts from
    // zero accumulators (an inline constant of the MFMA), so the 144
    // accumulator registers need no clearing; the next step's loads go to the
    // operand registers once this step's MFMAs have read them.
    if (ks > 0) {
      v4i fa[3], fb[3];
#pragma unroll
      for (int f = 0; f < 3; ++f) { fa[f] = A[f * fs]; fb[f] = B[f * fs]; }
      bb = pgld_mfma_fp4(fa[0], fb[0], zero16f());
      bc = pgld_mfma_fp4(fa[0], fb[1], zero16f());
      bv = pgld_mfma_fp4(fa[0], fb[2], zero16f());
      cbm = pgld_mfma_fp4(fa[1], fb[0], zero16f());
      cc = pgld_mfma_fp4(fa[1], fb[1], zero16f());
      cv = pgld_mfma_fp4(fa[1], fb[2], zero16f());
      vb = pgld_mfma_fp4(fa[2], fb[0], zero16f());
      vc = pgld_mfma_fp4(fa[2], fb[1], zero16f());
      vv = pgld_mfma_fp4(fa[2], fb[2], zero16f());
      if (ks > 1)
#pragma unroll
        for (int f = 0; f < 3; ++f) { fa[f] = A[f * fs + 64]; fb[f] = B[f * fs + 64]; }
      for (int s = 1; s < ks; ++s) {
        bb = pgld_mfma_fp4(fa[0], fb[0], bb);
        bc = pgld_mfma_fp4(fa[0], fb[1], bc);
        bv = pgld_mfma_fp4(fa[0], fb[2], bv);
        cbm = pgld_mfma_fp4(fa[1], fb[0], cbm);
        cc = pgld_mfma_fp4(fa[1], fb[1], cc);
        cv = pgld_mfma_fp4(fa[1], fb[2], cv);
        vb = pgld_mfma_fp4(fa[2], fb[0], vb);
        vc = pgld_mfma_fp4(fa[2], fb[1], vc);
        vv = pgld_mfma_fp4(fa[2], fb[2], vv);
        if (s + 1 < ks) {
          const int n = (s + 1) * 64;
#pragma unroll
          for (int f = 0; f < 3; ++f) { fa[f] = A[f * fs + n]; fb[f] = B[f * fs + n]; }
        }
      }
    } else {
      bb = bc = cbm = cc = bv = cv = vb = vc = vv = zero16f();
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int n11 = (int)bb[r], n12 = (int)bc[r], n21 = (int)cbm[r], n22 = (int)cc[r];
      const int b_v = (int)bv[r], c_v = (int)cv[r];
      const int n10 = b_v - n11 - n12, n20 = c_v - n21 - n22;
      const int n01 = (int)vb[r] - n11 - n21, n02 = (int)vc[r] - n12 - n22;
      const int n00 = (int)vv[r] - b_v - c_v - n01 - n02;
      const int c[9] = {n22, n21, n20, n12, n11, n10, n02, n01, n00};
      geno_pack_slot(w, r, c);
    }
  } else {
    bb = bc = cbm = cc = zero16f();
    constexpr int PF = 2;
    v4i fa[PF][2], fb[PF][2];
#pragma unroll
    for (int d = 0; d < PF; ++d)
      if (d < ks)
#pragma unroll
        for (int f = 0; f < 2; ++f) { fa[d][f] = A[f * fs + d * 64]; fb[d][f] = B[f * fs + d * 64]; }
    for (int s0 = 0; s0 < ks; s0 += PF) {
#pragma unroll
      for (int d = 0; d < PF; ++d) {
        const int s = s0 + d;
        if (s < ks) {
          bb = pgld_mfma_fp4(fa[d][0], fb[d][0], bb);
          bc = pgld_mfma_fp4(fa[d][0], fb[d][1], bc);
          cbm = pgld_mfma_fp4(fa[d][1], fb[d][0], cbm);
          cc = pgld_mfma_fp4(fa[d][1], fb[d][1], cc);
          if (s + PF < ks) {
            const int n = (s + PF) * 64;
#pragma unroll
            for (int f = 0; f < 2; ++f) { fa[d][f] = A[f * fs + n]; fb[d][f] = B[f * fs + n]; }
          }
        }
      }
    }
    const int32_t* st = a.site_stats + (int64_t)p * 3 * a.V;
    const int64_t j = cb * 32 + (lane & 31);
    const int vbj = j < a.V ? st[j] : 0, vcj = j < a.V ? st[a.V + j] : 0;
    const int K = a.K[p];
    // rows as in tile_counts: i = ib + o with o constant per r
    const int64_t ib = rb * 32 + 4 * (lane >> 5);
    const int32_t* si = st + ib;
    const int32_t* sci = st + a.V + ib;
    const int lim = (int)(a.V - ib < 32 ? a.V - ib : 32);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int o = 8 * (r >> 2) + (r & 3);
      const int bvi = o < lim ? si[o] : 0, cvi = o < lim ? sci[o] : 0;
      const int n11 = (int)bb[r], n12 = (int)bc[r], n21 = (int)cbm[r], n22 = (int)cc[r];
      const int n10 = bvi - n11 - n12, n20 = cvi - n21 - n22;
      const int n01 = vbj - n11 - n21, n02 = vcj - n12 - n22;
      const int n00 = K - bvi - cvi - n01 - n02;
      const int c[9] = {n22, n21, n20, n12, n11, n10, n02, n01, n00};
      geno_pack_slot(w, r, c);
    }
  }
}

// ---------------------------------------------------------------------------
// Traits: what a tile's counts are and how a pair's statistics follow.
// ---------------------------------------------------------------------------
// Entry rt_off + max(n - rt_lo, 0): n below the table's range only occurs in
// padded slots (n = 0), whose statistics are weighted out.  FOLD adds the
// offset before the clamp (two instructions fewer per population and pair;
// the four-population kernel keeps the other form, which it spills less with).
template <bool FOLD = true>
__device__ __forceinline__ const double* rt_entry(const FusedParams& a, const double* srt, int p,
                                                  int n) {
  if constexpr (FOLD) {
    const int e = n + (a.rt_off[p] - a.rt_lo[p]);
    return srt + 4 * (e < a.rt_off[p] ? a.rt_off[p] : e);
  } else {
    int e = n - a.rt_lo[p];
    e = e < 0 ? 0 : e;
    return srt + 4 * (a.rt_off[p] + e);
  }
}

// Statistic groups of a traits class T: groups 0 .. NG-1 hold the LD
// statistics (GS each, consecutive), and with R2 one more group holds
// (sum of r^2 over pairs where it is defined, number of such pairs) of the
// first population.  Output row per bin: NSO = NS (+2) sums, then the pair count.
template <int P, bool R2 = false>
struct HapTraits {
  static constexpr int NP = P;
  static constexpr bool WITH_R2 = R2;
  // Workgroups per CU (waves per SIMD): as many as the live state allows
  // without spilling in the tile loop -- three for one population (168
  // registers), two for two (256), one for three or four populations, which
  // keep more counts and use the full register file.
  static constexpr int MINB = P == 1 ? 3 : P == 2 ? 2 : 1;
  static constexpr int NS = P == 1 ? HAP_NSTATS_P1 : P == 2 ? HAP_NSTATS_P2 : P == 3 ? HAP_NSTATS_P3
                                                                                     : HAP_NSTATS_P4;
  static constexpr int NG = P == 1 ? HAP_NGROUPS_P1 : P == 2 ? HAP_NGROUPS_P2 : P == 3 ? HAP_NGROUPS_P3
                                                                                       : HAP_NGROUPS_P4;
  // With one statistic group the r^2 pair joins it (statistics NS, NS+1),
  // so the population bases are derived once per pair; otherwise it is a
  // group of its own after the NG groups.
  static constexpr bool MERGE_R2 = R2 && (P == 1 ? HAP_NGROUPS_P1 : 2) == 1;
  static constexpr int GS = (P == 1 ? HAP_GS_P1 : P == 2 ? HAP_GS_P2 : P == 3 ? HAP_GS_P3
                                                                          : HAP_GS_P4) +
                            (MERGE_R2 ? 2 : 0);  // per group
  static constexpr int NSO = NS + (R2 ? 2 : 0);
  static constexpr int NGT = NG + (R2 && !MERGE_R2 ? 1 : 0);
  // k-steps of the LDS operand ring per wave (0: register ring)
  static constexpr int RING = P >= 4 ? 4 : 0;
  static constexpr int RING_BYTES = RING * 4 * 1024;  // per wave
  // the group's sums hold statistic q times 2^e (raw numerators; 0 past the
  // statistics and for the r^2 slots)
  template <int G>
  __device__ __forceinline__ static int scale_exp(int q) {
    if constexpr (G < NG) return hap_scale_group<P, G>(q);
    else return 0;
  }
  struct Counts { v16u w0[P], w1[P]; };  // slot words of tile_counts<., RING>
  // population p's basis for slot r
  __device__ __forceinline__ static HapPop pop(const FusedParams& a, const double* srt,
                                               const Counts& c, int p, int r) {
    const uint32_t w0 = c.w0[p][r], w1 = c.w1[p][r];
    if constexpr (MINB < 3) {  // tile_counts' MARG
      const int vv = (int)(w1 >> 16);
      return hap_pop_marg((int)(w0 & 0xffff), (int)(w0 >> 16), (int)(w1 & 0xffff), vv,
                          rt_entry(a, srt, p, vv));
    } else {
      const int n11 = w0 & 0xffff, n10 = w0 >> 16, n01 = w1 & 0xffff, n00 = w1 >> 16;
      return hap_pop(n11, n10, n01, n00, 0,
                     rt_entry<(P < 4)>(a, srt, p, n11 + n10 + n01 + n00));
    }
  }
  __device__ __forceinline__ static void counts(const FusedParams& a, int64_t rb, int64_t cb, int lane,
                                                Counts& c, v4i* ring) {
#pragma unroll
    for (int p = 0; p < P; ++p)
      tile_counts<(MINB >= 3), RING>(a, p, rb, cb, lane, c.w0[p], c.w1[p], ring);
  }
  // statistics G*GS .. G*GS+GS-1 of the pair in accumulator slot r
  template <int G>
  __device__ __forceinline__ static void stats(const FusedParams& a, const double* srt, const Counts& c, int r,
                               double* x) {
    HapPop pp[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
      pp[p] = pop(a, srt, c, p, r);
    }
    if constexpr (G < NG) {
      hap_stats_group<P, G>(pp, x);
      if constexpr (MERGE_R2) hap_r2(pp[0], x + GS - 2);
    } else {
      hap_r2(pp[0], x);
    }
  }
  // acc[q] += wt * statistic G*GS+q (wt 0 or 1)
  template <int G>
  __device__ __forceinline__ static void accum(const FusedParams& a, const double* srt,
                                               const Counts& c, int r, double* acc, double wt) {
    if constexpr (G < NG) {
      HapPop pp[P];
#pragma unroll
      for (int p = 0; p < P; ++p) {
        pp[p] = pop(a, srt, c, p, r);
      }
      hap_acc_group<P, G>(pp, acc, wt);
      if constexpr (MERGE_R2) {
        double x[2];
        hap_r2(pp[0], x);
        acc[GS - 2] = fma(x[0], wt, acc[GS - 2]);
        acc[GS - 1] = fma(x[1], wt, acc[GS - 1]);
      }
    } else {
      double x[2];
      stats<G>(a, srt, c, r, x);
      acc[0] = fma(x[0], wt, acc[0]);
      acc[1] = fma(x[1], wt, acc[1]);
    }
  }
};

// W = 3: nine 10-bit counts in three words per slot (population sizes up to
// 1023); W = 5: 16-bit counts in five words (up to 32767).
template <int P, int W, bool R2 = false>
struct GenoTraits {
  static constexpr int NP = P;
  static constexpr bool WITH_R2 = R2;
  // One population fits two waves per SIMD; more need the full register file
  // of a one-wave-per-SIMD launch.
  static constexpr int MINB = P == 1 ? 2 : 1;
  static constexpr int NS = P == 1 ? GENO_NSTATS_P1 : P == 2 ? GENO_NSTATS_P2 : P == 3 ? GENO_NSTATS_P3
                                                                                        : GENO_NSTATS_P4;
  static constexpr int NG = P == 1 ? GENO_NGROUPS_P1 : P == 2 ? GENO_NGROUPS_P2 : P == 3 ? GENO_NGROUPS_P3
                                                                                          : GENO_NGROUPS_P4;
  // the r^2 pair joins a single statistic group (as in HapTraits)
  static constexpr bool MERGE_R2 = R2 && (P == 1 ? GENO_NGROUPS_P1 : 2) == 1;
  static constexpr int GS = (P == 1 ? GENO_GS_P1 : P == 2 ? GENO_GS_P2 : P == 3 ? GENO_GS_P3
                                                                           : GENO_GS_P4) +
                            (MERGE_R2 ? 2 : 0);
  static constexpr int NSO = NS + (R2 ? 2 : 0);
  static constexpr int NGT = NG + (R2 && !MERGE_R2 ? 1 : 0);
  static constexpr int RING = 0;
  static constexpr int RING_BYTES = 0;
  template <int G>
  __device__ __forceinline__ static int scale_exp(int) { return 0; }
  struct Counts { v16u w[P][W]; };
  __device__ __forceinline__ static void counts(const FusedParams& a, int64_t rb, int64_t cb, int lane,
                                                Counts& c, v4i*) {
#pragma unroll
    for (int p = 0; p < P; ++p) geno_tile_counts<W>(a, p, rb, cb, lane, c.w[p]);
  }
  template <int G>
  __device__ __forceinline__ static void stats(const FusedParams& a, const double* srt, const Counts& c, int r,
                               double* x) {
    GenoPop pp[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
      int g[9];
      if constexpr (W == 3) {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          const uint32_t v = c.w[p][q][r];
          g[3 * q] = v & 0x3ff;
          g[3 * q + 1] = (v >> 10) & 0x3ff;
          g[3 * q + 2] = v >> 20;
        }
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const uint32_t v = c.w[p][q][r];
          g[2 * q] = v & 0xffff;
          g[2 * q + 1] = v >> 16;
        }
        g[8] = c.w[p][4][r];
      }
      int y[9];
      geno_centered(g, y);  // y[0] = valid individuals
      pp[p] = geno_pop(g, y, 0, rt_entry(a, srt, p, y[0]));
    }
    if constexpr (G < NG) {
      geno_stats_group<P, G>(pp, x);
      if constexpr (MERGE_R2) geno_r2(pp[0], x + GS - 2);
    } else {
      geno_r2(pp[0], x);
    }
  }
  template <int G>
  __device__ __forceinline__ static void accum(const FusedParams& a, const double* srt,
                                               const Counts& c, int r, double* acc, double wt) {
    constexpr int GSZ = G < NG ? GS : 2;
    double x[GSZ];
    stats<G>(a, srt, c, r, x);
#pragma unroll
    for (int q = 0; q < GSZ; ++q) acc[q] = fma(x[q], wt, acc[q]);
  }
};

// ---------------------------------------------------------------------------
// Binning.  Each wave owns an FP64 staging array stage[bin][slot] for the GS
// statistics of the group being evaluated, and a pair counter per bin, in
// LDS.  Inside a tile, lanes add their running per-bin sums to the staging
// array (LDS FP64 atomics, wave-private region); at the end of the group's
// pass the wave converts the touched bins to 3-limb fixed point, adds them to
// the workgroup's integer limbs and clears the staging entries.  The FP64
// rounding is thus confined to the 1024 pairs of one tile, and the integer
// limb sums across tiles, waves and workgroups are exact, so the result does
// not depend on how tiles are dealt to waves.
// ---------------------------------------------------------------------------
template <int GS, bool COUNT>
__device__ __forceinline__ void stage_add(double* stg, int* scnt, int bin, const double* acc,
                                          int count) {
  double* d = stg + bin * GS;
#pragma unroll
  for (int q = 0; q < GS; ++q) atomicAdd(d + q, acc[q]);
  if constexpr (COUNT) atomicAdd(scnt + bin, count);
}

// Move staged bins [blo, bhi] of this wave into the workgroup limbs of
// statistics s0 .. s0+GS-1 (and the pair counters).
template <class T, int G, int NS, int GS>
__device__ __forceinline__ void stage_drain(double* stg, int* scnt, unsigned long long* slimb,
                                            int lane, int blo, int bhi, int s0) {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  const int E = (bhi - blo + 1) * (GS + 1);
  for (int e = lane; e < E; e += 64) {
    const int b = blo + e / (GS + 1), q = e - (b - blo) * (GS + 1);
    if (q < GS) {
      const double v = stg[b * GS + q];
      if (v != 0.0) {
        stg[b * GS + q] = 0.0;
        unsigned long long* dst = slimb + ((int64_t)b * (NS + 1) + s0 + q) * 3;
        long long x2, x1, x0;
        pgld_to_limbs(v, x2, x1, x0, pgld_limb_scale(T::template scale_exp<G>(q)));
        atomicAdd(dst + 0, (unsigned long long)x0);
        atomicAdd(dst + 1, (unsigned long long)x1);
        atomicAdd(dst + 2, (unsigned long long)x2);
      }
    } else {
      const int c = scnt[b];
      if (c) {
        scnt[b] = 0;
        atomicAdd(slimb + ((int64_t)b * (NS + 1) + NS) * 3, (unsigned long long)c);
      }
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
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

// Signed coordinate difference c[j] - c[i]; for i < j it is the pair's
// distance (coordinates are non-decreasing along the sites), and over a tile
// it is smallest at (i1, j0) and largest at (i0, j1).
__device__ __forceinline__ double coord_diff(const FusedParams& a, int64_t i, int64_t j) {
  return a.coord_is_pos ? (a.pos[j] - a.pos[i]) : (a.coord[j] - a.coord[i]);
}

struct TileCtx {
  const double* sedges;
  const double* srt;
  double* stg;   // this wave's staging [nb][GS]
  int* scw;      // this wave's pair counters [nb]
  unsigned long long* slimb;
  int lane, bmin, bmax;
  bool fast;   // every pair valid and in bin bmin
  bool dual;   // otherwise: every valid pair in bin blo or blo + 1
  int blo;
  int64_t i0, j0;
};

// Wave totals of a group's per-lane sums for one bin, added to the workgroup
// limbs (lane q < GS: statistic S0+q); with COUNT, `pairs` joins the bin's
// pair counter.
template <class T, int G, int NS, int GS, int S0, bool COUNT>
__device__ __forceinline__ void reduce_to_limbs(const TileCtx& t, double (&acc)[GS], int bin,
                                                unsigned long long pairs) {
  const int lane = t.lane;
  // one butterfly reduce-scatter of the GS sums (padded to NV): lane l ends
  // with statistic bitreverse(l mod NV); the first padding slot carries the
  // pair count
  constexpr int NV = GS <= 2 ? 2 : GS <= 4 ? 4 : GS <= 8 ? 8 : GS <= 16 ? 16 : GS <= 32 ? 32 : 64;
  constexpr int LOG = NV == 2 ? 1 : NV == 4 ? 2 : NV == 8 ? 3 : NV == 16 ? 4 : NV == 32 ? 5 : 6;
  static_assert(GS <= 64 && (!COUNT || GS < NV), "group size and count slot");
  double v[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = q < GS ? acc[q] : 0.0;
  const double tot = pgld_wave_sums<NV>(v, lane);
  const int q = pgld_bitrev(lane & (NV - 1), LOG);
  if (lane < NV && q < GS) {
    unsigned long long* dst = t.slimb + ((int64_t)bin * (NS + 1) + S0 + q) * 3;
    long long x2, x1, x0;
    pgld_to_limbs(tot, x2, x1, x0, pgld_limb_scale(T::template scale_exp<G>(q)));
    atomicAdd(dst + 0, (unsigned long long)x0);
    atomicAdd(dst + 1, (unsigned long long)x1);
    atomicAdd(dst + 2, (unsigned long long)x2);
  } else if (COUNT && lane < NV && q == GS) {
    atomicAdd(t.slimb + ((int64_t)bin * (NS + 1) + NS) * 3, pairs);
  }
}

// Bins of the lane's 16 pairs of a tile that is not single-bin: 8 bits per
// slot r in word r >> 2, 0xFF where the pair is not counted (i >= j, beyond
// the cutoff, outside the bins).  All loads are issued before the epilogue.
__device__ __forceinline__ void tile_bins(const FusedParams& a, const TileCtx& t,
                                          uint32_t (&bw)[4]) {
  const int nb = a.n_bins, lane = t.lane;
  const int64_t j = t.j0 + (lane & 31);
  const bool jok = j < a.V;
  const double pos_j = jok ? a.pos[j] : 0.0;
  const double co_j = (jok && !a.coord_is_pos) ? a.coord[j] : pos_j;
  int guess = nb - 1;
  // row i = ib + o with o = 8q + k a constant (mfma_row(lane, 4q + k)): the
  // loads take immediate offsets from one base and i < j compares o with jr
  const int64_t ib = t.i0 + 4 * (lane >> 5);
  const double* pos_i = a.pos + ib;
  const double* co_i = a.coord + ib;
  const int jr = (int)(j - ib < 32 ? j - ib : 32);  // o < jr <=> i < j (o < 32)
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    uint32_t w = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int o = 8 * q + k;
      uint32_t b8 = 0xFF;
      if (jok && o < jr) {
        const double pi = pos_i[o];
        const double d = a.coord_is_pos ? (pos_j - pi) : fabs(co_j - co_i[o]);
        const int b = pgld_bin_scan(d, t.sedges, nb, guess, nb - 1);
        if (pos_j <= pi + a.max_dist && b >= 0 && b < nb) {
          b8 = (uint32_t)b;
          guess = b;
        }
      }
      w |= b8 << (8 * k);
    }
    bw[q] = w;
  }
}

// One pass over the tile's 1024 pairs for statistic group G.  Single-bin
// tiles are reduced to limbs at the tile end, so FP64 rounding never spans
// more than one tile (and the result does not depend on the schedule).
template <class T, int G>
__device__ __forceinline__ void tile_group(const FusedParams& a, const TileCtx& t,
                                           const typename T::Counts& cts) {
  // NS below is the output row width; GS the size of this group, S0 its start
  constexpr int NS = T::NSO;
  constexpr int GS = G < T::NG ? T::GS : 2;
  constexpr int S0 = G < T::NG ? G * T::GS : T::NS;
  // slot loops unrolled by two (two pairs' FP64 chains interleave) for the
  // traits that do not spill with it: one population, or two haplotype
  // populations at two waves per SIMD
  constexpr bool U2 = T::NP == 1 || (T::NP == 2 && T::MINB == 2);
  const int nb = a.n_bins, lane = t.lane;
  double acc[GS];
#pragma unroll
  for (int q = 0; q < GS; ++q) acc[q] = 0.0;
  if (t.fast) {
    // every pair of the tile is valid and falls in bin bmin
    if constexpr (U2) {  // 2 x 500 haplotypes, 200k sites: 3.21 -> 3.18 ms
#pragma unroll 2
      for (int r = 0; r < 16; ++r) T::template accum<G>(a, t.srt, cts, r, acc, 1.0);
    } else {
#pragma unroll 1
      for (int r = 0; r < 16; ++r) T::template accum<G>(a, t.srt, cts, r, acc, 1.0);
    }
    reduce_to_limbs<T, G, NS, GS, S0, G == 0>(t, acc, t.bmin, 1024ull);
  } else if (t.dual) {
    // Two bins: each pair's statistics go to both accumulators with 0/1
    // weights (the estimators are finite for every slot: counts of padded
    // sites are zero and the reciprocal table covers every n), so the wave
    // does the fast path's work plus one FMA per statistic, and the per-bin
    // pair counts are ballots.
    uint32_t bw[4];
    tile_bins(a, t, bw);
    double acc1[GS];
#pragma unroll
    for (int q = 0; q < GS; ++q) acc1[q] = 0.0;
    unsigned long long cnt0 = 0, cnt1 = 0;
    if constexpr (U2) {
#pragma unroll 2
      for (int r = 0; r < 16; ++r) {
        const int rq = r >> 2;
        const uint32_t w = rq == 0 ? bw[0] : rq == 1 ? bw[1] : rq == 2 ? bw[2] : bw[3];
        const int b = (int)((w >> (8 * (r & 3))) & 0xFFu);
        const bool in0 = b == t.blo, in1 = b == t.blo + 1;
        if constexpr (G == 0) {
          cnt0 += (unsigned long long)__popcll(__ballot(in0));
          cnt1 += (unsigned long long)__popcll(__ballot(in1));
        }
        T::template accum<G>(a, t.srt, cts, r, acc, in0 ? 1.0 : 0.0);
        T::template accum<G>(a, t.srt, cts, r, acc1, in1 ? 1.0 : 0.0);
      }
    } else {
#pragma unroll 1
      for (int r = 0; r < 16; ++r) {
        const int rq = r >> 2;
        const uint32_t w = rq == 0 ? bw[0] : rq == 1 ? bw[1] : rq == 2 ? bw[2] : bw[3];
        const int b = (int)((w >> (8 * (r & 3))) & 0xFFu);
        const bool in0 = b == t.blo, in1 = b == t.blo + 1;
        if constexpr (G == 0) {
          cnt0 += (unsigned long long)__popcll(__ballot(in0));
          cnt1 += (unsigned long long)__popcll(__ballot(in1));
        }
        T::template accum<G>(a, t.srt, cts, r, acc, in0 ? 1.0 : 0.0);
        T::template accum<G>(a, t.srt, cts, r, acc1, in1 ? 1.0 : 0.0);
      }
    }
    reduce_to_limbs<T, G, NS, GS, S0, G == 0>(t, acc, t.blo, cnt0);
    if (t.blo + 1 < nb) reduce_to_limbs<T, G, NS, GS, S0, G == 0>(t, acc1, t.blo + 1, cnt1);
  } else {
    uint32_t bw[4];
    tile_bins(a, t, bw);
    int cur = -1, count = 0;
#pragma unroll 1
    for (int r = 0; r < 16; ++r) {
      const int rq = r >> 2;
      const uint32_t w = rq == 0 ? bw[0] : rq == 1 ? bw[1] : rq == 2 ? bw[2] : bw[3];
      const int b = (int)((w >> (8 * (r & 3))) & 0xFFu);
      if (b != 0xFF) {
        double x[GS];
#pragma unroll
        for (int q = 0; q < GS; ++q) x[q] = 0.0;
        T::template accum<G>(a, t.srt, cts, r, x, 1.0);
        if (b != cur) {
          if (cur >= 0) stage_add<GS, G == 0>(t.stg, t.scw, cur, acc, count);
          cur = b;
          count = 0;
#pragma unroll
          for (int q = 0; q < GS; ++q) acc[q] = 0.0;
        }
#pragma unroll
        for (int q = 0; q < GS; ++q) acc[q] += x[q];
        ++count;
      }
    }
    if (cur >= 0) stage_add<GS, G == 0>(t.stg, t.scw, cur, acc, count);
    stage_drain<T, G, NS, GS>(t.stg, t.scw, t.slimb, lane, t.bmin < 0 ? 0 : t.bmin,
                        t.bmax >= nb ? nb - 1 : t.bmax, S0);
  }
}

template <class T, int G>
__device__ __forceinline__ void tile_groups(const FusedParams& a, const TileCtx& t,
                                            const typename T::Counts& cts) {
  if constexpr (G < T::NGT) {
    tile_group<T, G>(a, t, cts);
    tile_groups<T, G + 1>(a, t, cts);
  }
}

template <class T>
__global__ void __launch_bounds__(256, T::MINB) fused_kernel(FusedParams a) {
  constexpr int NS = T::NSO, GS = T::GS;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  double* sedges = reinterpret_cast<double*>(smem);
  double* srt = sedges + PGLD_MAX_BINS + 8;
  const int nb = a.n_bins;
  double* sstage = srt + 4 * a.rt_len;                       // [4 waves][nb][GS]
  unsigned long long* slimb =
      reinterpret_cast<unsigned long long*>(sstage + 4 * nb * GS);  // [nb][NS+1][3]
  int* scnt = reinterpret_cast<int*>(slimb + nb * (NS + 1) * 3);    // [4 waves][nb]
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  // this wave's operand ring (16-byte aligned, after the tables)
  v4i* ring = nullptr;
  if (T::RING > 0 && a.use_ring) {
    const uintptr_t end = reinterpret_cast<uintptr_t>(scnt + 4 * nb);
    ring = reinterpret_cast<v4i*>((end + 15) & ~uintptr_t(15)) + wave * (T::RING_BYTES / 16);
  }
  for (int e = tid; e <= nb; e += blockDim.x) sedges[e] = a.edges[e];
  for (int p = 0; p < T::NP; ++p) {
    const int len = a.K[p] - a.rt_lo[p] + 1;
    for (int e = tid; e < 4 * len; e += blockDim.x)
      srt[4 * a.rt_off[p] + e] = a.rtab[4 * a.rt_lo[p] + e];
  }
  for (int e = tid; e < 4 * nb * GS; e += blockDim.x) sstage[e] = 0.0;
  for (int e = tid; e < nb * (NS + 1) * 3; e += blockDim.x) slimb[e] = 0ull;
  for (int e = tid; e < 4 * nb; e += blockDim.x) scnt[e] = 0;
  __syncthreads();
  TileCtx t;
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
    T::counts(a, rb, cb, lane, cts, ring);
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
    t.bmin = pgld_bin_scan(d_min, sedges, nb, 0, nb - 1);
    t.bmax = pgld_bin_scan(d_max, sedges, nb, t.bmin < 0 ? 0 : t.bmin, nb - 1);
    t.fast = (cb > rb) && (j1 == j0 + 31) && in_cut &&
             t.bmin == t.bmax && t.bmin >= 0 && t.bmin < nb;
    t.blo = t.bmin < 0 ? 0 : t.bmin;
    t.dual = !t.fast && t.blo < nb && (t.bmax >= nb ? nb - 1 : t.bmax) - t.blo <= 1;
    tile_groups<T, 0>(a, t, cts);
    local = next;
  }
  // nothing may stay staged
  stage_drain<T, 0, NS, GS>(t.stg, t.scw, slimb, lane, 0, nb - 1, 0);
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

template <class T>
static int launch_fused(const PanelLayout& L, const pgld_binned_args* g, hipStream_t s) {
  constexpr int NS = T::NSO;
  constexpr int P = T::NP;
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
  const size_t fixed = (PGLD_MAX_BINS + 8) * sizeof(double) + (size_t)off * 4 * sizeof(double);
  const size_t per_bin = (size_t)4 * T::GS * sizeof(double) +
                         (size_t)(NS + 1) * 3 * sizeof(unsigned long long) + 4 * sizeof(int);
  if (fixed + per_bin > kLds) return fail(PGLD_EINVAL, "fused kernel: LDS tables exceed 160 KiB");
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
    size_t lds = fixed + (size_t)c.n_bins * per_bin;
    // the operand ring only where the tables leave room for it
    const size_t ring_bytes = T::RING > 0 ? 16 + (size_t)4 * T::RING_BYTES : 0;
    c.use_ring = ring_bytes > 0 && lds + ring_bytes <= kLds;
    if (c.use_ring) lds += ring_bytes;
    PGLD_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(fused_kernel<T>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(fused_kernel<T>, dim3(grid), dim3(256), lds, s, c);
    PGLD_LAUNCH_CHECK();
  }
  return PGLD_OK;
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
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  v4i* ring = nullptr;
  if (T::RING > 0 && a.use_ring) ring = reinterpret_cast<v4i*>(smem) + wave * (T::RING_BYTES / 16);
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
    T::counts(a, rb, cb, lane, cts, ring);
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
  const size_t lds = T::RING > 0 ? (size_t)4 * T::RING_BYTES : 0;
  a.use_ring = T::RING > 0;
  if (lds) {
    PGLD_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(count_band_kernel<T>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  }
  int grid = g->grid > 0 ? g->grid : T::MINB * num_cus();
  grid = ((grid + 7) / 8) * 8;
  const int64_t need = ((g->n_tiles + 31) / 32) * 8;
  if (grid > need) grid = (int)need;
  if (grid < 8) grid = 8;
  hipLaunchKernelGGL(count_band_kernel<T>, dim3(grid), dim3(256), lds, s, a, sums);
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
};

template <class T, int G>
__device__ __forceinline__ void window_group(const WindowParams& w, const double* srt,
                                             const typename T::Counts& cts, int lane, bool full,
                                             int64_t i0, int64_t j0, int64_t lo, int64_t hi,
                                             unsigned long long* dst_w) {
  constexpr int NS = T::NSO;
  constexpr int GS = G < T::NG ? T::GS : 2;
  constexpr int S0 = G < T::NG ? G * T::GS : T::NS;
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
  // two slots per iteration: two pairs' FP64 chains interleave
#pragma unroll 2
  for (int r = 0; r < 16; ++r) {
    const int o = 8 * (r >> 2) + (r & 3);
    // every slot's statistics are finite (padded sites have zero counts), so
    // slots outside the window take weight 0 instead of a divergent branch
    const bool in = full || (o >= lo_o && o < j_o && jin);
    T::template accum<G>(w.f, srt, cts, r, acc, in ? 1.0 : 0.0);
    count += in;
  }
  // butterfly reduce-scatter (see reduce_to_limbs); for group 0 the pair
  // count rides in the first padding slot (an exact small integer)
  constexpr int NV = GS <= 2 ? 2 : GS <= 4 ? 4 : GS <= 8 ? 8 : GS <= 16 ? 16 : GS <= 32 ? 32 : 64;
  constexpr int LOG = NV == 2 ? 1 : NV == 4 ? 2 : NV == 8 ? 3 : NV == 16 ? 4 : NV == 32 ? 5 : 6;
  static_assert(G != 0 || GS < NV, "count slot");
  double v[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = q < GS ? acc[q] : (q == GS && G == 0 ? (double)count : 0.0);
  const double tot = pgld_wave_sums<NV>(v, lane);
  const int q = pgld_bitrev(lane & (NV - 1), LOG);
  if (lane < NV && q < GS) {
    unsigned long long* dst = dst_w + (int64_t)(S0 + q) * 3;
    long long x2, x1, x0;
    pgld_to_limbs(tot, x2, x1, x0, pgld_limb_scale(T::template scale_exp<G>(q)));
    if (x0) atomicAdd(dst + 0, (unsigned long long)x0);
    if (x1) atomicAdd(dst + 1, (unsigned long long)x1);
    if (x2) atomicAdd(dst + 2, (unsigned long long)x2);
  } else if (G == 0 && lane < NV && q == GS) {
    const unsigned long long c = (unsigned long long)tot;
    if (c) atomicAdd(dst_w + (int64_t)NS * 3, c);
  }
}

template <class T, int G>
__device__ __forceinline__ void window_groups(const WindowParams& w, const double* srt,
                                              const typename T::Counts& cts, int lane, bool full,
                                              int64_t i0, int64_t j0, int64_t lo, int64_t hi,
                                              unsigned long long* dst_w) {
  if constexpr (G < T::NGT) {
    window_group<T, G>(w, srt, cts, lane, full, i0, j0, lo, hi, dst_w);
    window_groups<T, G + 1>(w, srt, cts, lane, full, i0, j0, lo, hi, dst_w);
  }
}

template <class T>
__global__ void __launch_bounds__(256, T::MINB) window_kernel(WindowParams w) {
  constexpr int NS = T::NSO;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  double* srt = reinterpret_cast<double*>(smem);
  const FusedParams& a = w.f;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  for (int p = 0; p < T::NP; ++p) {
    const int len = a.K[p] - a.rt_lo[p] + 1;
    for (int e = tid; e < 4 * len; e += blockDim.x)
      srt[4 * a.rt_off[p] + e] = a.rtab[4 * a.rt_lo[p] + e];
  }
  __syncthreads();
  // XCD-aware split as in fused_kernel: each XCD walks a contiguous stretch
  // of the work list, whose consecutive items share a window's panels.
  const int nx = 8;
  const int xcd = blockIdx.x % nx;
  const int waves_per_xcd = (gridDim.x / nx) * 4;
  const int64_t seg_lo = w.n_work * xcd / nx, seg_hi = w.n_work * (xcd + 1) / nx;
  uint32_t* ctr = a.sched ? a.sched + xcd * 32 : nullptr;
  for (int64_t local = (blockIdx.x / nx) * 4 + wave; seg_lo + local < seg_hi;) {
    const int64_t k = seg_lo + local;
    uint32_t claim = 0;
    if (ctr && lane == 0) claim = atomicAdd(ctr, 1u);
    const int64_t win = w.work[3 * k], rb = w.work[3 * k + 1], cb = w.work[3 * k + 2];
    const int64_t lo = w.win[2 * win], hi = w.win[2 * win + 1];
    typename T::Counts cts;
    T::counts(a, rb, cb, lane, cts, nullptr);  // one population: no LDS ring
    const int64_t i0 = rb * 32, j0 = cb * 32;
    const bool full = cb > rb && i0 >= lo && j0 + 32 <= hi;
    window_groups<T, 0>(w, srt, cts, lane, full, i0, j0, lo, hi,
                        w.out + (int64_t)win * (NS + 1) * 3);
    local = ctr ? waves_per_xcd + (int64_t)__builtin_amdgcn_readfirstlane(__shfl(claim, 0, 64))
                : local + waves_per_xcd;
  }
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
  const size_t lds = (size_t)a.rt_len * 4 * sizeof(double);
  if (lds > 160 * 1024) return fail(PGLD_EINVAL, "window kernel: reciprocal table exceeds LDS");
  PGLD_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(window_kernel<T>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  int grid = g->grid > 0 ? g->grid : T::MINB * num_cus();
  grid = ((grid + 7) / 8) * 8;
  const int64_t need = ((g->n_work + 31) / 32) * 8;
  if (grid > need) grid = (int)need;
  if (grid < 8) grid = 8;
  hipLaunchKernelGGL(window_kernel<T>, dim3(grid), dim3(256), lds, s, w);
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
  if (g->with_r2 != 0 && g->with_r2 != 1) return fail(PGLD_EINVAL, "with_r2 must be 0 or 1");
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
  if (args->with_r2) {  // one population (checked above)
    if (desc->mode == PGLD_MODE_HAP) return launch_fused<HapTraits<1, true>>(L, args, s);
    return L.pop_size[0] <= 1023 ? launch_fused<GenoTraits<1, 3, true>>(L, args, s)
                                 : launch_fused<GenoTraits<1, 5, true>>(L, args, s);
  }
  if (desc->mode == PGLD_MODE_HAP) {
    switch (desc->n_pops) {
      case 1: return launch_fused<HapTraits<1>>(L, args, s);
      case 2: return launch_fused<HapTraits<2>>(L, args, s);
      case 3: return launch_fused<HapTraits<3>>(L, args, s);
      case 4: return launch_fused<HapTraits<4>>(L, args, s);
      default: break;
    }
  } else {
    bool narrow = true;
    for (int p = 0; p < desc->n_pops; ++p) narrow = narrow && L.pop_size[p] <= 1023;
    switch (desc->n_pops) {
      case 1:
        return narrow ? launch_fused<GenoTraits<1, 3>>(L, args, s)
                      : launch_fused<GenoTraits<1, 5>>(L, args, s);
      case 2:
        return narrow ? launch_fused<GenoTraits<2, 3>>(L, args, s)
                      : launch_fused<GenoTraits<2, 5>>(L, args, s);
      case 3:
        return narrow ? launch_fused<GenoTraits<3, 3>>(L, args, s)
                      : launch_fused<GenoTraits<3, 5>>(L, args, s);
      case 4:
        return narrow ? launch_fused<GenoTraits<4, 3>>(L, args, s)
                      : launch_fused<GenoTraits<4, 5>>(L, args, s);
      default: break;
    }
  }
  return fail(PGLD_EUNSUPPORTED, "fused kernel: this (mode, n_pops) is not built yet");
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
  if (args->n_tiles < 0) return fail(PGLD_EINVAL, "negative n_tiles");
  if (args->n_tiles == 0) return PGLD_OK;
  hipStream_t s = as_stream(stream);
  switch (desc->n_pops) {
    case 1: return launch_count_band<HapTraits<1>>(L, args, d_sums, s);
    case 2: return launch_count_band<HapTraits<2>>(L, args, d_sums, s);
    case 3: return launch_count_band<HapTraits<3>>(L, args, d_sums, s);
    case 4: return launch_count_band<HapTraits<4>>(L, args, d_sums, s);
    default: break;
  }
  return fail(PGLD_EINVAL, "1-4 populations supported");
}

extern "C" int pgld_windows_fused(const pgld_panel_desc* desc, const pgld_window_args* args,
                                  void* stream) {
  PanelLayout L;
  int rc = make_layout(desc, &L);
  if (rc) return rc;
  if (!args) return fail(PGLD_EINVAL, "null window args");
  if (desc->n_pops != 1) return fail(PGLD_EUNSUPPORTED, "window sums are built for one population");
  if (args->n_work < 0 || args->n_windows < 0) return fail(PGLD_EINVAL, "negative sizes");
  if (args->n_work == 0) return PGLD_OK;
  if (!args->d_panels || !args->d_site_stats || !args->d_rtab || !args->d_work ||
      !args->d_windows || !args->d_out)
    return fail(PGLD_EINVAL, "null device pointer in window args");
  hipStream_t s = as_stream(stream);
  if (desc->mode == PGLD_MODE_HAP) return launch_windows<HapTraits<1, true>>(L, args, s);
  return L.pop_size[0] <= 1023 ? launch_windows<GenoTraits<1, 3, true>>(L, args, s)
                               : launch_windows<GenoTraits<1, 5, true>>(L, args, s);
}

