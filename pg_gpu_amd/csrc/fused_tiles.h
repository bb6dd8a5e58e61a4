// Tile machinery of the fused binned and window kernels (fused.hip):
// kernel parameters, MFMA tile counting, the haplotype / genotype traits
// (how a tile's counts are produced and how a pair's statistics follow), the
// per-bin staging and reduction, and the per-tile statistic passes.
#pragma once
#include <algorithm>
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

// k-steps of operand loads in flight in the missing-data counting loop
#ifndef PGLD_PF_MISSING
#define PGLD_PF_MISSING 3
#endif

// hh = A.B, hv = A.Bv, vh = Av.B, vv = Av.Bv over the ks k-steps.  Operand
// loads run PF k-steps ahead in a register ring indexed by the unrolled
// position only.  The main loop covers whole groups of PF k-steps with every
// load issued unconditionally (a load past the last k-step re-reads k-step 0,
// already in L1, and is never multiplied), so the loop body has no branch and
// the ring needs no register copies; the remaining ks % PF k-steps follow.
template <int PF>
__device__ __forceinline__ void mfma_products4(const v4i* A, const v4i* Av, const v4i* B,
                                               const v4i* Bv, int ks, v16f& hh, v16f& hv,
                                               v16f& vh, v16f& vv) {
  v4i ra[PF], rav[PF], rb[PF], rbv[PF];
#pragma unroll
  for (int d = 0; d < PF; ++d) {
    const int n = d < ks ? d * 64 : 0;
    ra[d] = A[n]; rav[d] = Av[n]; rb[d] = B[n]; rbv[d] = Bv[n];
  }
  const int full = ks / PF;
  if (full > 0) {
    // the first k-step starts from zero accumulators (an inline constant of
    // the MFMA): the 64 accumulator registers need no clearing
#pragma unroll
    for (int d = 0; d < PF; ++d) {
      hh = pgld_mfma_fp4(ra[d], rb[d], d == 0 ? zero16f() : hh);
      hv = pgld_mfma_fp4(ra[d], rbv[d], d == 0 ? zero16f() : hv);
      vh = pgld_mfma_fp4(rav[d], rb[d], d == 0 ? zero16f() : vh);
      vv = pgld_mfma_fp4(rav[d], rbv[d], d == 0 ? zero16f() : vv);
      const int s = PF + d;
      const int n = s < ks ? s * 64 : 0;
      ra[d] = A[n]; rav[d] = Av[n]; rb[d] = B[n]; rbv[d] = Bv[n];
    }
  } else {
    hh = hv = vh = vv = zero16f();
  }
  for (int g = 1; g < full; ++g) {
#pragma unroll
    for (int d = 0; d < PF; ++d) {
      hh = pgld_mfma_fp4(ra[d], rb[d], hh);
      hv = pgld_mfma_fp4(ra[d], rbv[d], hv);
      vh = pgld_mfma_fp4(rav[d], rb[d], vh);
      vv = pgld_mfma_fp4(rav[d], rbv[d], vv);
      const int s = (g + 1) * PF + d;
      const int n = s < ks ? s * 64 : 0;
      ra[d] = A[n]; rav[d] = Av[n]; rb[d] = B[n]; rbv[d] = Bv[n];
    }
  }
  const int rem = ks - full * PF;
#pragma unroll
  for (int d = 0; d < PF - 1; ++d) {
    if (d < rem) {
      hh = pgld_mfma_fp4(ra[d], rb[d], hh);
      hv = pgld_mfma_fp4(ra[d], rbv[d], hv);
      vh = pgld_mfma_fp4(rav[d], rb[d], vh);
      vv = pgld_mfma_fp4(rav[d], rbv[d], vv);
    }
  }
}

// acc[3 f + g] = A_f . B_g over the ks k-steps for three features f, g
// (fragments of feature f at A + f fs): the genotype counting products.  The
// same register ring as mfma_products4, PF k-steps of six operands deep (two
// for one population; one where several populations fill the registers).
template <int PF>
__device__ __forceinline__ void mfma_products9(const v4i* A, const v4i* B, int fs, int ks,
                                               v16f (&acc)[9]) {
  v4i ra[PF][3], rb[PF][3];
#pragma unroll
  for (int d = 0; d < PF; ++d) {
    const int n = d < ks ? d * 64 : 0;
#pragma unroll
    for (int f = 0; f < 3; ++f) { ra[d][f] = A[f * fs + n]; rb[d][f] = B[f * fs + n]; }
  }
  const int full = ks / PF;
  if (full > 0) {
#pragma unroll
    for (int d = 0; d < PF; ++d) {
#pragma unroll
      for (int f = 0; f < 3; ++f)
#pragma unroll
        for (int g = 0; g < 3; ++g)
          acc[3 * f + g] = pgld_mfma_fp4(ra[d][f], rb[d][g], d == 0 ? zero16f() : acc[3 * f + g]);
      const int s = PF + d;
      const int n = s < ks ? s * 64 : 0;
#pragma unroll
      for (int f = 0; f < 3; ++f) { ra[d][f] = A[f * fs + n]; rb[d][f] = B[f * fs + n]; }
    }
  } else {
#pragma unroll
    for (int q = 0; q < 9; ++q) acc[q] = zero16f();
  }
  for (int gi = 1; gi < full; ++gi) {
#pragma unroll
    for (int d = 0; d < PF; ++d) {
#pragma unroll
      for (int f = 0; f < 3; ++f)
#pragma unroll
        for (int g = 0; g < 3; ++g)
          acc[3 * f + g] = pgld_mfma_fp4(ra[d][f], rb[d][g], acc[3 * f + g]);
      const int s = (gi + 1) * PF + d;
      const int n = s < ks ? s * 64 : 0;
#pragma unroll
      for (int f = 0; f < 3; ++f) { ra[d][f] = A[f * fs + n]; rb[d][f] = B[f * fs + n]; }
    }
  }
  const int rem = ks - full * PF;
#pragma unroll
  for (int d = 0; d < PF - 1; ++d) {
    if (d < rem) {
#pragma unroll
      for (int f = 0; f < 3; ++f)
#pragma unroll
        for (int g = 0; g < 3; ++g)
          acc[3 * f + g] = pgld_mfma_fp4(ra[d][f], rb[d][g], acc[3 * f + g]);
    }
  }
}

// H.H over the ks k-steps of row operand A and column operand B, operand
// loads two k-steps ahead.
__device__ __forceinline__ v16f hh_products(const v4i* A, const v4i* B, int ks) {
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
  return hh;
}

// Counts of a population without missing calls: n11 = H.H from the MFMAs,
// the marginals from the site tallies.  MARG packs n11 | ki << 16 and
// kj | K << 16 (ki, kj the allele counts of the row and column sites), else
// the four class counts.  Returns kj.
template <bool MARG>
__device__ __forceinline__ int tile_counts_nomiss(const FusedParams& a, int p, int64_t rb,
                                                  int64_t cb, int lane, const v4i* A,
                                                  const v4i* B, int ks, v16u& w0, v16u& w1) {
  const v16f hh = hh_products(A, B, ks);
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
  return kj;
}

// Counts of population p for tile (rb, cb), packed per slot as
// w0 = n11 | n10 << 16, w1 = n01 | n00 << 16 (or the marginals, below); PFM
// k-steps of operand loads in flight with missing calls.
template <bool TWO_PASS, int PFM = PGLD_PF_MISSING>
__device__ __forceinline__ void tile_counts(const FusedParams& a, int p, int64_t rb, int64_t cb,
                                            int lane, v16u& w0, v16u& w1) {
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
  if (a.missing[p] && !TWO_PASS) {
    v16f hh, hv, vh, vv;
    mfma_products4<PFM>(A, A + ks * 64, B, B + ks * 64, ks, hh, hv, vh, vv);
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
    tile_counts_nomiss<MARG>(a, p, rb, cb, lane, A, B, ks, w0, w1);
  }
}

// ---------------------------------------------------------------------------
// Genotype tile counts.  Features per sample (pgld_pack): u = g - 1 in
// {-1, 0, 1}, q = u^2 and, with missing calls, v = [g >= 0] (u = q = 0 where
// g is missing).  Over the samples valid at both sites the nine products are
// the pair's centred dosage moments (row site first):
//   u.u = ab11, u.q = ab12, q.u = ab21, q.q = ab22,
//   u.v = a1, q.v = a2, v.u = b1, v.q = b2, v.v = n0.
// Without missing calls the marginal sums come from the site tallies
// (a1 = 2 n2 + n1 - K, a2 = K - n1, with n1, n2 the site's counts of
// genotypes 1 and 2).  One population keeps the moments (signed fields);
// several turn them into the nine class counts (moments order n22 n21 n20
// n12 n11 n10 n02 n01 n00) with the exact inverse transform
// [g=2] = (q + u)/2, [g=0] = (q - u)/2, [g=1] = v - q.
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

// Pack the moments of slot r: W = 3 (|values| <= 1023) as 11-bit signed and
// 10-bit unsigned fields, W = 5 as 16-bit fields (a1, b1 and the ab_kl with a
// first power of u signed).
template <int W>
__device__ __forceinline__ void geno_pack_moments(v16u (&w)[W], int r, const GenoMoments& m) {
  if constexpr (W == 3) {
    w[0][r] = ((uint32_t)m.ab11 & 0x7ffu) | (((uint32_t)m.ab12 & 0x7ffu) << 11) |
              ((uint32_t)m.a2 << 22);
    w[1][r] = ((uint32_t)m.a1 & 0x7ffu) | (((uint32_t)m.ab21 & 0x7ffu) << 11) |
              ((uint32_t)m.ab22 << 22);
    w[2][r] = ((uint32_t)m.b1 & 0x7ffu) | ((uint32_t)m.b2 << 11) | ((uint32_t)m.n0 << 21);
  } else {
    w[0][r] = ((uint32_t)m.ab11 & 0xffffu) | ((uint32_t)m.ab12 << 16);
    w[1][r] = ((uint32_t)m.a1 & 0xffffu) | ((uint32_t)m.ab21 << 16);
    w[2][r] = ((uint32_t)m.b1 & 0xffffu) | ((uint32_t)m.a2 << 16);
    w[3][r] = (uint32_t)m.ab22 | ((uint32_t)m.b2 << 16);
    w[4][r] = (uint32_t)m.n0;
  }
}

__device__ __forceinline__ int geno_sfield(uint32_t w, int off, int bits) {
  return ((int)(w << (32 - off - bits))) >> (32 - bits);
}

// The moments of slot r packed by geno_pack_moments.
template <int W>
__device__ __forceinline__ GenoMoments geno_unpack_moments(const v16u (&w)[W], int r) {
  GenoMoments m;
  if constexpr (W == 3) {
    const uint32_t x = w[0][r], y = w[1][r], z = w[2][r];
    m.ab11 = geno_sfield(x, 0, 11); m.ab12 = geno_sfield(x, 11, 11); m.a2 = (int)(x >> 22);
    m.a1 = geno_sfield(y, 0, 11); m.ab21 = geno_sfield(y, 11, 11); m.ab22 = (int)(y >> 22);
    m.b1 = geno_sfield(z, 0, 11); m.b2 = (int)((z >> 11) & 0x3ffu); m.n0 = (int)(z >> 21);
  } else {
    const uint32_t x = w[0][r], y = w[1][r], z = w[2][r], t = w[3][r];
    m.ab11 = geno_sfield(x, 0, 16); m.ab12 = geno_sfield(x, 16, 16);
    m.a1 = geno_sfield(y, 0, 16); m.ab21 = geno_sfield(y, 16, 16);
    m.b1 = geno_sfield(z, 0, 16); m.a2 = (int)(z >> 16);
    m.ab22 = (int)(t & 0xffffu); m.b2 = (int)(t >> 16);
    m.n0 = (int)w[4][r];
  }
  return m;
}

// CENTRED: pack the moments (one population); else the class counts.
template <int W, int PF, bool CENTRED>
__device__ __forceinline__ void geno_tile_counts(const FusedParams& a, int p, int64_t rb,
                                                 int64_t cb, int lane, v16u (&w)[W]) {
  const int ks = a.ksteps[p];
  const v4i* A = reinterpret_cast<const v4i*>(a.panels + rb * a.block_bytes + a.pop_off[p]) + lane;
  const v4i* B = reinterpret_cast<const v4i*>(a.panels + cb * a.block_bytes + a.pop_off[p]) + lane;
  const int fs = ks * 64;  // feature stride in v4i
  auto emit = [&](int r, const GenoMoments& m) {
    if constexpr (CENTRED) {
      geno_pack_moments(w, r, m);
    } else {
      int c[9];
      geno_class_counts(m, c);
      geno_pack_slot(w, r, c);
    }
  };
  if (a.missing[p]) {
    v16f acc[9];
    mfma_products9<PF>(A, B, fs, ks, acc);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      GenoMoments m;
      m.ab11 = (int)acc[0][r]; m.ab12 = (int)acc[1][r]; m.a1 = (int)acc[2][r];
      m.ab21 = (int)acc[3][r]; m.ab22 = (int)acc[4][r]; m.a2 = (int)acc[5][r];
      m.b1 = (int)acc[6][r]; m.b2 = (int)acc[7][r]; m.n0 = (int)acc[8][r];
      emit(r, m);
    }
  } else {
    v16f uu, uq, qu, qq;
    mfma_products4<PGLD_PF_MISSING>(A, A + fs, B, B + fs, ks, uu, uq, qu, qq);
    const int32_t* st = a.site_stats + (int64_t)p * 3 * a.V;
    const int64_t j = cb * 32 + (lane & 31);
    const int K = a.K[p];
    // a padded site (past the last one) counts as all heterozygous (u = q = 0,
    // matching its zero products), so every class count stays non-negative
    const int n1j = j < a.V ? st[j] : K, n2j = j < a.V ? st[a.V + j] : 0;
    const int b1 = 2 * n2j + n1j - K, b2 = K - n1j;
    // rows as in tile_counts: i = ib + o with o constant per r
    const int64_t ib = rb * 32 + 4 * (lane >> 5);
    const int32_t* si = st + ib;
    const int32_t* sci = st + a.V + ib;
    const int lim = (int)(a.V - ib < 32 ? a.V - ib : 32);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int o = 8 * (r >> 2) + (r & 3);
      const int n1i = o < lim ? si[o] : K, n2i = o < lim ? sci[o] : 0;
      GenoMoments m;
      m.ab11 = (int)uu[r]; m.ab12 = (int)uq[r]; m.ab21 = (int)qu[r]; m.ab22 = (int)qq[r];
      m.a1 = 2 * n2i + n1i - K; m.a2 = K - n1i;
      m.b1 = b1; m.b2 = b2; m.n0 = K;
      emit(r, m);
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
  // the group's sums hold statistic q times 2^e (raw numerators; 0 past the
  // statistics and for the r^2 slots)
  template <int G>
  __device__ __forceinline__ static int scale_exp(int q) {
    if constexpr (G < NG) return hap_scale_group<P, G>(q);
    else return 0;
  }
  // factor applied to a sum before it becomes limbs (sums here are values)
  template <int G>
  __device__ __forceinline__ static double stat_mult(const FusedParams&, const double*, int) {
    return 1.0;
  }
  static constexpr bool PC_OK = true;   // has the producer/consumer form
  // single-bin tiles with the slot loop fully unrolled
  static constexpr bool FULL_FAST = false;
  // statistics of pass G, and the output column of its q-th (the r^2 pair
  // after the group's statistics, or the r^2 group, lands at NS, NS + 1)
  template <int G>
  __device__ static constexpr int gsz() {
    if constexpr (G < NG) return hap_group_size<P, G>() + (MERGE_R2 ? 2 : 0);
    else return 2;
  }
  template <int G>
  __device__ __forceinline__ static int col(int q) {
    if constexpr (G < NG) return hap_stat_col<P, G>(q);
    else return NS + q;
  }
  struct Counts { v16u w0[P], w1[P]; };  // slot words of tile_counts
  // Mailbox form of a tile's counts (the producer/consumer kernel): word k of
  // slot r of lane l at mb[(16 k + r) * 64 + l], k = 2p (w0) or 2p + 1 (w1);
  // every index a constant, so the counts never leave registers otherwise.
  static constexpr int NWORDS = 2 * P;
  __device__ __forceinline__ static void put(const Counts& c, uint32_t* mb, int lane) {
#pragma unroll
    for (int p = 0; p < P; ++p)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        mb[((2 * p) * 16 + r) * 64 + lane] = c.w0[p][r];
        mb[((2 * p + 1) * 16 + r) * 64 + lane] = c.w1[p][r];
      }
  }
  __device__ __forceinline__ static void get(Counts& c, const uint32_t* mb, int lane) {
#pragma unroll
    for (int p = 0; p < P; ++p)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        c.w0[p][r] = mb[((2 * p) * 16 + r) * 64 + lane];
        c.w1[p][r] = mb[((2 * p + 1) * 16 + r) * 64 + lane];
      }
  }
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
                                                Counts& c) {
#pragma unroll
    for (int p = 0; p < P; ++p) tile_counts<(MINB >= 3)>(a, p, rb, cb, lane, c.w0[p], c.w1[p]);
  }
  // the counting waves of the producer/consumer kernel hold no statistics
  // state, so their operand ring runs PC_PF k-steps deep
#ifndef PGLD_PC_PF_HAP
#define PGLD_PC_PF_HAP 6
#endif
  static constexpr int PC_PF = PGLD_PC_PF_HAP;
  __device__ __forceinline__ static void counts_pc(const FusedParams& a, int64_t rb, int64_t cb,
                                                   int lane, Counts& c) {
#pragma unroll
    for (int p = 0; p < P; ++p)
      tile_counts<(MINB >= 3), PC_PF>(a, p, rb, cb, lane, c.w0[p], c.w1[p]);
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

// One haplotype population without missing calls.  Every pair has n = K
// valid samples, so the reciprocal falling factorials are constants: each
// tile's statistics are accumulated as exact integer numerators (below 2^53
// for K below ~2,800 with 1,024 pairs per tile; above that the sums round
// like any FP64 sum) and multiplied by 1/(K)_4 once per sum, where it becomes
// limbs (stat_mult).  The column site's allele counts are tile constants and
// a slot word holds n11 | ki << 16 only.
template <bool R2 = false>
struct HapNMTraits : HapTraits<1, R2> {
  using Base = HapTraits<1, R2>;
  static constexpr bool PC_OK = false;
  // no class counts or reciprocal lookups: four waves per SIMD fit (<= 128
  // registers)
  static constexpr int MINB = 4;
  struct Counts {
    v16u w0[1], w1[1];     // n11 | ki << 16 (w1 unused)
    double pB, qB, uB, PB, n;
  };
  __device__ __forceinline__ static void counts(const FusedParams& a, int64_t rb, int64_t cb,
                                                int lane, Counts& c) {
    const int ks = a.ksteps[0];
    const v4i* A = reinterpret_cast<const v4i*>(a.panels + rb * a.block_bytes + a.pop_off[0]) + lane;
    const v4i* B = reinterpret_cast<const v4i*>(a.panels + cb * a.block_bytes + a.pop_off[0]) + lane;
    const int kj = tile_counts_nomiss<true>(a, 0, rb, cb, lane, A, B, ks, c.w0[0], c.w1[0]);
    c.n = pgld_count_f64(a.K[0]);
    c.pB = pgld_count_f64(kj);
    c.qB = c.n - c.pB;
    c.uB = c.qB - c.pB;
    c.PB = c.pB * c.qB;
  }
  // 1/(K)_4 for the three LD sums (raw numerators), 1 for the r^2 pair
  template <int G>
  __device__ __forceinline__ static double stat_mult(const FusedParams& a, const double* srt, int q) {
    return q < 3 ? rt_entry(a, srt, 0, a.K[0])[3] : 1.0;
  }
  // the pair of slot r (counts, its basis fields)
  __device__ __forceinline__ static HapPop pop(const Counts& c, int r) {
    const uint32_t w = c.w0[0][r];
    HapPop p;
    p.c1 = pgld_count_f64((int)(w & 0xffffu));
    p.pA = pgld_count_f64((int)(w >> 16));
    p.n = c.n; p.pB = c.pB; p.qB = c.qB; p.uB = c.uB; p.PB = c.PB;
    p.qA = p.n - p.pA;
    p.c2 = p.pA - p.c1; p.c3 = p.pB - p.c1; p.c4 = p.qA - p.c3;
    p.x = p.c1 * p.c4; p.y = p.c2 * p.c3;
    p.D = p.y - p.x;
    p.s14 = p.c1 + p.c4; p.s23 = p.c2 + p.c3;
    p.uA = p.qA - p.pA;
    p.PA = p.pA * p.qA;
    p.w = fma(p.uA, p.uB, p.s23 - p.s14);
    p.r1 = p.r2 = p.r3 = p.r4 = 0.0;   // not used: the sums are scaled by stat_mult
    return p;
  }
  // acc[q] += wt * (raw numerator of statistic q; r^2 and its defined flag)
  template <int G>
  __device__ __forceinline__ static void accum(const FusedParams&, const double*, const Counts& c,
                                               int r, double* acc, double wt) {
    static_assert(Base::NGT == 1, "one statistic pass");
    const HapPop p = pop(c, r);
    acc[0] = fma(hap_dd_same_num(p), wt, acc[0]);
    acc[1] = fma(hap_dz_same_num(p), wt, acc[1]);
    acc[2] = fma(hap_pi2_same_num(p), wt, acc[2]);
    if constexpr (R2) {
      double x[2];
      hap_r2(p, x);
      acc[3] = fma(x[0], wt, acc[3]);
      acc[4] = fma(x[1], wt, acc[4]);
    }
  }
};

// One haplotype population without missing calls, as launched (fused and
// window kernels): n11 stays in its FP32 accumulator and the row sites'
// allele counts in words of their own, so nothing is packed per tile or work
// item and nothing unpacked per pair.  Both kernels keep four waves per SIMD
// (122-127 registers); HapNMTraits' packed words remain the layout of the
// 16-site window tiles.
template <bool R2>
struct HapNMRawTraits : HapNMTraits<R2> {
  struct Counts {
    v16f n11;
    v16u ki;
    double pB, qB, uB, PB, n;
  };
  __device__ __forceinline__ static void counts(const FusedParams& a, int64_t rb, int64_t cb,
                                                int lane, Counts& c) {
    const int ks = a.ksteps[0];
    const v4i* A = reinterpret_cast<const v4i*>(a.panels + rb * a.block_bytes + a.pop_off[0]) + lane;
    const v4i* B = reinterpret_cast<const v4i*>(a.panels + cb * a.block_bytes + a.pop_off[0]) + lane;
    c.n11 = hh_products(A, B, ks);
    const int32_t* cnt1 = a.site_stats;
    const int64_t j = cb * 32 + (lane & 31);
    const int kj = j < a.V ? cnt1[j] : 0;
    // row i = ib + o, o = mfma_row(lane, r) - 4 (lane >> 5), as in tile_counts
    const int64_t ib = rb * 32 + 4 * (lane >> 5);
    const int32_t* ci = cnt1 + ib;
    const int lim = (int)(a.V - ib < 32 ? a.V - ib : 32);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int o = 8 * (r >> 2) + (r & 3);
      c.ki[r] = o < lim ? (uint32_t)ci[o] : 0u;
    }
    c.n = pgld_count_f64(a.K[0]);
    c.pB = pgld_count_f64(kj);
    c.qB = c.n - c.pB;
    c.uB = c.qB - c.pB;
    c.PB = c.pB * c.qB;
  }
  __device__ __forceinline__ static HapPop pop(const Counts& c, int r) {
    HapPop p;
    p.c1 = (double)c.n11[r];
    p.pA = pgld_count_f64(c.ki[r]);
    p.n = c.n; p.pB = c.pB; p.qB = c.qB; p.uB = c.uB; p.PB = c.PB;
    p.qA = p.n - p.pA;
    p.c2 = p.pA - p.c1; p.c3 = p.pB - p.c1; p.c4 = p.qA - p.c3;
    p.x = p.c1 * p.c4; p.y = p.c2 * p.c3;
    p.D = p.y - p.x;
    p.s14 = p.c1 + p.c4; p.s23 = p.c2 + p.c3;
    p.uA = p.qA - p.pA;
    p.PA = p.pA * p.qA;
    p.w = fma(p.uA, p.uB, p.s23 - p.s14);
    p.r1 = p.r2 = p.r3 = p.r4 = 0.0;   // not used: the sums are scaled by stat_mult
    return p;
  }
  template <int G>
  __device__ __forceinline__ static void accum(const FusedParams&, const double*, const Counts& c,
                                               int r, double* acc, double wt) {
    const HapPop p = pop(c, r);
    acc[0] = fma(hap_dd_same_num(p), wt, acc[0]);
    acc[1] = fma(hap_dz_same_num(p), wt, acc[1]);
    acc[2] = fma(hap_pi2_same_num(p), wt, acc[2]);
    if constexpr (R2) {
      double x[2];
      hap_r2(p, x);
      acc[3] = fma(x[0], wt, acc[3]);
      acc[4] = fma(x[1], wt, acc[4]);
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
  // one population keeps the centred moments in its slot words; several keep
  // the class counts (their cross-population estimators use them)
  static constexpr bool CENTRED = P == 1;
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
  template <int G>
  __device__ __forceinline__ static int scale_exp(int) { return 0; }
  template <int G>
  __device__ __forceinline__ static double stat_mult(const FusedParams&, const double*, int) {
    return 1.0;
  }
  static constexpr bool PC_OK = true;
  static constexpr bool FULL_FAST = false;
  template <int G>
  __device__ static constexpr int gsz() {
    if constexpr (G < NG) return geno_group_size<P, G>() + (MERGE_R2 ? 2 : 0);
    else return 2;
  }
  template <int G>
  __device__ __forceinline__ static int col(int q) {
    if constexpr (G < NG) return geno_stat_col<P, G>(q);
    else return NS + q;
  }
  struct Counts { v16u w[P][W]; };
  // mailbox form (see HapTraits): word k = W p + q
  static constexpr int NWORDS = P * W;
  __device__ __forceinline__ static void put(const Counts& c, uint32_t* mb, int lane) {
#pragma unroll
    for (int p = 0; p < P; ++p)
#pragma unroll
      for (int q = 0; q < W; ++q)
#pragma unroll
        for (int r = 0; r < 16; ++r) mb[((W * p + q) * 16 + r) * 64 + lane] = c.w[p][q][r];
  }
  __device__ __forceinline__ static void get(Counts& c, const uint32_t* mb, int lane) {
#pragma unroll
    for (int p = 0; p < P; ++p)
#pragma unroll
      for (int q = 0; q < W; ++q)
#pragma unroll
        for (int r = 0; r < 16; ++r) c.w[p][q][r] = mb[((W * p + q) * 16 + r) * 64 + lane];
  }
  __device__ __forceinline__ static void counts(const FusedParams& a, int64_t rb, int64_t cb, int lane,
                                                Counts& c) {
#pragma unroll
    for (int p = 0; p < P; ++p) geno_tile_counts<W, P == 1 ? 2 : 1, CENTRED>(a, p, rb, cb, lane, c.w[p]);
  }
  // counting waves of the producer/consumer kernel (see HapTraits)
#ifndef PGLD_PC_PF_GENO
#define PGLD_PC_PF_GENO 3
#endif
  static constexpr int PC_PF = PGLD_PC_PF_GENO;
  __device__ __forceinline__ static void counts_pc(const FusedParams& a, int64_t rb, int64_t cb,
                                                   int lane, Counts& c) {
#pragma unroll
    for (int p = 0; p < P; ++p) geno_tile_counts<W, PC_PF, CENTRED>(a, p, rb, cb, lane, c.w[p]);
  }
  template <int G>
  __device__ __forceinline__ static void stats(const FusedParams& a, const double* srt, const Counts& c, int r,
                               double* x) {
    GenoPop pp[P];
    if constexpr (CENTRED) {
      // the slot words hold the moments themselves (the one-population
      // estimators and r^2 use nothing else)
      const GenoMoments m = geno_unpack_moments<W>(c.w[0], r);
      const int y[9] = {m.n0, m.b1, m.b2, m.a1, m.ab11, m.ab12, m.a2, m.ab21, m.ab22};
      const int g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      pp[0] = geno_pop(g, y, 0, rt_entry(a, srt, 0, m.n0));
    }
#pragma unroll
    for (int p = 0; p < (CENTRED ? 0 : P); ++p) {
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
    constexpr int GSZ = gsz<G>();
    double x[GSZ];
    stats<G>(a, srt, c, r, x);
#pragma unroll
    for (int q = 0; q < GSZ; ++q) acc[q] = fma(x[q], wt, acc[q]);
  }
};

// One genotype population in the per-wave kernel: the tile's nine moment
// products stay in their FP32 MFMA accumulators (exact integers) and the
// epilogue reads them directly, with no packing into count words.  The nine
// accumulators take 144 registers, which two waves per SIMD still hold.
template <bool R2>
struct GenoRawTraits {
  using Packed = GenoTraits<1, 3, R2>;
  static constexpr int NP = 1;
  static constexpr bool WITH_R2 = R2;
  static constexpr int MINB = 2;
  static constexpr int NS = Packed::NS;
  static constexpr int NG = Packed::NG;
  static constexpr bool MERGE_R2 = Packed::MERGE_R2;
  static constexpr int GS = Packed::GS;
  static constexpr int NSO = Packed::NSO;
  static constexpr int NGT = Packed::NGT;
  static constexpr bool PC_OK = false;
  // the single-bin slot loop fully unrolled: the accumulators are read with
  // static indices (0.722 -> 0.698 ms on 1,000 diploids, 100k sites)
  static constexpr bool FULL_FAST = true;
  template <int G>
  __device__ __forceinline__ static int scale_exp(int q) { return Packed::template scale_exp<G>(q); }
  template <int G>
  __device__ __forceinline__ static double stat_mult(const FusedParams& a, const double* srt, int q) {
    return Packed::template stat_mult<G>(a, srt, q);
  }
  template <int G>
  __device__ static constexpr int gsz() { return Packed::template gsz<G>(); }
  template <int G>
  __device__ __forceinline__ static int col(int q) { return Packed::template col<G>(q); }
  // m[3 f + g] = feature f of the row site . feature g of the column site
  // (u, q, v): ab11, ab12, a1, ab21, ab22, a2, b1, b2, n0
  struct Counts { v16f m[9]; };
  __device__ __forceinline__ static void counts(const FusedParams& a, int64_t rb, int64_t cb, int lane,
                                                Counts& c) {
    const int ks = a.ksteps[0];
    const v4i* A = reinterpret_cast<const v4i*>(a.panels + rb * a.block_bytes + a.pop_off[0]) + lane;
    const v4i* B = reinterpret_cast<const v4i*>(a.panels + cb * a.block_bytes + a.pop_off[0]) + lane;
    const int fs = ks * 64;  // feature stride in v4i
    if (a.missing[0]) {
      mfma_products9<2>(A, B, fs, ks, c.m);
      return;
    }
    mfma_products4<PGLD_PF_MISSING>(A, A + fs, B, B + fs, ks, c.m[0], c.m[1], c.m[3], c.m[4]);
    // no missing calls: the marginal sums from the site tallies (a padded
    // site counts as all heterozygous, u = q = 0)
    const int32_t* st = a.site_stats;
    const int K = a.K[0];
    const int64_t j = cb * 32 + (lane & 31);
    const int n1j = j < a.V ? st[j] : K, n2j = j < a.V ? st[a.V + j] : 0;
    const float b1 = (float)(2 * n2j + n1j - K), b2 = (float)(K - n1j);
    // opaque per tile: a loop-invariant 16-register splat would be hoisted
    // out of the tile loop and held (or spilled) across it
    float n0 = (float)K;
    asm volatile("" : "+v"(n0));
    const int64_t ib = rb * 32 + 4 * (lane >> 5);
    const int32_t* si = st + ib;
    const int32_t* sci = st + a.V + ib;
    const int64_t lim = a.V - ib;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int o = 8 * (r >> 2) + (r & 3);
      const int n1i = o < lim ? si[o] : K, n2i = o < lim ? sci[o] : 0;
      c.m[2][r] = (float)(2 * n2i + n1i - K);
      c.m[5][r] = (float)(K - n1i);
      c.m[6][r] = b1;
      c.m[7][r] = b2;
      c.m[8][r] = n0;
    }
  }
  template <int G>
  __device__ __forceinline__ static void stats(const FusedParams& a, const double* srt, const Counts& c, int r,
                                               double* x) {
    GenoPop p;
#pragma unroll
    for (int k = 0; k < 9; ++k) p.g[k] = 0.0;  // the one-population forms read only moments
    p.ab11 = (double)c.m[0][r]; p.ab12 = (double)c.m[1][r]; p.a1 = (double)c.m[2][r];
    p.ab21 = (double)c.m[3][r]; p.ab22 = (double)c.m[4][r]; p.a2 = (double)c.m[5][r];
    p.b1 = (double)c.m[6][r]; p.b2 = (double)c.m[7][r];
    const float n0f = c.m[8][r];
    p.n0 = (double)n0f;
    p.n = p.n0;
    const double2* t = reinterpret_cast<const double2*>(rt_entry(a, srt, 0, (int)n0f));
    const double2 ra = t[0], rb = t[1];
    p.r1 = ra.x; p.r2 = ra.y; p.r3 = rb.x; p.r4 = rb.y;
    if constexpr (G < NG) {
      geno_stats_group<1, G>(&p, x);
      if constexpr (MERGE_R2) geno_r2(p, x + GS - 2);
    } else {
      geno_r2(p, x);
    }
  }
  template <int G>
  __device__ __forceinline__ static void accum(const FusedParams& a, const double* srt,
                                               const Counts& c, int r, double* acc, double wt) {
    constexpr int GSZ = gsz<G>();
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
__device__ __forceinline__ void stage_drain(const FusedParams& a, const double* srt, double* stg,
                                            int* scnt, unsigned long long* slimb, int lane,
                                            int blo, int bhi) {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  const int E = (bhi - blo + 1) * (GS + 1);
  for (int e = lane; e < E; e += 64) {
    const int b = blo + e / (GS + 1), q = e - (b - blo) * (GS + 1);
    if (q < GS) {
      const double v = stg[b * GS + q];
      if (v != 0.0) {
        stg[b * GS + q] = 0.0;
        unsigned long long* dst = slimb + ((int64_t)b * (NS + 1) + T::template col<G>(q)) * 3;
        long long x2, x1, x0;
        pgld_to_limbs(v * T::template stat_mult<G>(a, srt, q), x2, x1, x0,
                      pgld_limb_scale(T::template scale_exp<G>(q)));
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
  const FusedParams* fp;
  const double* sedges;
  const double* srt;
  double* stg;   // this wave's staging [nb][GS]
  int* scw;      // this wave's pair counters [nb]
  unsigned long long* slimb;
  int lane, bmin, bmax;
  bool fast;   // every pair valid and in bin bmin
  bool clean;  // every pair valid (i < j, inside the cutoff and the bins)
  bool dual;   // otherwise: every valid pair in bin blo or blo + 1
  int blo;
  int64_t i0, j0;
};

// Wave totals of a group's per-lane sums for one bin, added to the workgroup
// limbs (lane q < GS: the group's statistic q, column T::col<G>(q)); with
// COUNT, `pairs` joins the bin's pair counter.
template <class T, int G, int NS, int GS, bool COUNT>
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
    unsigned long long* dst = t.slimb + ((int64_t)bin * (NS + 1) + T::template col<G>(q)) * 3;
    long long x2, x1, x0;
    pgld_to_limbs(tot * T::template stat_mult<G>(*t.fp, t.srt, q), x2, x1, x0,
                  pgld_limb_scale(T::template scale_exp<G>(q)));
    atomicAdd(dst + 0, (unsigned long long)x0);
    atomicAdd(dst + 1, (unsigned long long)x1);
    atomicAdd(dst + 2, (unsigned long long)x2);
  } else if (COUNT && lane < NV && q == GS) {
    atomicAdd(t.slimb + ((int64_t)bin * (NS + 1) + NS) * 3, pairs);
  }
}

// Bins of the lane's 16 pairs of a tile that is not single-bin: 8 bits per
// slot r in word r >> 2, 0xFF where the pair is not counted (i >= j, beyond
// the cutoff, outside the bins).  The corner pairs bound every pair's bin to
// [t.bmin, t.bmax], so a pair's bin is the window's first bin plus the number
// of the window's inner edges at or below its distance (the digitize rule for
// non-decreasing edges): wave-uniform edges, no per-lane search.  The 16 row
// coordinates are loaded at once (indices clamped into the sites).
#ifndef PGLD_BIN_WINDOW
#define PGLD_BIN_WINDOW 24  // widest window of inner edges counted; wider: per-lane scan
#endif
// tile_bins for a clean tile: off the diagonal, all 32 x 32 sites real,
// inside the cutoff and inside the bins (bins [lo, hi] wave-uniform), so every
// pair is counted.  The row coordinates are the same for all lanes of a half
// wave: scalar loads, one select per pair.
__device__ __forceinline__ void tile_bins_clean(const FusedParams& a, const TileCtx& t, int lo,
                                                int hi, uint32_t (&bw)[4]) {
  const int lane = t.lane;
  const double* c = a.coord_is_pos ? a.pos : a.coord;
  const double cj = c[t.j0 + (lane & 31)];
  const double* crow = c + t.i0;
  const bool upper = lane >= 32;  // rows o + 4
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    double d[8];
    int cnt[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const int r = 8 * h + s, o = 8 * (r >> 2) + (r & 3);
      const double ci = upper ? crow[o + 4] : crow[o];
      d[s] = cj - ci;  // coordinates are non-decreasing: the distance for i < j
      cnt[s] = lo;
    }
    for (int e = lo + 1; e <= hi; ++e) {
      const double ee = a.edges[e];
#pragma unroll
      for (int s = 0; s < 8; ++s) cnt[s] += d[s] >= ee ? 1 : 0;
    }
#pragma unroll
    for (int w2 = 0; w2 < 2; ++w2)
      bw[2 * h + w2] = (uint32_t)cnt[4 * w2] | ((uint32_t)cnt[4 * w2 + 1] << 8) |
                       ((uint32_t)cnt[4 * w2 + 2] << 16) | ((uint32_t)cnt[4 * w2 + 3] << 24);
  }
}

// CLEAN: clean tiles take tile_bins_clean.  Where its registers fit: one or
// two populations at two or four waves per SIMD (the three-wave and
// one-wave kernels spill with it).
template <class T>
__host__ __device__ constexpr bool clean_bins() {
  return T::NP <= 2 && (T::MINB == 2 || T::MINB == 4);
}
template <bool CLEAN>
__device__ __forceinline__ void tile_bins(const FusedParams& a, const TileCtx& t,
                                          uint32_t (&bw)[4]) {
  const int nb = a.n_bins, lane = t.lane;
  const int64_t j = t.j0 + (lane & 31);
  const bool jok = j < a.V;
  const int64_t jc = jok ? j : a.V - 1;
  const double pos_j = a.pos[jc];
  const double co_j = a.coord_is_pos ? pos_j : a.coord[jc];
  // row i = ib + o with o = 8q + k a constant (mfma_row(lane, 4q + k))
  const int64_t ib = t.i0 + 4 * (lane >> 5);
  const int jr = jok ? (int)(j - ib < 32 ? j - ib : 32) : 0;  // o < jr <=> i < j (o < 32)
  // candidate bins [lo, hi] of the tile (wave-uniform); hi < lo: none
  const int lo = __builtin_amdgcn_readfirstlane(t.bmin < 0 ? 0 : t.bmin);
  const int hi = __builtin_amdgcn_readfirstlane(t.bmax >= nb ? nb - 1 : t.bmax);
  const bool window = hi - lo <= PGLD_BIN_WINDOW;  // inner edges lo + 1 .. hi
  if constexpr (CLEAN) {
    if (t.clean && window) {
      tile_bins_clean(a, t, lo, hi, bw);
      return;
    }
  }
  int guess = nb - 1;
  // two halves of 8 slots (words 2h, 2h + 1) keep the live distances few
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    double d[8];
    uint32_t in_cut = 0;  // bit s: pos_j <= pos_i + max_dist
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const int r = 8 * h + s;
      const int64_t i = ib + 8 * (r >> 2) + (r & 3);
      const int64_t ic = i < a.V ? i : a.V - 1;
      const double pi = a.pos[ic];
      d[s] = a.coord_is_pos ? pos_j - pi : fabs(co_j - a.coord[ic]);
      in_cut |= (pos_j <= pi + a.max_dist ? 1u : 0u) << s;
    }
    if (window) {
      int c[8];
#pragma unroll
      for (int s = 0; s < 8; ++s) c[s] = lo;
      for (int e = lo + 1; e <= hi; ++e) {
        const double ee = a.edges[e];
#pragma unroll
        for (int s = 0; s < 8; ++s) c[s] += d[s] >= ee ? 1 : 0;
      }
      // lo <= n_bins and hi + 1 >= 0 index real edges (n_bins + 1 of them)
      const double e_lo = a.edges[lo], e_hi = a.edges[hi + 1];
#pragma unroll
      for (int w2 = 0; w2 < 2; ++w2) {
        uint32_t w = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int s = 4 * w2 + k, o = 8 * (2 * h + w2) + k;
          const bool ok = o < jr && ((in_cut >> s) & 1u) && d[s] >= e_lo && d[s] < e_hi;
          w |= (ok ? (uint32_t)c[s] : 0xFFu) << (8 * k);
        }
        bw[2 * h + w2] = w;
      }
    } else {
#pragma unroll
      for (int w2 = 0; w2 < 2; ++w2) {
        uint32_t w = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int s = 4 * w2 + k, o = 8 * (2 * h + w2) + k;
          uint32_t b8 = 0xFF;
          if (o < jr) {
            const int b = pgld_bin_scan(d[s], t.sedges, nb, guess, nb - 1);
            if (((in_cut >> s) & 1u) && b >= 0 && b < nb) {
              b8 = (uint32_t)b;
              guess = b;
            }
          }
          w |= b8 << (8 * k);
        }
        bw[2 * h + w2] = w;
      }
    }
  }
}

// One pass over the tile's 1024 pairs for statistic group G.  Single-bin
// tiles are reduced to limbs at the tile end, so FP64 rounding never spans
// more than one tile (and the result does not depend on the schedule).
// PRE: the tile's bin words (tile_bins) were computed beforehand and are
// passed in `pre` (the producer/consumer kernel: its counting wave does it).
template <class T, int G, bool PRE = false>
__device__ __forceinline__ void tile_group(const FusedParams& a, const TileCtx& t,
                                           const typename T::Counts& cts,
                                           const uint32_t (&pre)[4]) {
  // NS below is the output row width; GS the size of this group
  constexpr int NS = T::NSO;
  constexpr int GS = T::template gsz<G>();
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
    if constexpr (T::FULL_FAST) {  // static slot indices (no index-mode moves)
#pragma unroll
      for (int r = 0; r < 16; ++r) T::template accum<G>(a, t.srt, cts, r, acc, 1.0);
    } else if constexpr (U2) {  // 2 x 500 haplotypes, 200k sites: 3.21 -> 3.18 ms
#pragma unroll 2
      for (int r = 0; r < 16; ++r) T::template accum<G>(a, t.srt, cts, r, acc, 1.0);
    } else {
#pragma unroll 1
      for (int r = 0; r < 16; ++r) T::template accum<G>(a, t.srt, cts, r, acc, 1.0);
    }
    reduce_to_limbs<T, G, NS, GS, G == 0>(t, acc, t.bmin, 1024ull);
  } else if (t.dual) {
    // Two bins: each pair's statistics go to both accumulators with 0/1
    // weights (the estimators are finite for every slot: counts of padded
    // sites are zero and the reciprocal table covers every n), so the wave
    // does the fast path's work plus one FMA per statistic, and the per-bin
    // pair counts are ballots.
    uint32_t bw[4];
    if constexpr (PRE) {
#pragma unroll
      for (int q = 0; q < 4; ++q) bw[q] = pre[q];
    } else {
      tile_bins<clean_bins<T>()>(a, t, bw);
    }
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
    reduce_to_limbs<T, G, NS, GS, G == 0>(t, acc, t.blo, cnt0);
    if (t.blo + 1 < nb) reduce_to_limbs<T, G, NS, GS, G == 0>(t, acc1, t.blo + 1, cnt1);
  } else {
    uint32_t bw[4];
    if constexpr (PRE) {
#pragma unroll
      for (int q = 0; q < 4; ++q) bw[q] = pre[q];
    } else {
      tile_bins<clean_bins<T>()>(a, t, bw);
    }
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
    stage_drain<T, G, NS, GS>(a, t.srt, t.stg, t.scw, t.slimb, lane, t.bmin < 0 ? 0 : t.bmin,
                              t.bmax >= nb ? nb - 1 : t.bmax);
  }
}

template <class T, int G, bool PRE = false>
__device__ __forceinline__ void tile_groups(const FusedParams& a, const TileCtx& t,
                                            const typename T::Counts& cts,
                                            const uint32_t (&pre)[4]) {
  if constexpr (G < T::NGT) {
    tile_group<T, G, PRE>(a, t, cts, pre);
    tile_groups<T, G + 1, PRE>(a, t, cts, pre);
  }
}


}  // namespace pgld
