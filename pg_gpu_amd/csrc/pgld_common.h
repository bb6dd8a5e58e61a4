// Shared definitions for the pgld HIP library (gfx950 / MI355X).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define PGLD_MAX_POPS 4
#define PGLD_MAX_BINS 64
#define PGLD_SITE_BLOCK 32      // sites per packed site block (MFMA 32x32 M/N)
#define PGLD_KSTEP 64           // samples per packed k-step (MFMA 32x32x64 K)
#define PGLD_FRAG_BYTES 1024    // one 32 sites x 64 samples fragment of 4-bit values
#define PGLD_FP4_ONE 2u         // 1.0 in FP4 E2M1 (sign 0, exponent 01, mantissa 0)
#define PGLD_FP4_NEG_ONE 0xAu   // -1.0 in FP4 E2M1 (sign bit set)

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v8i __attribute__((ext_vector_type(8)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef float v16f __attribute__((ext_vector_type(16)));

// Counting product of two 0/1 indicator fragments: 4-bit E2M1 operands
// (1.0 or 0), FP32 accumulation.  Every partial sum is an integer below 2^24,
// so the accumulated counts are exact.  The E8M0 block scales are 127 (2^0).
__device__ __forceinline__ v16f pgld_mfma_fp4(v4i a, v4i b, v16f c) {
  const v8i a8 = {a[0], a[1], a[2], a[3], 0, 0, 0, 0};
  const v8i b8 = {b[0], b[1], b[2], b[3], 0, 0, 0, 0};
  return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a8, b8, c, 4, 4, 0, 127, 0, 127);
}

__device__ __forceinline__ v16f zero16f() {
  v16f z;
#pragma unroll
  for (int q = 0; q < 16; ++q) z[q] = 0.0f;
  return z;
}

// Fixed-point bin accumulator: value = (L2 * 2^64 + L1 * 2^32 + L0) * 2^-72.
// A double is split as sign * (h + m * 2^-32 + s * 2^-64) * 2^-8 with h, m, s
// integers (|v| * 2^8 = h + frac, frac * 2^32 = m + frac', frac' * 2^32 ~ s);
// every step is exact for |v| >= 0, only the last digit is rounded (to
// nearest), so one conversion errs by at most 2^-73.  The three signed limbs
// are added with plain integer adds; integer addition is associative, so the
// sums do not depend on the order in which workgroups, waves or lanes add
// their contributions.

// Exact conversion of a count to FP64 (one v_cvt_f64_u32).  Written as an
// asm statement so that the result is opaque to the compiler's integer
// reasoning: FP64 products and sums of counts then stay single FP64
// instructions instead of being rewritten as integer multiplies followed by
// conversions.
__device__ __forceinline__ double pgld_count_f64(uint32_t v) {
  double d;
  asm("v_cvt_f64_u32 %0, %1" : "=v"(d) : "v"(v));
  return d;
}

// Signed twin of pgld_count_f64 (one v_cvt_f64_i32).
__device__ __forceinline__ double pgld_int_f64(int32_t v) {
  double d;
  asm("v_cvt_f64_i32 %0, %1" : "=v"(d) : "v"(v));
  return d;
}

#define PGLD_LIMBS 3

// 2^(8-e) as a double (0 <= e < 16): the limb scale of a sum held as
// v = 2^e * value
__device__ __forceinline__ double pgld_limb_scale(int e) {
  return __builtin_bit_cast(double, (unsigned long long)(1023 + 8 - e) << 52);
}

__device__ __forceinline__ void pgld_to_limbs(double v, long long& x2, long long& x1,
                                              long long& x0, double scale = 256.0) {
  const double a = fabs(v);
  const double t = a * scale;                 // exact power-of-two scaling
  const double h = floor(t);
  const double r = (t - h) * 4294967296.0;    // fraction in [0, 2^32), exact
  const double m = floor(r);
  const double s = (r - m) * 4294967296.0;    // in [0, 2^32), exact
  long long y2 = (long long)h, y1 = (long long)m, y0 = (long long)rint(s);
  if (v < 0.0) { y2 = -y2; y1 = -y1; y0 = -y0; }
  x2 = y2; x1 = y1; x0 = y0;
}

// Carry-normalise three signed limb sums so that L0, L1 lie in [0, 2^32).
__device__ __forceinline__ void pgld_normalize_limbs(long long& l2, long long& l1, long long& l0) {
  long long c = l0 >> 32;  // arithmetic shift: floor division by 2^32
  l0 -= c * 4294967296ll;
  l1 += c;
  c = l1 >> 32;
  l1 -= c * 4294967296ll;
  l2 += c;
}

// DPP move of both halves of a double (row-local lane permutation `ctrl`).
template <int CTRL>
__device__ __forceinline__ double pgld_dpp_f64(double v) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(uint32_t)u, CTRL, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(uint32_t)(u >> 32), CTRL, 0xf, 0xf, true);
  return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// Wave-wide (wave64) sum of a double, valid in lanes 0..31.  Within each
// 16-lane row: DPP quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_ror 4 and
// row_ror 8 leave the row total in every lane of the row; a ds_swizzle
// exchange of the two rows of each 32-lane half and a readlane of lane 32
// complete the sum.  The summation order is fixed, so the result is
// deterministic.
__device__ __forceinline__ double pgld_wave_sum(double v) {
  v += pgld_dpp_f64<0xB1>(v);   // quad_perm [1,0,3,2]: lane ^ 1
  v += pgld_dpp_f64<0x4E>(v);   // quad_perm [2,3,0,1]: lane ^ 2
  v += pgld_dpp_f64<0x124>(v);  // row_ror 4
  v += pgld_dpp_f64<0x128>(v);  // row_ror 8
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  // ds_swizzle bitmask mode: and 0x1f, or 0, xor 0x10 -> lane ^ 16 within 32
  const int lo = __builtin_amdgcn_ds_swizzle((int)(uint32_t)u, 0x401F);
  const int hi = __builtin_amdgcn_ds_swizzle((int)(uint32_t)(u >> 32), 0x401F);
  v += __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
  const uint64_t w = __builtin_bit_cast(uint64_t, v);
  const uint32_t hl = __builtin_amdgcn_readlane((int)(uint32_t)w, 32);
  const uint32_t hh = __builtin_amdgcn_readlane((int)(uint32_t)(w >> 32), 32);
  return v + __builtin_bit_cast(double, ((uint64_t)hh << 32) | hl);
}

// Exchange a double with lane ^ M (M = 1, 2: DPP quad_perm; 4, 8, 16:
// ds_swizzle bitmask xor; 32: ds_bpermute).
template <int M>
__device__ __forceinline__ double pgld_xchg_f64(double v) {
  if constexpr (M == 1) {
    return pgld_dpp_f64<0xB1>(v);
  } else if constexpr (M == 2) {
    return pgld_dpp_f64<0x4E>(v);
  } else if constexpr (M == 4 || M == 8 || M == 16) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const int lo = __builtin_amdgcn_ds_swizzle((int)(uint32_t)u, (M << 10) | 0x1F);
    const int hi = __builtin_amdgcn_ds_swizzle((int)(uint32_t)(u >> 32), (M << 10) | 0x1F);
    return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
  } else {
    static_assert(M == 32, "lane masks 1..32");
    return __shfl_xor(v, 32, 64);
  }
}

// One butterfly step of a wave reduce-scatter: lanes with bit M clear keep
// v[0..CNT) and send v[CNT..2 CNT) to lane ^ M, which keeps the upper half;
// each adds what it receives to what it keeps.
template <int CNT, int M>
__device__ __forceinline__ void pgld_butterfly(double* v, int lane) {
  const bool up = (lane & M) != 0;
#pragma unroll
  for (int k = 0; k < CNT; ++k) {
    const double keep = up ? v[k + CNT] : v[k];
    const double send = up ? v[k] : v[k + CNT];
    v[k] = keep + pgld_xchg_f64<M>(send);
  }
}

// All butterfly steps of masks M, 2M, ..., NV/2.
template <int NV, int M>
__device__ __forceinline__ void pgld_butterflies(double* v, int lane) {
  if constexpr (M < NV) {
    pgld_butterfly<NV / (2 * M), M>(v, lane);
    pgld_butterflies<NV, 2 * M>(v, lane);
  }
}
// Plain exchange-and-add over the lane masks M .. 32.
template <int M>
__device__ __forceinline__ double pgld_xsum(double t) {
  if constexpr (M <= 32) return pgld_xsum<2 * M>(t + pgld_xchg_f64<M>(t));
  else return t;
}

// Wave (wave64) sums of NV (2 .. 64, a power of two) per-lane doubles with
// one butterfly reduce-scatter: lane l ends with the total of value
// bitreverse(l mod NV) (over log2 NV bits) -- every lane of the wave, as the
// remaining lane bits are summed by plain exchanges.  A fixed summation tree,
// so the result is deterministic; about 7 NV instructions instead of 16 NV
// for NV separate wave sums.
template <int NV>
__device__ __forceinline__ double pgld_wave_sums(double (&v)[NV], int lane) {
  static_assert(NV == 2 || NV == 4 || NV == 8 || NV == 16 || NV == 32 || NV == 64, "NV: 2 .. 64");
  pgld_butterflies<NV, 1>(v, lane);
  return pgld_xsum<NV>(v[0]);
}

__device__ __forceinline__ int pgld_bitrev(int x, int bits) {
  return (int)(__builtin_bitreverse32((uint32_t)x) >> (32 - bits));
}

__device__ __forceinline__ int pgld_wave_sum_int(int v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// digitize(d, edges) - 1 restricted to a known window [lo, hi] of bin ids:
// returns b in [lo, hi] with edges[b] <= d < edges[b+1], or -1 / n_bins
// outside [edges[0], edges[n_bins]).  Linear scan: the window is small.
__device__ __forceinline__ int pgld_bin_scan(double d, const double* edges, int n_bins,
                                             int lo, int hi) {
  if (!(d >= edges[0])) return -1;          // also catches NaN
  if (!(d < edges[n_bins])) return n_bins;
  int b = lo < 0 ? 0 : lo;
  while (b > 0 && d < edges[b]) --b;
  while (b < n_bins - 1 && d >= edges[b + 1]) ++b;
  (void)hi;
  return b;
}

// The same digitize for a wave-uniform distance d, from one edge per lane
// (lane l holds edges[l], +inf past the last edge; e64 = edges[64], the one
// edge no lane holds, when n_bins = 64): the number of edges at or below d,
// minus one.  Every lane of the wave must be active.
__device__ __forceinline__ int pgld_bin_ballot(double d, double lane_edge, int n_bins, double e64) {
  int cnt = __popcll(__ballot(d >= lane_edge));
  if (n_bins == 64 && d >= e64) ++cnt;
  return cnt - 1;
}

// Centred dosage moments of a site pair over the samples valid at both
// (genotype panels: products of the features u = g - 1, q = u^2, v = [g >= 0];
// row site first): ab11 = u.u, ab12 = u.q, ab21 = q.u, ab22 = q.q, a1 = u.v,
// a2 = q.v, b1 = v.u, b2 = v.q, n0 = v.v.
struct GenoMoments { int n0, a1, a2, b1, b2, ab11, ab12, ab21, ab22; };

// The nine class counts (moments order) of a pair from its moments; every
// numerator is an exact multiple of its divisor and non-negative.
__device__ __forceinline__ void geno_class_counts(const GenoMoments& m, int (&c)[9]) {
  const int uu = m.ab11, uq = m.ab12, qu = m.ab21, qq = m.ab22;
  c[0] = (qq + qu + uq + uu) >> 2;           // n22
  c[1] = (m.a2 + m.a1 - qq - uq) >> 1;       // n21 = (qv + uv - qq - uq) / 2
  c[2] = (qq - qu + uq - uu) >> 2;           // n20
  c[3] = (m.b2 + m.b1 - qq - qu) >> 1;       // n12 = (vq + vu - qq - qu) / 2
  c[4] = m.n0 - m.b2 - m.a2 + qq;            // n11 = vv - vq - qv + qq
  c[5] = (m.b2 - m.b1 - qq + qu) >> 1;       // n10 = (vq - vu - qq + qu) / 2
  c[6] = (qq + qu - uq - uu) >> 2;           // n02
  c[7] = (m.a2 - m.a1 - qq + uq) >> 1;       // n01 = (qv - uv - qq + uq) / 2
  c[8] = (qq - qu - uq + uu) >> 2;           // n00
}
