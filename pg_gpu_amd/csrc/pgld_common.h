// Shared definitions for the pgld HIP library (gfx950 / MI355X).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define PGLD_MAX_POPS 4
#define PGLD_MAX_BINS 64
#define PGLD_SITE_BLOCK 32      // sites per packed site block (MFMA 32x32 M/N)
#define PGLD_KSTEP 32           // samples per packed k-step (MFMA 32x32x32 K)
#define PGLD_FRAG_BYTES 1024    // one 32 sites x 32 samples int8 fragment

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

// Fixed-point bin accumulator: value = (L2 * 2^64 + L1 * 2^32 + L0) * 2^-72.
// A double is split as sign * (h + m * 2^-32 + s * 2^-64) * 2^-8 with h, m, s
// integers (|v| * 2^8 = h + frac, frac * 2^32 = m + frac', frac' * 2^32 ~ s);
// every step is exact for |v| >= 0, only the last digit is rounded (to
// nearest), so one conversion errs by at most 2^-73.  The three signed limbs
// are added with plain integer adds; integer addition is associative, so the
// sums do not depend on the order in which workgroups, waves or lanes add
// their contributions.

// Exact conversion of a count (< 2^31) to FP64 through the bit pattern of
// 2^52 + v.  The result is opaque to the compiler's integer reasoning, so
// FP64 products and sums of counts stay single FP64 instructions instead of
// being rewritten as integer multiplies followed by conversions.
__device__ __forceinline__ double pgld_count_f64(uint32_t v) {
  return __hiloint2double(0x43300000, (int)v) - 4503599627370496.0;
}

#define PGLD_LIMBS 3

__device__ __forceinline__ void pgld_to_limbs(double v, long long& x2, long long& x1,
                                              long long& x0) {
  const double a = fabs(v);
  const double t = a * 256.0;                 // exact power-of-two scaling
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

// Wave-wide (wave64) sum of a double; every lane receives the total.
__device__ __forceinline__ double pgld_wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
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
