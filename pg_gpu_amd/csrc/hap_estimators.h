// Haplotype two-locus LD estimators (unbiased DD, Dz, pi2 raw calls).
//
// Device restatement of the reference's per-call formulas
// (pg_gpu/haplotype_kernels.py:69-304, the same arithmetic as
// pg_gpu/ld_statistics.py:1116-1617).  Each raw call is an integer polynomial
// in the four haplotype counts of the populations it involves, divided by a
// product of falling factorials (n)_k of their valid sample sizes.  The
// polynomials are evaluated in FP64 where every intermediate is an integer
// (or a multiple of 1/4) below 2^53, i.e. exactly; the division is replaced by
// a multiplication with 1/(n)_k read from a table (0 where (n)_k <= 0, which
// reproduces the reference's "den > 0 else 0" guard).
#pragma once
#include "pgld_common.h"

struct HapPop {
  double c1, c2, c3, c4;   // n11, n10, n01, n00
  double D;                // c2*c3 - c1*c4 (reference sign)
  double pA, qA, pB, qB;   // allele counts at locus A (left) and B (right)
  double r1, r2, r3, r4;   // 1/(n)_1 .. 1/(n)_4, 0 when not positive
  // shared products of the estimators below (integers)
  double x, y;             // c1*c4, c2*c3
  double s14, s23;         // c1 + c4, c2 + c3
  double uA, uB;           // qA - pA, qB - pB
  double PA, PB;           // pA*qA, pB*qB
  double w;                // uA*uB + s23 - s14
  double n;                // c1 + c2 + c3 + c4
};

// The shared products of a population's counts (c1..c4 and the r's set).
__device__ __forceinline__ void hap_basis(HapPop& p) {
  p.x = p.c1 * p.c4; p.y = p.c2 * p.c3;
  p.D = p.y - p.x;
  p.pA = p.c1 + p.c2; p.qA = p.c3 + p.c4;
  p.pB = p.c1 + p.c3; p.qB = p.c2 + p.c4;
  p.s14 = p.c1 + p.c4; p.s23 = p.c2 + p.c3;
  p.uA = p.qA - p.pA; p.uB = p.qB - p.pB;
  p.PA = p.pA * p.qA; p.PB = p.pB * p.qB;
  p.w = fma(p.uA, p.uB, p.s23 - p.s14);
  p.n = p.pA + p.qA;
}

// The same population from the marginals the indicator products give
// directly: c1 = n11 (H.H), pA = n11 + n10 (H.V), pB = n11 + n01 (V.H) and
// n = n11 + n10 + n01 + n00 (V.V); rt is the table entry of n.  Every field is
// the same integer as hap_pop's (exact in FP64 whatever the grouping).
__device__ __forceinline__ HapPop hap_pop_marg(int hh, int hv, int vh, int vv,
                                               const double* __restrict__ rt) {
  HapPop p;
  p.c1 = pgld_count_f64(hh); p.pA = pgld_count_f64(hv);
  p.pB = pgld_count_f64(vh); p.n = pgld_count_f64(vv);
  p.qA = p.n - p.pA; p.qB = p.n - p.pB;
  p.c2 = p.pA - p.c1; p.c3 = p.pB - p.c1; p.c4 = p.qA - p.c3;
  p.x = p.c1 * p.c4; p.y = p.c2 * p.c3;
  p.D = p.y - p.x;
  p.s14 = p.c1 + p.c4; p.s23 = p.c2 + p.c3;
  p.uA = p.qA - p.pA; p.uB = p.qB - p.pB;
  p.PA = p.pA * p.qA; p.PB = p.pB * p.qB;
  p.w = fma(p.uA, p.uB, p.s23 - p.s14);
  const double2* t = reinterpret_cast<const double2*>(rt);
  const double2 a = t[0], b = t[1];
  p.r1 = a.x; p.r2 = a.y; p.r3 = b.x; p.r4 = b.y;
  return p;
}

// rtab holds 4 doubles per n: 1/(n)_1, 1/(n)_2, 1/(n)_3, 1/(n)_4.
__device__ __forceinline__ HapPop hap_pop(int n11, int n10, int n01, int n00, int n,
                                          const double* __restrict__ rtab) {
  HapPop p;
  p.c1 = pgld_count_f64(n11); p.c2 = pgld_count_f64(n10);
  p.c3 = pgld_count_f64(n01); p.c4 = pgld_count_f64(n00);
  hap_basis(p);
  const double2* t = reinterpret_cast<const double2*>(rtab + 4 * n);
  const double2 a = t[0], b = t[1];
  p.r1 = a.x; p.r2 = a.y; p.r3 = b.x; p.r4 = b.y;
  return p;
}

// Each raw call is num * R: num the reference's polynomial in the counts
// (times an exact power of two), regrouped over the shared products above
// (identities in the counts, e.g.
// c1(c1-1)c4(c4-1) + c2(c2-1)c3(c3-1) - 2c1c2c3c4 = D^2 - x(s14-1) - y(s23-1)),
// and R the product of the reciprocal falling factorials 1/(n)_k of its
// populations.  Every intermediate of num is an integer (or a multiple of
// 1/16) below 2^53, so num equals the reference's numerator exactly whatever
// the grouping; hap_X = hap_X_num * R.  The fused kernel folds R into its
// accumulation (one FMA per statistic, see the generated hap_acc_* groups).
// Where num carries a constant power of two, hap_X_raw is num without it
// (num = raw * 2^-k): the fused kernel accumulates raw numerators and applies
// 2^-k once, where a tile's sums become fixed-point limbs.  Scaling by a power
// of two commutes with every rounding (no value comes near the subnormal
// range), so the limbs are the ones the scaled accumulation would give.

// x (s14 - 1) + y (s23 - 1), shared by the single-population DD and pi2
__device__ __forceinline__ double hap_x1(const HapPop& a) {
  return fma(a.x, a.s14 - 1.0, a.y * (a.s23 - 1.0));
}

// ---- DD -------------------------------------------------------------------
__device__ __forceinline__ double hap_dd_same_num(const HapPop& a) {
  return fma(a.D, a.D, -hap_x1(a));
}
__device__ __forceinline__ double hap_dd_same(const HapPop& a) { return hap_dd_same_num(a) * a.r4; }
__device__ __forceinline__ double hap_dd_pair_num(const HapPop& a, const HapPop& b) {
  return a.D * b.D;
}
__device__ __forceinline__ double hap_dd_pair(const HapPop& a, const HapPop& b) {
  return hap_dd_pair_num(a, b) * (a.r2 * b.r2);
}

// ---- Dz -------------------------------------------------------------------
// (c1c4 - c2c3)((qA-pA)(qB-pB) + (c2+c3) - (c1+c4)) + 2(c2c3 + c1c4)
__device__ __forceinline__ double hap_dz_same_num(const HapPop& a) {
  return fma(-a.D, a.w, 2.0 * (a.x + a.y));
}
__device__ __forceinline__ double hap_dz_same(const HapPop& a) { return hap_dz_same_num(a) * a.r4; }
// Dz(i,i,k)
__device__ __forceinline__ double hap_dz_iik_num(const HapPop& a, const HapPop& b) {
  return (a.uA * (-a.D)) * b.uB;
}
__device__ __forceinline__ double hap_dz_iik(const HapPop& a, const HapPop& b) {
  return hap_dz_iik_num(a, b) * (a.r3 * b.r1);
}
// (Dz(i,i,k) + Dz(i,k,i)) / 2, the averaged statistic: both share -D_i and
// the reciprocals, so their numerators merge
__device__ __forceinline__ double hap_dz_iik_iji_avg_raw(const HapPop& a, const HapPop& b) {
  return (-a.D) * fma(a.uA, b.uB, a.uB * b.uA);
}
__device__ __forceinline__ double hap_dz_iik_iji_avg_num(const HapPop& a, const HapPop& b) {
  return hap_dz_iik_iji_avg_raw(a, b) * 0.5;
}
// Dz(i,j,i)
__device__ __forceinline__ double hap_dz_iji_num(const HapPop& a, const HapPop& b) {
  return (a.uB * (-a.D)) * b.uA;
}
__device__ __forceinline__ double hap_dz_iji(const HapPop& a, const HapPop& b) {
  return hap_dz_iji_num(a, b) * (a.r3 * b.r1);
}
// Dz(i,j,j)
__device__ __forceinline__ double hap_dz_ijj_num(const HapPop& a, const HapPop& b) {
  return (-a.D) * b.w;
}
__device__ __forceinline__ double hap_dz_ijj(const HapPop& a, const HapPop& b) {
  return hap_dz_ijj_num(a, b) * (a.r2 * b.r2);
}
// Dz(i,j,k), all distinct
__device__ __forceinline__ double hap_dz_ijk_num(const HapPop& a, const HapPop& b,
                                                 const HapPop& c) {
  return -(a.D * b.uA * c.uB);
}
__device__ __forceinline__ double hap_dz_ijk(const HapPop& a, const HapPop& b, const HapPop& c) {
  return hap_dz_ijk_num(a, b, c) * (a.r2 * b.r1 * c.r1);
}

// ---- pi2 ------------------------------------------------------------------
// pA pB qA qB - c1c4 (c1 + 3c2 + 3c3 + c4 - 1) - c2c3 (3c1 + c2 + c3 + 3c4 - 1)
// = PA PB - [x (s14 - 1) + y (s23 - 1)] - 3 (x s23 + y s14)
__device__ __forceinline__ double hap_pi2_same_num(const HapPop& a) {
  const double z = fma(a.x, a.s23, a.y * a.s14);
  return fma(-3.0, z, fma(a.PA, a.PB, -hap_x1(a)));
}
__device__ __forceinline__ double hap_pi2_same(const HapPop& a) { return hap_pi2_same_num(a) * a.r4; }
// pop t three times, pop s once (all "triple" cases share this formula):
// (s.pB (PA qB - m) + s.qB (PA pB + m - 2 PA)) / 2 with PA, qB, pB, m of t and
// m = pA c4 + c2 qA
__device__ __forceinline__ double hap_pi2_triple_raw(const HapPop& t, const HapPop& s) {
  const double m = fma(t.pA, t.c4, t.c2 * t.qA);
  const double u = fma(t.PA, t.qB, -m);
  const double v = fma(-2.0, t.PA, fma(t.PA, t.pB, m));
  return fma(s.qB, v, s.pB * u);
}
__device__ __forceinline__ double hap_pi2_triple_num(const HapPop& t, const HapPop& s) {
  return hap_pi2_triple_raw(t, s) * 0.5;
}
__device__ __forceinline__ double hap_pi2_triple(const HapPop& t, const HapPop& s) {
  return hap_pi2_triple_num(t, s) * (s.r1 * t.r3);
}
// pi2(i,i,k,k)
__device__ __forceinline__ double hap_pi2_iikk_num(const HapPop& a, const HapPop& b) {
  return a.PA * b.PB;
}
__device__ __forceinline__ double hap_pi2_iikk(const HapPop& a, const HapPop& b) {
  return hap_pi2_iikk_num(a, b) * (a.r2 * b.r2);
}
// pi2(i,i,k,l), k != l
__device__ __forceinline__ double hap_pi2_iikl_raw(const HapPop& a, const HapPop& b,
                                                   const HapPop& c) {
  return a.PA * fma(b.pB, c.qB, b.qB * c.pB);
}
__device__ __forceinline__ double hap_pi2_iikl_num(const HapPop& a, const HapPop& b,
                                                   const HapPop& c) {
  return hap_pi2_iikl_raw(a, b, c) * 0.5;
}
__device__ __forceinline__ double hap_pi2_iikl(const HapPop& a, const HapPop& b, const HapPop& c) {
  return hap_pi2_iikl_num(a, b, c) * (a.r2 * b.r1 * c.r1);
}
// pi2(i,j,k,k), i != j
__device__ __forceinline__ double hap_pi2_ijkk_raw(const HapPop& a, const HapPop& b,
                                                   const HapPop& c) {
  return c.PB * fma(a.pA, b.qA, a.qA * b.pA);
}
__device__ __forceinline__ double hap_pi2_ijkk_num(const HapPop& a, const HapPop& b,
                                                   const HapPop& c) {
  return hap_pi2_ijkk_raw(a, b, c) * 0.5;
}
__device__ __forceinline__ double hap_pi2_ijkk(const HapPop& a, const HapPop& b, const HapPop& c) {
  return hap_pi2_ijkk_num(a, b, c) * (c.r2 * a.r1 * b.r1);
}
// pi2(i,j,i,j) and pi2(i,j,j,i).  The reference's quartic + cubic
// polynomial in the counts of A and B equals, with n = c1+c2+c3+c4 of each,
//   (w_A w_B - (n_A - 1)(n_B - 1)(uA_A uA_B + uB_A uB_B - n_A n_B)) / 4
// (uA = qA - pA, uB = qB - pB, w = uA uB + c2 + c3 - c1 - c4 are the centered
// allele and haplotype sums up to sign), an identity in the eight counts.
__device__ __forceinline__ double hap_pi2_ijij_raw(const HapPop& A, const HapPop& B) {
  const double m = (A.n - 1.0) * (B.n - 1.0);
  const double x = fma(-A.n, B.n, fma(A.uB, B.uB, A.uA * B.uA));
  return fma(-m, x, A.w * B.w);
}
__device__ __forceinline__ double hap_pi2_ijij_num(const HapPop& A, const HapPop& B) {
  return hap_pi2_ijij_raw(A, B) * 0.0625;
}
__device__ __forceinline__ double hap_pi2_ijij(const HapPop& A, const HapPop& B) {
  return hap_pi2_ijij_num(A, B) * (A.r2 * B.r2);
}
// One pop (S) shared between the two locus pairs, A/B the other two pops.
// The reference's polynomial equals, in the same centered sums,
//   (uA_A uB_B w_S - (n_S - 1)(uA_S uA_A n_B + uB_S uB_B n_A - n_S n_A n_B)) / 4.
__device__ __forceinline__ double hap_pi2_shared_raw(const HapPop& S, const HapPop& A,
                                                     const HapPop& B) {
  const double x = fma(-S.n, A.n * B.n, fma(S.uB * B.uB, A.n, (S.uA * A.uA) * B.n));
  return fma(-(S.n - 1.0), x, (A.uA * B.uB) * S.w);
}
__device__ __forceinline__ double hap_pi2_shared_num(const HapPop& S, const HapPop& A,
                                                     const HapPop& B) {
  return hap_pi2_shared_raw(S, A, B) * 0.0625;
}
__device__ __forceinline__ double hap_pi2_shared(const HapPop& S, const HapPop& A, const HapPop& B) {
  return hap_pi2_shared_num(S, A, B) * (S.r2 * A.r1 * B.r1);
}
// pi2(i,j,k,l), all distinct
__device__ __forceinline__ double hap_pi2_alldiff_raw(const HapPop& a, const HapPop& b,
                                                      const HapPop& c, const HapPop& d) {
  return fma(a.pA, b.qA, a.qA * b.pA) * fma(c.pB, d.qB, c.qB * d.pB);
}
__device__ __forceinline__ double hap_pi2_alldiff_num(const HapPop& a, const HapPop& b,
                                                      const HapPop& c, const HapPop& d) {
  return hap_pi2_alldiff_raw(a, b, c, d) * 0.25;
}
__device__ __forceinline__ double hap_pi2_alldiff(const HapPop& a, const HapPop& b,
                                                  const HapPop& c, const HapPop& d) {
  return hap_pi2_alldiff_num(a, b, c, d) * (a.r1 * b.r1 * c.r1 * d.r1);
}

// ---- r^2 (pg_gpu/haplotype_kernels.py:35-46, ld_statistics.r_squared) ------
// r^2 = D^2 / (pA qA pB qB) with frequencies; in counts
// (c1 c4 - c2 c3)^2 / ((c1+c2)(c3+c4)(c1+c3)(c2+c4)).  Undefined (NaN in the
// reference) when a locus is monomorphic among the valid samples, i.e. when
// the denominator is 0.  x[0] = r^2 (0 when undefined), x[1] = 1 when defined.
// The ratio is formed in FP32 (the numerator and denominator are exact).
__device__ __forceinline__ void hap_r2(const HapPop& a, double* x) {
  const double num = -a.D;
  const double den = a.PA * a.PB;
  const float nf = (float)num;
  const float r2 = nf * nf * __builtin_amdgcn_rcpf((float)den);
  const bool ok = den > 0.0;
  x[0] = ok ? (double)r2 : 0.0;
  x[1] = ok ? 1.0 : 0.0;
}
