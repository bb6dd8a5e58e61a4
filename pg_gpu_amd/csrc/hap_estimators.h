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
};

// rtab holds 4 doubles per n: 1/(n)_1, 1/(n)_2, 1/(n)_3, 1/(n)_4.
__device__ __forceinline__ HapPop hap_pop(int n11, int n10, int n01, int n00, int n,
                                          const double* __restrict__ rtab) {
  HapPop p;
  p.c1 = pgld_count_f64(n11); p.c2 = pgld_count_f64(n10);
  p.c3 = pgld_count_f64(n01); p.c4 = pgld_count_f64(n00);
  p.D = p.c2 * p.c3 - p.c1 * p.c4;
  p.pA = p.c1 + p.c2; p.qA = p.c3 + p.c4;
  p.pB = p.c1 + p.c3; p.qB = p.c2 + p.c4;
  const double2* t = reinterpret_cast<const double2*>(rtab + 4 * n);
  const double2 a = t[0], b = t[1];
  p.r1 = a.x; p.r2 = a.y; p.r3 = b.x; p.r4 = b.y;
  return p;
}

// ---- DD -------------------------------------------------------------------
__device__ __forceinline__ double hap_dd_same(const HapPop& a) {
  const double num = a.c1 * (a.c1 - 1.0) * a.c4 * (a.c4 - 1.0) +
                     a.c2 * (a.c2 - 1.0) * a.c3 * (a.c3 - 1.0) -
                     2.0 * a.c1 * a.c2 * a.c3 * a.c4;
  return num * a.r4;
}
__device__ __forceinline__ double hap_dd_pair(const HapPop& a, const HapPop& b) {
  return (a.D * b.D) * (a.r2 * b.r2);
}

// ---- Dz -------------------------------------------------------------------
__device__ __forceinline__ double hap_dz_same(const HapPop& a) {
  const double x = a.c1 * a.c4 - a.c2 * a.c3;
  const double num = x * (a.qA - a.pA) * (a.qB - a.pB) +
                     x * ((a.c2 + a.c3) - (a.c1 + a.c4)) + 2.0 * (a.c2 * a.c3 + a.c1 * a.c4);
  return num * a.r4;
}
// Dz(i,i,k)
__device__ __forceinline__ double hap_dz_iik(const HapPop& a, const HapPop& b) {
  const double num = (a.qA - a.pA) * (-a.D) * (b.qB - b.pB);
  return num * (a.r3 * b.r1);
}
// Dz(i,j,i)
__device__ __forceinline__ double hap_dz_iji(const HapPop& a, const HapPop& b) {
  const double num = (a.qB - a.pB) * (-a.D) * (b.qA - b.pA);
  return num * (a.r3 * b.r1);
}
// Dz(i,j,j)
__device__ __forceinline__ double hap_dz_ijj(const HapPop& a, const HapPop& b) {
  const double t = (b.c2 + b.c3) - (b.c1 + b.c4);
  const double num = (-a.D) * (t + (b.qB - b.pB) * (b.qA - b.pA));
  return num * (a.r2 * b.r2);
}
// Dz(i,j,k), all distinct
__device__ __forceinline__ double hap_dz_ijk(const HapPop& a, const HapPop& b, const HapPop& c) {
  const double num = -(a.D * (b.pA - b.qA) * (c.pB - c.qB));
  return num * (a.r2 * b.r1 * c.r1);
}

// ---- pi2 ------------------------------------------------------------------
__device__ __forceinline__ double hap_pi2_same(const HapPop& a) {
  const double num = a.pA * a.pB * a.qB * a.qA -
                     a.c1 * a.c4 * (-1.0 + a.c1 + 3.0 * a.c2 + 3.0 * a.c3 + a.c4) -
                     a.c2 * a.c3 * (-1.0 + 3.0 * a.c1 + a.c2 + a.c3 + 3.0 * a.c4);
  return num * a.r4;
}
// pop t three times, pop s once (all "triple" cases share this formula)
__device__ __forceinline__ double hap_pi2_triple(const HapPop& t, const HapPop& s) {
  const double tpA = t.pA, tqA = t.qA;
  const double num = (s.pB * (tpA * (t.c2 + t.c4) * tqA - tpA * t.c4 - t.c2 * tqA) +
                      s.qB * (tpA * (t.c1 + t.c3) * tqA + tpA * t.c4 + t.c2 * tqA) -
                      2.0 * tpA * tqA * s.qB) * 0.5;
  return num * (s.r1 * t.r3);
}
// pi2(i,i,k,k)
__device__ __forceinline__ double hap_pi2_iikk(const HapPop& a, const HapPop& b) {
  return (a.pA * a.qA) * (b.pB * b.qB) * (a.r2 * b.r2);
}
// pi2(i,i,k,l), k != l
__device__ __forceinline__ double hap_pi2_iikl(const HapPop& a, const HapPop& b, const HapPop& c) {
  return (a.pA * a.qA) * (b.pB * c.qB + b.qB * c.pB) * 0.5 * (a.r2 * b.r1 * c.r1);
}
// pi2(i,j,k,k), i != j
__device__ __forceinline__ double hap_pi2_ijkk(const HapPop& a, const HapPop& b, const HapPop& c) {
  return (c.pB * c.qB) * (a.pA * b.qA + a.qA * b.pA) * 0.5 * (c.r2 * a.r1 * b.r1);
}
// pi2(i,j,i,j) and pi2(i,j,j,i)
__device__ __forceinline__ double hap_pi2_ijij(const HapPop& A, const HapPop& B) {
  const double a1 = A.c1, a2 = A.c2, a3 = A.c3, a4 = A.c4;
  const double b1 = B.c1, b2 = B.c2, b3 = B.c3, b4 = B.c4;
  const double quartic = A.qB * A.qA * B.pA * B.pB + A.pB * A.qA * B.pA * B.qB +
                         A.pA * A.qB * B.pB * B.qA + A.pA * A.pB * B.qB * B.qA;
  const double cubic =
      b1 * (a4 - a2 * a3 - a2 * a4 - a3 * a4 - a4 * a4 - a4 * b1) +
      b2 * (a3 - a1 * a3 - a3 * a3 - a1 * a4 - a3 * a4 - a3 * b1 - a4 * b1 - a3 * b2) +
      b3 * (a2 - a1 * a2 - a2 * a2 - a1 * a4 - a2 * a4 - a2 * b1 - a4 * b1 - a1 * b2 -
            a4 * b2 - a2 * b3) +
      b4 * (a1 - a1 * a1 - a1 * a2 - a1 * a3 - a2 * a3 - a2 * b1 - a3 * b1 - a1 * b2 -
            a3 * b2 - a1 * b3 - a2 * b3 - a1 * b4);
  return (quartic + cubic) * 0.25 * (A.r2 * B.r2);
}
// one pop (S) shared between the two locus pairs, A/B the other two pops
__device__ __forceinline__ double hap_pi2_shared(const HapPop& S, const HapPop& A, const HapPop& B) {
  const double s1 = S.c1, s2 = S.c2, s3 = S.c3, s4 = S.c4;
  const double inner = (s3 - 1.0) * B.c1 + s3 * B.c2 - B.c3 + s3 * B.c3 + s3 * B.c4 + s1 * B.qB;
  const double num =
      s4 * s4 * A.pA * B.pB + s2 * s2 * A.qA * B.pB +
      (s1 + s3 - 1.0) * (s3 * A.pA + s1 * A.qA) * B.qB +
      s4 * (s1 * A.qA * B.pB + (A.c1 + A.c2) * inner) +
      s2 * (s4 * (A.pA + A.qA) * B.pB + s3 * ((A.c1 + A.c2) * B.pB + A.qA * B.qB) +
            A.qA * ((s1 - 1.0) * B.c1 - B.c3 + s1 * (B.c2 + B.c3 + B.c4)));
  return num * 0.25 * (S.r2 * A.r1 * B.r1);
}
// pi2(i,j,k,l), all distinct
__device__ __forceinline__ double hap_pi2_alldiff(const HapPop& a, const HapPop& b,
                                                  const HapPop& c, const HapPop& d) {
  return (a.pA * b.qA + a.qA * b.pA) * (c.pB * d.qB + c.qB * d.pB) * 0.25 *
         (a.r1 * b.r1 * c.r1 * d.r1);
}

// ---- r^2 (pg_gpu/haplotype_kernels.py:35-46, ld_statistics.r_squared) ------
// r^2 = D^2 / (pA qA pB qB) with frequencies; in counts
// (c1 c4 - c2 c3)^2 / ((c1+c2)(c3+c4)(c1+c3)(c2+c4)).  Undefined (NaN in the
// reference) when a locus is monomorphic among the valid samples, i.e. when
// the denominator is 0.  x[0] = r^2 (0 when undefined), x[1] = 1 when defined.
// The ratio is formed in FP32 (the numerator and denominator are exact).
__device__ __forceinline__ void hap_r2(const HapPop& a, double* x) {
  const double num = a.c1 * a.c4 - a.c2 * a.c3;
  const double den = (a.pA * a.qA) * (a.pB * a.qB);
  const float nf = (float)num;
  const float r2 = nf * nf * __builtin_amdgcn_rcpf((float)den);
  const bool ok = den > 0.0;
  x[0] = ok ? (double)r2 : 0.0;
  x[1] = ok ? 1.0 : 0.0;
}
