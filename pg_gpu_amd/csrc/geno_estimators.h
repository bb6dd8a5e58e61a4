// Genotype (unphased diploid) two-locus LD estimators, production form.
//
// Device restatement of pg_gpu/genotype_kernels.py: the per-population
// features of _PopDataGeno (:412-492) and the per-case formulas of the
// DD / Dz / pi2 kernels (:27-405) with the case dispatch of
// compute_all_{dd,dz,pi2}_geno (:561-878).  Genotype classes are in moments
// order g[0] = n22, g[1] = n21, ..., g[8] = n00.  The three single-population
// estimators are integer polynomials of degree 3-4 in the class counts; their
// sub-polynomials are generated (geno_polys.inc) as nested FP64 expressions
// that are exact for class counts below ~10^4.  Divisions by falling
// factorials use the reciprocal table (0 where the factorial is not
// positive); where the reference divides without a guard (Dz between
// populations, normalised frequencies) a population with fewer than the
// needed valid individuals contributes 0 here instead of inf/NaN.
#pragma once
#include "pgld_common.h"
#include "geno_polys.inc"

struct GenoPop {
  double g[9];
  double n;                      // valid individuals (= sum of the classes)
  double r1, r2, r3, r4;         // 1/(n)_1 .. 1/(n)_4, 0 when not positive
};

__device__ __forceinline__ GenoPop geno_pop(const int* c /* moments order */, int n,
                                            const double* __restrict__ rtab) {
  GenoPop p;
#pragma unroll
  for (int m = 0; m < 9; ++m) p.g[m] = pgld_count_f64(c[m]);
  p.n = n;
  const double2* t = reinterpret_cast<const double2*>(rtab + 4 * n);
  const double2 a = t[0], b = t[1];
  p.r1 = a.x; p.r2 = a.y; p.r3 = b.x; p.r4 = b.y;
  return p;
}

// ---- features (each an inline function so unused ones cost nothing) -------
__device__ __forceinline__ double gx_AB(const GenoPop& a) {
  return a.g[0] + 0.5 * a.g[1] + 0.5 * a.g[3] + 0.25 * a.g[4];
}
__device__ __forceinline__ double gx_Ab(const GenoPop& a) {
  return 0.5 * a.g[1] + a.g[2] + 0.25 * a.g[4] + 0.5 * a.g[5];
}
__device__ __forceinline__ double gx_aB(const GenoPop& a) {
  return 0.5 * a.g[3] + 0.25 * a.g[4] + a.g[6] + 0.5 * a.g[7];
}
__device__ __forceinline__ double gx_ab(const GenoPop& a) {
  return 0.25 * a.g[4] + 0.5 * a.g[5] + 0.5 * a.g[7] + a.g[8];
}
__device__ __forceinline__ double g_D(const GenoPop& a) {
  return gx_AB(a) * gx_ab(a) - gx_Ab(a) * gx_aB(a);
}
__device__ __forceinline__ double g_pA(const GenoPop& a) {
  return a.g[0] + a.g[1] + a.g[2] + 0.5 * (a.g[3] + a.g[4] + a.g[5]);
}
__device__ __forceinline__ double g_qA(const GenoPop& a) {
  return 0.5 * (a.g[3] + a.g[4] + a.g[5]) + a.g[6] + a.g[7] + a.g[8];
}
__device__ __forceinline__ double g_pB(const GenoPop& a) {
  return a.g[0] + 0.5 * a.g[1] + a.g[3] + 0.5 * a.g[4] + a.g[6] + 0.5 * a.g[7];
}
__device__ __forceinline__ double g_qB(const GenoPop& a) {
  return 0.5 * a.g[1] + a.g[2] + 0.5 * a.g[4] + a.g[5] + 0.5 * a.g[7] + a.g[8];
}
__device__ __forceinline__ double g_A(const GenoPop& a) {
  return -a.g[0] + a.g[2] - a.g[3] + a.g[5] - a.g[6] + a.g[8];
}
__device__ __forceinline__ double g_B(const GenoPop& a) {
  return -a.g[0] - a.g[1] - a.g[2] + a.g[6] + a.g[7] + a.g[8];
}
// 1 / max(4 n (n-1), 1)
__device__ __forceinline__ double g_rnn1(const GenoPop& a) { return a.n >= 2.0 ? 0.25 * a.r2 : 1.0; }
__device__ __forceinline__ double g_BL(const GenoPop& a) {
  return (4.0 * g_pA(a) * g_qA(a) - (a.g[3] + a.g[4] + a.g[5])) * g_rnn1(a);
}
__device__ __forceinline__ double g_BR(const GenoPop& a) {
  return (4.0 * g_pB(a) * g_qB(a) - (a.g[1] + a.g[4] + a.g[7])) * g_rnn1(a);
}
// normalised frequencies p = pA/n, r = qA/n, q = pB/n, s = qB/n
__device__ __forceinline__ double g_p(const GenoPop& a) { return g_pA(a) * a.r1; }
__device__ __forceinline__ double g_r(const GenoPop& a) { return g_qA(a) * a.r1; }
__device__ __forceinline__ double g_q(const GenoPop& a) { return g_pB(a) * a.r1; }
__device__ __forceinline__ double g_s(const GenoPop& a) { return g_qB(a) * a.r1; }

// ---- DD -------------------------------------------------------------------
__device__ __forceinline__ double geno_dd_same(const GenoPop& a) {
  const double Dg = g_D(a);
  return 4.0 * (geno_dd_poly(a.g) * (1.0 / 16.0) + Dg * Dg) * a.r4;
}
__device__ __forceinline__ double geno_dd_pair(const GenoPop& a, const GenoPop& b) {
  return 4.0 * g_D(a) * g_D(b) * (a.r2 * b.r2);
}

// ---- Dz -------------------------------------------------------------------
__device__ __forceinline__ double geno_dz_same(const GenoPop& a) {
  return 2.0 * (0.25 * geno_dz_quartic(a.g) + g_A(a) * g_B(a) * g_D(a)) * a.r4;
}
// Dz(i,i,k): 2 U_i A_k / (n_k (n_i)_3)
__device__ __forceinline__ double geno_dz_iik(const GenoPop& a, const GenoPop& b) {
  const double* g = a.g;
  const double MA = 2.0 * g[1] * g[6] + 4.0 * g[2] * g[6] + 2.0 * g[4] * g[6] + 4.0 * g[5] * g[6] -
                    2.0 * g[0] * g[7] + 2.0 * g[2] * g[7] - 2.0 * g[3] * g[7] + 2.0 * g[5] * g[7] -
                    4.0 * g[0] * g[8] - 2.0 * g[1] * g[8] - 4.0 * g[3] * g[8] - 2.0 * g[4] * g[8];
  const double U = g_D(a) * (1.0 + g_B(a)) + 0.25 * MA;
  return 2.0 * U * g_A(b) * (a.r3 * b.r1);
}
// Dz(i,j,i): 2 V_i B_j / (n_j (n_i)_3)
__device__ __forceinline__ double geno_dz_iji(const GenoPop& a, const GenoPop& b) {
  const double* g = a.g;
  const double MB = 2.0 * g[2] * g[3] + 2.0 * g[2] * g[4] - 2.0 * g[0] * g[5] - 2.0 * g[1] * g[5] +
                    4.0 * g[2] * g[6] + 2.0 * g[5] * g[6] + 4.0 * g[2] * g[7] + 2.0 * g[5] * g[7] -
                    4.0 * g[0] * g[8] - 4.0 * g[1] * g[8] - 2.0 * g[3] * g[8] - 2.0 * g[4] * g[8];
  const double Vc = g_D(a) * (1.0 + g_A(a)) + 0.25 * MB;
  return 2.0 * Vc * g_B(b) * (a.r3 * b.r1);
}
// Dz(i,j,j): 2 D_i W_j / ((n_i)_2 (n_j)_2)
__device__ __forceinline__ double geno_dz_ijj(const GenoPop& a, const GenoPop& b) {
  const double W = (-b.g[0] + b.g[2] + b.g[6] - b.g[8]) + g_A(b) * g_B(b);
  return 2.0 * g_D(a) * W * (a.r2 * b.r2);
}
// Dz(i,j,k), all distinct: 2 D_i B_j A_k / ((n_i)_2 n_j n_k)
__device__ __forceinline__ double geno_dz_ijk(const GenoPop& a, const GenoPop& b, const GenoPop& c) {
  return 2.0 * g_D(a) * g_B(b) * g_A(c) * (a.r2 * b.r1 * c.r1);
}

// ---- pi2 ------------------------------------------------------------------
__device__ __forceinline__ double geno_pi2_same(const GenoPop& a) {
  const double a1 = g_pA(a), a2 = g_pB(a), a3 = g_qB(a), a4 = g_qA(a);
  return (a1 * a2 * a3 * a4 + geno_pi2_corr(a.g) * (1.0 / 16.0)) * a.r4;
}

// pop t three times; the first two indices are t: (t,t,t,s) / (t,t,s,t)
__device__ __forceinline__ double geno_pi2_iiij(const GenoPop& t, const GenoPop& s) {
  const double* g = t.g;
  const double P = g_pA(t), R = g_qA(t), Q = g_pB(t), S = g_qB(t);
  const double L0 = g[4] + 2.0 * g[5] + 2.0 * g[7] + 4.0 * g[8];
  const double L1 = g[4] + 2.0 * g[5];
  const double negR3 = 3.0 * (g[3] + g[4] + g[5]) + 4.0 * (g[6] + g[7] + g[8]);
  const double negR5 = 5.0 * (g[3] + g[4] + g[5]) + 8.0 * (g[6] + g[7] + g[8]);
  const double b5A = -2.0 * g[1] - 4.0 * g[2] + 4.0 * g[3] + g[4] - 2.0 * g[5] + 8.0 * g[6] +
                     4.0 * g[7] + 4.0;
  const double b6A = -g[4] - 2.0 * g[5] - 4.0 * g[6] - 4.0 * g[7] - 4.0 * g[8];
  const double AQ3 = -(S * negR3) * 0.125 + P * S * R * 0.5 + (R * L0) * 0.125 + L1 * 0.125 -
                     (P * L0) * 0.125;
  const double AS3 = -(Q * negR5) * 0.125 + P * Q * R * 0.5 + (R * b5A) * 0.125 + b6A * 0.125 +
                     (P * L0) * 0.125;
  return (AQ3 * g_pB(s) + AS3 * g_qB(s)) * (t.r3 * s.r1);
}
// pop t three times, otherwise: (t,s,t,t) / (s,t,t,t)
__device__ __forceinline__ double geno_pi2_ijii(const GenoPop& t, const GenoPop& s) {
  const double* g = t.g;
  const double P = g_pA(t), R = g_qA(t), Q = g_pB(t), S = g_qB(t);
  const double L0 = g[4] + 2.0 * g[5] + 2.0 * g[7] + 4.0 * g[8];
  const double M1 = g[4] + 2.0 * g[7];
  const double negL3 = 3.0 * (g[1] + g[4] + g[7]) + 4.0 * (g[2] + g[5] + g[8]);
  const double negL5 = 5.0 * (g[1] + g[4] + g[7]) + 8.0 * (g[2] + g[5] + g[8]);
  const double b5B = 4.0 * g[1] + 8.0 * g[2] - 2.0 * g[3] + g[4] + 4.0 * g[5] - 4.0 * g[6] -
                     2.0 * g[7] + 4.0;
  const double b6B = -4.0 * g[2] - g[4] - 4.0 * g[5] - 2.0 * g[7] - 4.0 * g[8];
  const double BP3 = -(R * negL3) * 0.125 + Q * S * R * 0.5 + (S * L0) * 0.125 + M1 * 0.125 -
                     (Q * L0) * 0.125;
  const double BR3 = -(P * negL5) * 0.125 + P * Q * S * 0.5 + (S * b5B) * 0.125 + b6B * 0.125 +
                     (Q * L0) * 0.125;
  return (BP3 * g_pA(s) + BR3 * g_qA(s)) * (t.r3 * s.r1);
}
__device__ __forceinline__ double geno_pi2_iikk(const GenoPop& a, const GenoPop& b) {
  return g_BL(a) * g_BR(b);
}
__device__ __forceinline__ double geno_pi2_iikl(const GenoPop& a, const GenoPop& b, const GenoPop& c) {
  return g_BL(a) * 0.5 * (g_q(b) * g_s(c) + g_s(b) * g_q(c));
}
__device__ __forceinline__ double geno_pi2_ijkk(const GenoPop& a, const GenoPop& b, const GenoPop& c) {
  return 0.5 * (g_p(a) * g_r(b) + g_r(a) * g_p(b)) * g_BR(c);
}
// one pop (S) shared between the two locus pairs; A first pair, B second pair
__device__ __forceinline__ double geno_pi2_shared(const GenoPop& S, const GenoPop& A, const GenoPop& B) {
  const double rn = g_rnn1(S);
  const double pA = g_pA(S), qA = g_qA(S), pB = g_pB(S), qB = g_qB(S);
  const double C00 = (pA * pB - gx_AB(S)) * rn, C01 = (pA * qB - gx_Ab(S)) * rn;
  const double C10 = (qA * pB - gx_aB(S)) * rn, C11 = (qA * qB - gx_ab(S)) * rn;
  const double sk = g_s(B), qk = g_q(B);
  return g_r(A) * (C00 * sk + C01 * qk) + g_p(A) * (C10 * sk + C11 * qk);
}
__device__ __forceinline__ double geno_pi2_ijij(const GenoPop& I, const GenoPop& J) {
  const double Pi = g_pA(I), Qi = g_pB(I), Ri = g_qA(I), Si = g_qB(I);
  const double Pj = g_pA(J), Qj = g_pB(J), Rj = g_qA(J), Sj = g_qB(J);
  const double X01i = gx_Ab(I), X10i = gx_aB(I), X11i = gx_ab(I);
  const double X11j = gx_ab(J);
  const double Bi = gx_AB(I) - X01i - X10i + X11i;
  const double Bj = gx_AB(J) - gx_Ab(J) - gx_aB(J) + X11j;
  const double PQi = Pi * Qi, PRi = Pi * Ri, PSi = Pi * Si, QRi = Qi * Ri, RSi = Ri * Si;
  const double PQj = Pj * Qj, PSj = Pj * Sj, QRj = Qj * Rj, RSj = Rj * Sj;
  const double Tci = Ri * (1.0 + X01i - X10i), Uci = Si * (1.0 + Ri);
  double num = -0.25 * RSi * Qj - 0.25 * X11i * PQj + 0.25 * RSi * PQj - 0.25 * Ri * PSj;
  num += -0.25 * PQi * X11j - 0.25 * RSi * X11j + 0.25 * PSi * X11j + 0.25 * QRi * X11j;
  num += -0.25 * PRi * Sj + 0.25 * Tci * Sj + 0.25 * X11i * PSj + 0.25 * QRi * PSj;
  num += 0.25 * Uci * Rj - 0.25 * X01i * QRj + 0.25 * PSi * QRj;
  num += -0.25 * PSi * Rj - 0.25 * Pi * RSj + 0.25 * X01i * RSj + 0.25 * PQi * RSj;
  num += 0.25 * X11i * Bj + 0.25 * Bi * X11j - 1.0 * X11i * X11j;
  return num * (I.r2 * J.r2);
}
__device__ __forceinline__ double geno_pi2_alldiff(const GenoPop& a, const GenoPop& b,
                                                   const GenoPop& c, const GenoPop& d) {
  return 0.25 * (g_p(a) * g_r(b) + g_r(a) * g_p(b)) * (g_q(c) * g_s(d) + g_s(c) * g_q(d));
}

// ---- r^2 of dosages (pairwise-complete Pearson) ----------------------------
// Over the individuals valid at both loci, x and y the 0/1/2 dosages:
// r = (n Sxy - Sx Sy) / sqrt((n Sxx - Sx^2)(n Syy - Sy^2)).  Without missing
// calls this is Rogers-Huff r (pg_gpu/ld_statistics.py:492-521).  Moments
// order g[0..8] = n22 n21 n20 n12 n11 n10 n02 n01 n00 (first digit: locus i).
// x[0] = r^2 (0 when a locus has no dosage variance), x[1] = 1 when defined.
__device__ __forceinline__ void geno_r2(const GenoPop& p, double* x) {
  const double* g = p.g;
  const double X2 = g[0] + g[1] + g[2], X1 = g[3] + g[4] + g[5];
  const double Y2 = g[0] + g[3] + g[6], Y1 = g[1] + g[4] + g[7];
  const double sx = 2.0 * X2 + X1, sxx = 4.0 * X2 + X1;
  const double sy = 2.0 * Y2 + Y1, syy = 4.0 * Y2 + Y1;
  const double sxy = 4.0 * g[0] + 2.0 * (g[1] + g[3]) + g[4];
  const double n = p.n;
  const double num = n * sxy - sx * sy;            // exact integers below 2^53
  const double vx = n * sxx - sx * sx, vy = n * syy - sy * sy;
  const float nf = (float)num;
  const float r2 = nf * nf * __builtin_amdgcn_rcpf((float)vx * (float)vy);
  const bool ok = vx > 0.0 && vy > 0.0;
  x[0] = ok ? (double)r2 : 0.0;
  x[1] = ok ? 1.0 : 0.0;
}
