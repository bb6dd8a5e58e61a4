// Genotype (unphased diploid) two-locus LD estimators, production form.
//
// Device restatement of pg_gpu/genotype_kernels.py: the per-population
// features of _PopDataGeno (:412-492) and the per-case formulas of the
// DD / Dz / pi2 kernels (:27-405) with the case dispatch of
// compute_all_{dd,dz,pi2}_geno (:561-878).  Genotype classes are in moments
// order g[0] = n22, g[1] = n21, ..., g[8] = n00.  The three single-population
// estimators are integer polynomials of degree 3-4 in the class counts,
// evaluated here in centered dosage moments (below) as FP64 expressions whose
// intermediates are integers, exact for populations below ~10^4 individuals
// (the reference's sub-polynomials, tabulated in oracle/geno_kernel_polys.json,
// are checked against these forms by the test suite).  Divisions by falling
// factorials use the reciprocal table (0 where the factorial is not
// positive); where the reference divides without a guard (Dz between
// populations, normalised frequencies) a population with fewer than the
// needed valid individuals contributes 0 here instead of inf/NaN.
#pragma once
#include "pgld_common.h"

// Centered dosage moments.  With x, y in {0, 1, 2} the dosages (derived
// allele counts) of an individual at the left and right locus, and the sums
// over the valid individuals,
//   n0 = sum 1,  a1 = sum (x-1),  a2 = sum (x-1)^2,  b1 = sum (y-1),
//   b2 = sum (y-1)^2,  ab_kl = sum (x-1)^k (y-1)^l  (k, l = 1, 2),
// the nine class counts map one-to-one (an integer change of variables) onto
// these nine integers, and the reference's estimator polynomials, rewritten in
// them, shrink from 92-155 monomials to 11-21: e.g. D = (n0 ab11 - a1 b1)/4,
// pA = (n0 + a1)/2, and the three single-population sub-polynomials below.
// All of them have integer coefficients, so the FP64 evaluation stays exact.
struct GenoPop {
  double g[9];
  double n;                      // valid individuals for the falling factorials
  double r1, r2, r3, r4;         // 1/(n)_1 .. 1/(n)_4, 0 when not positive
  double n0, a1, a2, b1, b2, ab11, ab12, ab21, ab22;  // centered moments
};

// Centered moments y[0..8] = n0, b1, b2, a1, ab11, ab12, a2, ab21, ab22 of the
// classes c in moments order (c[3 u + v]: left genotype 2 - u, right 2 - v):
// a separable transform, rows first (weights 1, x-1, (x-1)^2 over x = 2, 1, 0)
// then columns.  T is int or double.
template <class T>
__device__ __forceinline__ void geno_centered(const T* c, T* y) {
  T R0[3], R1[3], R2[3];
#pragma unroll
  for (int v = 0; v < 3; ++v) {
    R2[v] = c[v] + c[6 + v];
    R0[v] = R2[v] + c[3 + v];
    R1[v] = c[v] - c[6 + v];
  }
  const T* R[3] = {R0, R1, R2};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    y[3 * k + 2] = R[k][0] + R[k][2];
    y[3 * k] = y[3 * k + 2] + R[k][1];
    y[3 * k + 1] = R[k][0] - R[k][2];
  }
}

__device__ __forceinline__ void geno_set_centered(GenoPop& p, const double* y) {
  p.n0 = y[0]; p.b1 = y[1]; p.b2 = y[2];
  p.a1 = y[3]; p.ab11 = y[4]; p.ab12 = y[5];
  p.a2 = y[6]; p.ab21 = y[7]; p.ab22 = y[8];
}

// c: class counts in moments order; y: their centered moments
// (geno_centered); rtab + 4 n: the reciprocal falling factorials for n.
__device__ __forceinline__ GenoPop geno_pop(const int* c, const int* y, int n,
                                            const double* __restrict__ rtab) {
  GenoPop p;
#pragma unroll
  for (int m = 0; m < 9; ++m) p.g[m] = pgld_count_f64(c[m]);
  double yd[9];
#pragma unroll
  for (int m = 0; m < 9; ++m) yd[m] = pgld_int_f64(y[m]);
  geno_set_centered(p, yd);
  p.n = yd[0];
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
// 4 D = n0 ab11 - a1 b1
__device__ __forceinline__ double g_E(const GenoPop& a) { return fma(a.ab11, a.n0, -(a.a1 * a.b1)); }
__device__ __forceinline__ double g_D(const GenoPop& a) { return g_E(a) * 0.25; }
// derived allele counts (halves of dosage sums) at each locus
__device__ __forceinline__ double g_pA(const GenoPop& a) { return (a.n0 + a.a1) * 0.5; }
__device__ __forceinline__ double g_qA(const GenoPop& a) { return (a.n0 - a.a1) * 0.5; }
__device__ __forceinline__ double g_pB(const GenoPop& a) { return (a.n0 + a.b1) * 0.5; }
__device__ __forceinline__ double g_qB(const GenoPop& a) { return (a.n0 - a.b1) * 0.5; }
// the reference's A and B contrasts
__device__ __forceinline__ double g_A(const GenoPop& a) { return -a.b1; }
__device__ __forceinline__ double g_B(const GenoPop& a) { return -a.a1; }
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
// Products shared by the three single-population estimators.
struct GenoShared {
  double n, ab, E, u, t4, t56, s11, a1s, b1s;
};
__device__ __forceinline__ GenoShared geno_shared(const GenoPop& a) {
  GenoShared s;
  s.n = a.n0;
  s.ab = a.a1 * a.b1;
  s.E = fma(a.ab11, a.n0, -s.ab);                       // 4 D
  s.a1s = a.a1 * a.a1;
  s.b1s = a.b1 * a.b1;
  s.u = (a.a2 * a.b2 - s.a1s * a.b2) - a.a2 * s.b1s;    // a2 b2 - a1^2 b2 - a2 b1^2
  s.t4 = s.ab * a.ab11;                                 // a1 b1 ab11
  s.t56 = fma(a.a1, a.ab12, a.ab21 * a.b1);             // a1 ab12 + ab21 b1
  s.s11 = a.ab11 * a.ab11;
  return s;
}
// DD: (poly + (4D)^2) / (4 (n)_4) with the reference's poly
// (genotype_kernels.py:27-71) = 2(t4 + (n-1) t56) + u - ab11^2 (3n-2) - ab22 n(n-1)
__device__ __forceinline__ double geno_dd_same(const GenoPop& a) {
  const GenoShared s = geno_shared(a);
  const double nm1 = s.n - 1.0;
  double poly = fma(2.0, fma(nm1, s.t56, s.t4), s.u);
  poly = fma(-s.s11, fma(3.0, s.n, -2.0), poly);
  poly = fma(-a.ab22, s.n * nm1, poly);
  return (fma(s.E, s.E, poly) * 0.25) * a.r4;
}
__device__ __forceinline__ double geno_dd_pair(const GenoPop& a, const GenoPop& b) {
  return 4.0 * g_D(a) * g_D(b) * (a.r2 * b.r2);
}

// ---- Dz -------------------------------------------------------------------
// Dz: (quartic + a1 b1 (4D)) / (2 (n)_4) with the reference's quartic
// (genotype_kernels.py:235-272) = t4 - u - (n+1) t56 - ab11^2 (n-1) + 2 ab22 n
__device__ __forceinline__ double geno_dz_same(const GenoPop& a) {
  const GenoShared s = geno_shared(a);
  double q = fma(-(s.n + 1.0), s.t56, s.t4 - s.u);
  q = fma(-s.s11, s.n - 1.0, q);
  q = fma(2.0 * a.ab22, s.n, q);
  return (fma(s.ab, s.E, q) * 0.5) * a.r4;
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
// pi2: (16 pA qA pB qB + corr) / (16 (n)_4), 16 pA qA pB qB =
// (n^2 - a1^2)(n^2 - b1^2), with the reference's corr (genotype_kernels.py:
// 364-405) = u + 4 (t56 - t4) + (a1^2 + b1^2)(5n - 6) + (a2 + b2)(n-2)(n-3)
//            + 2 ab11^2 - 6 ab22 - n (6n^2 - 11n + 6)
__device__ __forceinline__ double geno_pi2_same(const GenoPop& a) {
  const GenoShared s = geno_shared(a);
  const double n = s.n, nn = n * n;
  double c = fma(4.0, s.t56 - s.t4, s.u);
  c = fma(s.a1s + s.b1s, fma(5.0, n, -6.0), c);
  c = fma(a.a2 + a.b2, fma(n, n - 5.0, 6.0), c);
  c = fma(2.0, s.s11, c);
  c = fma(-6.0, a.ab22, c);
  c = fma(-n, fma(n, fma(6.0, n, -11.0), 6.0), c);
  return (fma(nn - s.a1s, nn - s.b1s, c) * (1.0 / 16.0)) * a.r4;
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
  // n Sxy - Sx Sy = n0 ab11 - a1 b1; n Sxx - Sx^2 = n0 a2 - a1^2 (exact integers)
  const double num = g_E(p);
  const double vx = fma(p.n0, p.a2, -(p.a1 * p.a1)), vy = fma(p.n0, p.b2, -(p.b1 * p.b1));
  const float nf = (float)num;
  const float r2 = nf * nf * __builtin_amdgcn_rcpf((float)vx * (float)vy);
  const bool ok = vx > 0.0 && vy > 0.0;
  x[0] = ok ? (double)r2 : 0.0;
  x[1] = ok ? 1.0 : 0.0;
}
