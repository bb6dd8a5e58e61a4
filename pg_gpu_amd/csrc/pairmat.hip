// Dense per-pair LD matrices for a rectangle of sites (the tile consumers of
// pg_gpu/ld_statistics.py: ZnS, omega, pairwise r^2).  One wave per 32 x 32
// block: FP4 MFMA counts on the packed panels, then one value per pair.
//
//   PGLD_PAIR_SIGMA_D2  unbiased D^2 / pi^2 (Ragsdale & Gravel 2019) of
//                       haplotypes, as _tile_sigma_d2
//                       (pg_gpu/ld_statistics.py:311-341): plane 0 the ratio
//                       of the falling-factorial numerators where n >= 4 and
//                       pi2's numerator is not 0 (else 0), plane 1 = 1 where
//                       n >= 4 (the pairs ZnS counts)
//   PGLD_PAIR_R2_NAIVE  frequency r^2 of haplotypes as _tile_r2_naive /
//                       HaplotypeMatrix.pairwise_r2 (ld_statistics.py:301-309,
//                       haplotype_matrix.py:1092-1170): per-site frequencies
//                       over each site's valid samples, p_AB over samples
//                       valid at both, 0 where a site is monomorphic
//   PGLD_PAIR_R2_DOSAGE r^2 of mean-centred dosages with missing calls set to
//                       zero, as _r2_matrix_diploid (ld_statistics.py:973-1021)
//   PGLD_PAIR_R2        r^2 of haplotypes from the pair's own counts over the
//                       samples valid at both sites, as ld_statistics.r_squared
//                       on tally_gpu_haplotypes counts (ld_statistics.py:
//                       190-217, haplotype_matrix.py:1235-1317): NaN where a
//                       site is monomorphic among those samples
#include "pgld_internal.h"

namespace pgld {

struct PairBlockArgs {
  const int8_t* panels;
  int64_t block_bytes, pop_off;
  int ks, missing, K, kind;
  const int32_t* st;  // [3][V] tallies of the population: #1, #2, #missing
  int64_t V, i0, ni, j0, nj;
  double* out;        // [2][ni][nj]
};

__global__ void __launch_bounds__(64) pair_block_hap_kernel(PairBlockArgs a) {
  const int lane = threadIdx.x;
  const int64_t rb = a.i0 / 32 + blockIdx.y, cb = a.j0 / 32 + blockIdx.x;
  const v4i* A = reinterpret_cast<const v4i*>(a.panels + rb * a.block_bytes + a.pop_off) + lane;
  const v4i* B = reinterpret_cast<const v4i*>(a.panels + cb * a.block_bytes + a.pop_off) + lane;
  v16f hh = zero16f(), hv = zero16f(), vh = zero16f(), vv = zero16f();
  for (int s = 0; s < a.ks; ++s) {
    const v4i ah = A[s * 64], bh = B[s * 64];
    hh = pgld_mfma_fp4(ah, bh, hh);
    if (a.missing) {
      const v4i av = A[(a.ks + s) * 64], bv = B[(a.ks + s) * 64];
      hv = pgld_mfma_fp4(ah, bv, hv);
      vh = pgld_mfma_fp4(av, bh, vh);
      vv = pgld_mfma_fp4(av, bv, vv);
    }
  }
  const int64_t j = cb * 32 + (lane & 31);
  const int64_t plane = a.ni * a.nj;
  for (int r = 0; r < 16; ++r) {
    const int64_t i = rb * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    if (i >= a.i0 + a.ni || j >= a.j0 + a.nj || i >= a.V || j >= a.V) continue;
    const int ci = a.st[i], cj = a.st[j];
    const int ni_v = a.K - a.st[2 * a.V + i], nj_v = a.K - a.st[2 * a.V + j];
    int hvr, vhr, vvr;
    if (a.missing) {
      hvr = (int)hv[r]; vhr = (int)vh[r]; vvr = (int)vv[r];
    } else {
      hvr = ci; vhr = cj; vvr = a.K;
    }
    const double c1 = (double)(int)hh[r], c2 = (double)(hvr - (int)hh[r]);
    const double c3 = (double)(vhr - (int)hh[r]), c4 = (double)(vvr - hvr - vhr + (int)hh[r]);
    const double n = (double)vvr;
    double v = 0.0, ok = 0.0;
    if (a.kind == PGLD_PAIR_SIGMA_D2) {
      const double dd = c1 * (c1 - 1.0) * c4 * (c4 - 1.0) + c2 * (c2 - 1.0) * c3 * (c3 - 1.0) -
                        2.0 * c1 * c2 * c3 * c4;
      const double s12 = c1 + c2, s13 = c1 + c3, s24 = c2 + c4, s34 = c3 + c4;
      const double pi2 = s12 * s13 * s24 * s34 -
                         c1 * c4 * (-1.0 + c1 + 3.0 * c2 + 3.0 * c3 + c4) -
                         c2 * c3 * (-1.0 + 3.0 * c1 + c2 + c3 + 3.0 * c4);
      const bool valid = n >= 4.0;
      v = (valid && pi2 != 0.0) ? dd / pi2 : 0.0;
      ok = valid ? 1.0 : 0.0;
    } else if (a.kind == PGLD_PAIR_R2) {
      const double den = (c1 + c2) * (c3 + c4) * (c1 + c3) * (c2 + c4);
      const double num = c1 * c4 - c2 * c3;
      v = den > 0.0 ? num * num / den : __builtin_nan("");
      ok = den > 0.0 ? 1.0 : 0.0;
    } else {  // PGLD_PAIR_R2_NAIVE
      const double pi = ni_v > 0 ? (double)ci / (double)ni_v : 0.0;
      const double pj = nj_v > 0 ? (double)cj / (double)nj_v : 0.0;
      const double pab = n > 0.0 ? c1 / n : 0.0;
      const double D = pab - pi * pj;
      const double den = (pi * (1.0 - pi)) * (pj * (1.0 - pj));
      v = den > 0.0 ? D * D / den : 0.0;
      ok = den > 0.0 ? 1.0 : 0.0;
    }
    const int64_t o = (i - a.i0) * a.nj + (j - a.j0);
    a.out[o] = v;
    a.out[plane + o] = ok;
  }
}

__global__ void __launch_bounds__(64) pair_block_geno_kernel(PairBlockArgs a) {
  const int lane = threadIdx.x;
  const int64_t rb = a.i0 / 32 + blockIdx.y, cb = a.j0 / 32 + blockIdx.x;
  const v4i* A = reinterpret_cast<const v4i*>(a.panels + rb * a.block_bytes + a.pop_off) + lane;
  const v4i* B = reinterpret_cast<const v4i*>(a.panels + cb * a.block_bytes + a.pop_off) + lane;
  const int nf = a.missing ? 3 : 2;
  v16f pr[3][3];
  for (int x = 0; x < 3; ++x)
    for (int y = 0; y < 3; ++y) pr[x][y] = zero16f();
  for (int s = 0; s < a.ks; ++s) {
    v4i fa[3], fb[3];
    for (int f = 0; f < nf; ++f) {
      fa[f] = A[(f * a.ks + s) * 64];
      fb[f] = B[(f * a.ks + s) * 64];
    }
    for (int x = 0; x < nf; ++x)
      for (int y = 0; y < nf; ++y) pr[x][y] = pgld_mfma_fp4(fa[x], fb[y], pr[x][y]);
  }
  const int64_t j = cb * 32 + (lane & 31);
  const int64_t plane = a.ni * a.nj;
  for (int r = 0; r < 16; ++r) {
    const int64_t i = rb * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    if (i >= a.i0 + a.ni || j >= a.j0 + a.nj || i >= a.V || j >= a.V) continue;
    // sums over the individuals valid at both sites from the centred
    // feature products (u = x - 1, v = [x >= 0]; dosage x = u + 1):
    // sum x = u.v + v.v, sum x y = u.u + u.v + v.u + v.v
    const double uu = pr[0][0][r];
    double uv, vu, vv;
    if (a.missing) {
      uv = pr[0][2][r]; vu = pr[2][0][r]; vv = pr[2][2][r];
    } else {
      uv = 2.0 * a.st[a.V + i] + a.st[i] - a.K;
      vu = 2.0 * a.st[a.V + j] + a.st[j] - a.K;
      vv = a.K;
    }
    const double sx = uv + vv, sy = vu + vv;
    const double sxy = uu + uv + vu + vv;
    // per-site means and sums of squares over each site's valid individuals
    const double ni_v = (double)(a.K - a.st[2 * a.V + i]), nj_v = (double)(a.K - a.st[2 * a.V + j]);
    const double si = a.st[i] + 2.0 * a.st[a.V + i], sj = a.st[j] + 2.0 * a.st[a.V + j];
    const double sii = a.st[i] + 4.0 * a.st[a.V + i], sjj = a.st[j] + 4.0 * a.st[a.V + j];
    const double mi = ni_v > 0.0 ? si / ni_v : 0.0, mj = nj_v > 0.0 ? sj / nj_v : 0.0;
    const double cov = sxy - mj * sx - mi * sy + mi * mj * vv;
    const double vari = sii - 2.0 * mi * si + mi * mi * ni_v;
    const double varj = sjj - 2.0 * mj * sj + mj * mj * nj_v;
    const double den = sqrt(vari * varj);
    double v = 0.0;
    if (den > 0.0) {
      const double rr = cov / den;
      v = rr * rr;
    }
    const int64_t o = (i - a.i0) * a.nj + (j - a.j0);
    a.out[o] = v;
    a.out[plane + o] = den > 0.0 ? 1.0 : 0.0;
  }
}

}  // namespace pgld

using namespace pgld;

extern "C" int pgld_pair_block(const pgld_panel_desc* desc, const int8_t* d_panels,
                               const int32_t* d_site_stats, int32_t pop, int32_t kind, int64_t i0,
                               int64_t ni, int64_t j0, int64_t nj, double* d_out, void* stream) {
  PanelLayout L;
  int rc = make_layout(desc, &L);
  if (rc) return rc;
  if (pop < 0 || pop >= desc->n_pops) return fail(PGLD_EINVAL, "pop out of range");
  if (i0 % 32 || j0 % 32 || i0 < 0 || j0 < 0 || ni < 0 || nj < 0)
    return fail(PGLD_EINVAL, "i0/j0 must be non-negative multiples of 32");
  if (i0 + ni > L.n_sites || j0 + nj > L.n_sites) return fail(PGLD_EINVAL, "block out of range");
  const bool hap_kind = kind == PGLD_PAIR_SIGMA_D2 || kind == PGLD_PAIR_R2_NAIVE ||
                        kind == PGLD_PAIR_R2;
  if (!hap_kind && kind != PGLD_PAIR_R2_DOSAGE) return fail(PGLD_EINVAL, "unknown pair kind");
  if (hap_kind != (desc->mode == PGLD_MODE_HAP))
    return fail(PGLD_EINVAL, "pair kind does not match the panel mode");
  if (!d_panels || !d_site_stats || !d_out) return fail(PGLD_EINVAL, "null device pointer");
  if (ni == 0 || nj == 0) return PGLD_OK;
  PairBlockArgs a;
  a.panels = d_panels;
  a.block_bytes = L.block_bytes;
  a.pop_off = L.pop_off[pop];
  a.ks = L.ksteps[pop];
  a.missing = L.missing[pop];
  a.K = L.pop_size[pop];
  a.kind = kind;
  a.st = d_site_stats + (int64_t)pop * 3 * L.n_sites;
  a.V = L.n_sites;
  a.i0 = i0; a.ni = ni; a.j0 = j0; a.nj = nj;
  a.out = d_out;
  dim3 grid((unsigned)((nj + 31) / 32), (unsigned)((ni + 31) / 32));
  if (hap_kind)
    hipLaunchKernelGGL(pair_block_hap_kernel, grid, dim3(64), 0, as_stream(stream), a);
  else
    hipLaunchKernelGGL(pair_block_geno_kernel, grid, dim3(64), 0, as_stream(stream), a);
  PGLD_LAUNCH_CHECK();
  return PGLD_OK;
}
