// Once-per-dataset passes over the int8 sample-major matrix: the biallelic
// site filters and the haplotype <-> genotype conversions of the containers.
// All three are HBM-bound byte work: one thread per site (filters) or per
// output byte (conversions), consecutive threads on consecutive sites of a
// row, so every row is read with coalesced loads.
#include "pgld_internal.h"

namespace pgld {

// keep[s] = 1 if site s passes `rule` over the selected rows (all rows when
// rows == nullptr):
//   PGLD_FILTER_HAP_BIALLELIC: no value >= 2, at least one 0 and one 1
//     (HaplotypeMatrix.apply_biallelic_filter, haplotype_matrix.py:751-765);
//   PGLD_FILTER_GENO_BIALLELIC: with n = #(x >= 0) and alt = sum of x >= 0,
//     0 < alt < 2 n and n >= 2 (GenotypeMatrix.apply_biallelic_filter,
//     genotype_matrix.py:516-523, and the union filter of moments_ld.py:256-276).
__global__ void __launch_bounds__(256) biallelic_mask_kernel(
    const int8_t* __restrict__ mat, int64_t n_sites, int64_t ld,
    const int32_t* __restrict__ rows, int32_t n_sel, int32_t rule, uint8_t* __restrict__ keep) {
  const int64_t site = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (site >= n_sites) return;
  int n0 = 0, n1 = 0, nhi = 0, nv = 0, alt = 0;
  for (int r = 0; r < n_sel; ++r) {
    const int64_t row = rows ? rows[r] : r;
    const int v = mat[row * ld + site];
    n0 += (v == 0);
    n1 += (v == 1);
    nhi += (v >= 2);
    nv += (v >= 0);
    alt += v > 0 ? v : 0;
  }
  bool k;
  if (rule == PGLD_FILTER_HAP_BIALLELIC)
    k = nhi == 0 && n0 > 0 && n1 > 0;
  else
    k = alt > 0 && alt < 2 * nv && nv >= 2;
  keep[site] = k ? 1 : 0;
}

// geno[i][s] = max(h[2i][s], 0) + max(h[2i+1][s], 0), or -1 if either is < 0.
__global__ void __launch_bounds__(256) hap_to_geno_kernel(const int8_t* __restrict__ hap,
                                                          int64_t ld_hap, int64_t n_ind,
                                                          int64_t n_sites,
                                                          int8_t* __restrict__ geno,
                                                          int64_t ld_geno) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_sites) return;
  for (int64_t i = blockIdx.y; i < n_ind; i += gridDim.y) {
    const int a = hap[2 * i * ld_hap + s], b = hap[(2 * i + 1) * ld_hap + s];
    geno[i * ld_geno + s] =
        (a < 0 || b < 0) ? (int8_t)-1 : (int8_t)((a > 0 ? a : 0) + (b > 0 ? b : 0));
  }
}

// hap[2i][s] = [g >= 1], hap[2i+1][s] = [g >= 2], both -1 where g < 0.
__global__ void __launch_bounds__(256) geno_to_hap_kernel(const int8_t* __restrict__ geno,
                                                          int64_t ld_geno, int64_t n_ind,
                                                          int64_t n_sites, int8_t* __restrict__ hap,
                                                          int64_t ld_hap) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_sites) return;
  for (int64_t i = blockIdx.y; i < n_ind; i += gridDim.y) {
    const int g = geno[i * ld_geno + s];
    hap[2 * i * ld_hap + s] = g < 0 ? (int8_t)-1 : (int8_t)(g >= 1);
    hap[(2 * i + 1) * ld_hap + s] = g < 0 ? (int8_t)-1 : (int8_t)(g >= 2);
  }
}

static int check_matrix(const void* p, int64_t n_rows, int64_t n_sites, int64_t ld) {
  if (n_rows < 0 || n_sites < 0) return fail(PGLD_EINVAL, "negative matrix shape");
  if (n_rows > 0 && n_sites > 0 && !p) return fail(PGLD_EINVAL, "null matrix pointer");
  if (n_rows > 1 && ld < n_sites) return fail(PGLD_EINVAL, "row stride smaller than n_sites");
  return PGLD_OK;
}

}  // namespace pgld

using namespace pgld;

extern "C" int pgld_biallelic_mask(const int8_t* d_mat, int64_t n_rows, int64_t n_sites,
                                   int64_t ld, const int32_t* d_rows, int32_t n_sel,
                                   int32_t rule, uint8_t* d_keep, void* stream) {
  int rc = check_matrix(d_mat, n_rows, n_sites, ld);
  if (rc) return rc;
  if (rule != PGLD_FILTER_HAP_BIALLELIC && rule != PGLD_FILTER_GENO_BIALLELIC)
    return fail(PGLD_EINVAL, "unknown filter rule");
  if (n_sel < 0 || (!d_rows && n_sel != n_rows) || (n_sel > 0 && n_rows == 0))
    return fail(PGLD_EINVAL, "selected rows do not match the matrix");
  if (n_sites == 0) return PGLD_OK;
  if (!d_keep) return fail(PGLD_EINVAL, "null keep mask");
  hipLaunchKernelGGL(biallelic_mask_kernel, dim3((unsigned)((n_sites + 255) / 256)), dim3(256),
                     0, as_stream(stream), d_mat, n_sites, ld, d_rows, n_sel, rule, d_keep);
  PGLD_LAUNCH_CHECK();
  return PGLD_OK;
}

extern "C" int pgld_hap_to_geno(const int8_t* d_hap, int64_t n_hap, int64_t n_sites,
                                int64_t ld_hap, int8_t* d_geno, int64_t ld_geno, void* stream) {
  int rc = check_matrix(d_hap, n_hap, n_sites, ld_hap);
  if (rc) return rc;
  if (n_hap % 2) return fail(PGLD_EINVAL, "need an even number of haplotypes");
  const int64_t n_ind = n_hap / 2;
  if (n_ind == 0 || n_sites == 0) return PGLD_OK;
  rc = check_matrix(d_geno, n_ind, n_sites, ld_geno);
  if (rc) return rc;
  const unsigned gy = (unsigned)(n_ind < 4096 ? n_ind : 4096);
  hipLaunchKernelGGL(hap_to_geno_kernel, dim3((unsigned)((n_sites + 255) / 256), gy),
                     dim3(256), 0, as_stream(stream), d_hap, ld_hap, n_ind, n_sites, d_geno,
                     ld_geno);
  PGLD_LAUNCH_CHECK();
  return PGLD_OK;
}

extern "C" int pgld_geno_to_hap(const int8_t* d_geno, int64_t n_ind, int64_t n_sites,
                                int64_t ld_geno, int8_t* d_hap, int64_t ld_hap, void* stream) {
  int rc = check_matrix(d_geno, n_ind, n_sites, ld_geno);
  if (rc) return rc;
  if (n_ind == 0 || n_sites == 0) return PGLD_OK;
  rc = check_matrix(d_hap, 2 * n_ind, n_sites, ld_hap);
  if (rc) return rc;
  const unsigned gy = (unsigned)(n_ind < 4096 ? n_ind : 4096);
  hipLaunchKernelGGL(geno_to_hap_kernel, dim3((unsigned)((n_sites + 255) / 256), gy),
                     dim3(256), 0, as_stream(stream), d_geno, ld_geno, n_ind, n_sites, d_hap,
                     ld_hap);
  PGLD_LAUNCH_CHECK();
  return PGLD_OK;
}
