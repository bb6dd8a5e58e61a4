// Heterozygosity statistics H_i_j summed over sites, from the per-site
// tallies of pgld_scan_sites (moments_ld.py:_compute_heterozygosity, :346-388).
//
// Per population p and site s: alt = number of alt alleles, n = allele sample
// size.  Haplotypes: alt = count of value 1 and n = population size (missing
// calls count as reference alleles, as in the reference).  Genotypes:
// alt = c1 + 2 c2 and n = 2 (K - missing).
//   H_i_i = sum_s 2 ref_i alt_i / (n_i (n_i - 1)),
//   H_i_j = sum_s (ref_i alt_j + alt_i ref_j) / (n_i n_j).
// Each block sums its sites in a fixed-order tree, then adds the block sum
// as fixed-point limbs (reproducible); non-finite terms are counted.
#include "pgld_internal.h"

namespace pgld {

__global__ void __launch_bounds__(256) het_kernel(const int32_t* __restrict__ stats, int P,
                                                  int64_t V, int mode, int4 K,
                                                  unsigned long long* __restrict__ out,
                                                  int* __restrict__ nonfinite) {
  __shared__ double red[256];
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double alt[PGLD_MAX_POPS], n[PGLD_MAX_POPS];
  const int Ks[4] = {K.x, K.y, K.z, K.w};
  for (int p = 0; p < P; ++p) {
    if (s < V) {
      const int32_t* st = stats + (int64_t)p * 3 * V;
      if (mode == PGLD_MODE_HAP) {
        alt[p] = st[s];
        n[p] = Ks[p];
      } else {
        alt[p] = st[s] + 2.0 * st[V + s];
        n[p] = 2.0 * (Ks[p] - st[2 * V + s]);
      }
    } else {
      alt[p] = 0.0;
      n[p] = 0.0;
    }
  }
  int e = 0;
  for (int i = 0; i < P; ++i) {
    for (int j = i; j < P; ++j, ++e) {
      double v = 0.0;
      if (s < V) {
        if (i == j)
          v = 2.0 * (n[i] - alt[i]) * alt[i] / (n[i] * (n[i] - 1.0));
        else
          v = ((n[i] - alt[i]) * alt[j] + alt[i] * (n[j] - alt[j])) / (n[i] * n[j]);
      }
      if (!isfinite(v)) {
        atomicAdd(nonfinite + e, 1);
        v = 0.0;
      }
      red[threadIdx.x] = v;
      __syncthreads();
      for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
      }
      if (threadIdx.x == 0) {
        long long x2, x1, x0;
        pgld_to_limbs(red[0], x2, x1, x0);
        atomicAdd(out + 3 * e + 0, (unsigned long long)x0);
        atomicAdd(out + 3 * e + 1, (unsigned long long)x1);
        atomicAdd(out + 3 * e + 2, (unsigned long long)x2);
      }
      __syncthreads();
    }
  }
}

}  // namespace pgld

using namespace pgld;

extern "C" int pgld_heterozygosity(const pgld_panel_desc* desc, const int32_t* d_site_stats,
                                   long long* d_out, int32_t* d_nonfinite, void* stream) {
  PanelLayout L;
  int rc = make_layout(desc, &L);
  if (rc) return rc;
  if (L.n_sites == 0) return PGLD_OK;
  const int4 K = make_int4(L.pop_size[0], L.pop_size[1], L.pop_size[2], L.pop_size[3]);
  hipLaunchKernelGGL(het_kernel, dim3((unsigned)((L.n_sites + 255) / 256)), dim3(256), 0,
                     as_stream(stream), d_site_stats, desc->n_pops, L.n_sites, desc->mode, K,
                     reinterpret_cast<unsigned long long*>(d_out), d_nonfinite);
  PGLD_LAUNCH_CHECK();
  return PGLD_OK;
}
