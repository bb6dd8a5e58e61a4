// Genotype instantiations of the fused binned and window kernels.  One
// population (the per-wave binned form and the windows) keeps its nine
// moment products in the MFMA accumulators; otherwise the counts are packed
// per slot: nine 10-bit fields in three words for populations of at most
// 1,023 individuals, 16-bit fields in five words above.
#include "fused_kernels.h"

namespace pgld {

int launch_binned_geno(const PanelLayout& L, const pgld_binned_args* args, hipStream_t s) {
  // one population: the per-wave form reads the moment accumulators directly
  // (any population size); the producer/consumer form hands packed words over
  if (L.n_pops == 1 && args->form != 1)
    return args->with_r2 ? launch_fused<GenoRawTraits<true>>(L, args, s)
                         : launch_fused<GenoRawTraits<false>>(L, args, s);
  if (args->with_r2)  // one population
    return L.pop_size[0] <= 1023 ? launch_fused<GenoTraits<1, 3, true>>(L, args, s)
                                 : launch_fused<GenoTraits<1, 5, true>>(L, args, s);
  bool narrow = true;
  for (int p = 0; p < L.n_pops; ++p) narrow = narrow && L.pop_size[p] <= 1023;
  switch (L.n_pops) {
    case 1:
      return narrow ? launch_fused<GenoTraits<1, 3>>(L, args, s)
                    : launch_fused<GenoTraits<1, 5>>(L, args, s);
    case 2:
      return narrow ? launch_fused<GenoTraits<2, 3>>(L, args, s)
                    : launch_fused<GenoTraits<2, 5>>(L, args, s);
    case 3:
      return narrow ? launch_fused<GenoTraits<3, 3>>(L, args, s)
                    : launch_fused<GenoTraits<3, 5>>(L, args, s);
    case 4:
      return narrow ? launch_fused<GenoTraits<4, 3>>(L, args, s)
                    : launch_fused<GenoTraits<4, 5>>(L, args, s);
    default: break;
  }
  return fail(PGLD_EINVAL, "1-4 populations supported");
}

int launch_windows_geno(const PanelLayout& L, const pgld_window_args* args, hipStream_t s) {
  // the moment accumulators read directly, as in the per-wave binned kernel
  return launch_windows<GenoRawTraits<true>>(L, args, s);
}

}  // namespace pgld
