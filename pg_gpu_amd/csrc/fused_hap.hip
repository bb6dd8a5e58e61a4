// Haplotype instantiations of the fused binned, counting-stage and window
// kernels (one translation unit per mode so the two compile in parallel).
#include "fused_kernels.h"

namespace pgld {

int launch_binned_hap(const PanelLayout& L, const pgld_binned_args* args, hipStream_t s) {
  // one population without missing calls: constant reciprocals, counts kept
  // as they come (HapNMRawTraits)
  const bool nm = L.n_pops == 1 && !L.missing[0];
  if (args->with_r2)  // one population
    return nm ? launch_fused<HapNMRawTraits<true>>(L, args, s)
              : launch_fused<HapTraits<1, true>>(L, args, s);
  switch (L.n_pops) {
    case 1: return nm ? launch_fused<HapNMRawTraits<false>>(L, args, s)
                      : launch_fused<HapTraits<1>>(L, args, s);
    case 2: return launch_fused<HapTraits<2>>(L, args, s);
    case 3: return launch_fused<HapTraits<3>>(L, args, s);
    case 4: return launch_fused<HapTraits<4>>(L, args, s);
    default: break;
  }
  return fail(PGLD_EINVAL, "1-4 populations supported");
}

int launch_count_band_hap(const PanelLayout& L, const pgld_binned_args* args,
                          unsigned long long* sums, hipStream_t s) {
  switch (L.n_pops) {
    case 1: return launch_count_band<HapTraits<1>>(L, args, sums, s);
    case 2: return launch_count_band<HapTraits<2>>(L, args, sums, s);
    case 3: return launch_count_band<HapTraits<3>>(L, args, sums, s);
    case 4: return launch_count_band<HapTraits<4>>(L, args, sums, s);
    default: break;
  }
  return fail(PGLD_EINVAL, "1-4 populations supported");
}

int launch_windows_hap(const PanelLayout& L, const pgld_window_args* args, hipStream_t s) {
  return L.missing[0] ? launch_windows<HapTraits<1, true>>(L, args, s)
                      : launch_windows<HapNMRawTraits<true>>(L, args, s);
}

}  // namespace pgld
