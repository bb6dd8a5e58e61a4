// Host-side helpers shared by the pgld translation units.
#pragma once
#include <string>
#include "../../include/pgld.h"
#include "pgld_common.h"

namespace pgld {

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define PGLD_HIP_CHECK(expr)                                                        \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess)                                                           \
      return ::pgld::fail(PGLD_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define PGLD_LAUNCH_CHECK() PGLD_HIP_CHECK(hipGetLastError())

// Device-visible layout of the packed panels.
struct PanelLayout {
  int32_t n_pops;
  int32_t mode;
  int32_t n_feat[PGLD_MAX_POPS];      // features per pop (1..3)
  int32_t ksteps[PGLD_MAX_POPS];      // 64-sample steps per pop
  int32_t pop_size[PGLD_MAX_POPS];
  int32_t missing[PGLD_MAX_POPS];
  int64_t pop_off[PGLD_MAX_POPS];     // byte offset of the pop inside a site block
  int64_t block_bytes;                // bytes per 32-site block
  int64_t n_blocks;
  int64_t n_sites;
};

int make_layout(const pgld_panel_desc* d, PanelLayout* L);

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Launchers of the fused kernels per mode (fused_hap.hip, fused_geno.hip);
// the arguments are validated by the entry points in fused.hip.
int launch_binned_hap(const PanelLayout& L, const pgld_binned_args* args, hipStream_t s);
int launch_binned_geno(const PanelLayout& L, const pgld_binned_args* args, hipStream_t s);
int launch_count_band_hap(const PanelLayout& L, const pgld_binned_args* args,
                          unsigned long long* sums, hipStream_t s);
int launch_windows_hap(const PanelLayout& L, const pgld_window_args* args, hipStream_t s);
int launch_windows_geno(const PanelLayout& L, const pgld_window_args* args, hipStream_t s);

}  // namespace pgld
