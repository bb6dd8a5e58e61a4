"""ctypes binding of libpgld.so (the C ABI declared in include/pgld.h).

The library must have been built (``python pg_gpu_amd/csrc/build.py`` or
``__graft_entry__.build()``); there is no fallback implementation.  Errors
returned by the library are raised as the Python exception the reference
raises for the same condition (ValueError for bad arguments).
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_int32, c_int64, c_void_p
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libpgld.so"

PGLD_OK, PGLD_EINVAL, PGLD_EDEVICE, PGLD_EUNSUPPORTED = 0, 1, 2, 3
MODE_HAP, MODE_GENO = 0, 1
PAIR_SIGMA_D2, PAIR_R2_NAIVE, PAIR_R2_DOSAGE, PAIR_R2 = 0, 1, 2, 3
FILTER_HAP_BIALLELIC, FILTER_GENO_BIALLELIC = 0, 1
MAX_POPS = 4
MAX_BINS = 64


class PanelDesc(ctypes.Structure):
    _fields_ = [
        ("mode", c_int32),
        ("n_pops", c_int32),
        ("n_sites", c_int64),
        ("pop_size", c_int32 * MAX_POPS),
        ("has_missing", c_int32 * MAX_POPS),
    ]


class BinnedArgs(ctypes.Structure):
    _fields_ = [
        ("d_panels", c_void_p),
        ("d_site_stats", c_void_p),
        ("d_mat", c_void_p),
        ("ld", c_int64),
        ("d_rows", c_void_p),
        ("row_offset", c_int32 * (MAX_POPS + 1)),
        ("d_pos", c_void_p),
        ("d_coord", c_void_p),
        ("coord_is_pos", c_int32),
        ("max_dist", c_double),
        ("d_edges", c_void_p),
        ("n_bins", c_int32),
        ("d_rtab", c_void_p),
        ("d_tiles", c_void_p),
        ("n_tiles", c_int64),
        ("d_out", c_void_p),
        ("grid", c_int32),
        ("n_lo", c_int32 * MAX_POPS),
        ("with_r2", c_int32),
        ("d_sched", c_void_p),
        ("form", c_int32),
    ]


class WindowArgs(ctypes.Structure):
    _fields_ = [
        ("d_panels", c_void_p),
        ("d_site_stats", c_void_p),
        ("d_mat", c_void_p),
        ("ld", c_int64),
        ("d_rows", c_void_p),
        ("d_rtab", c_void_p),
        ("d_windows", c_void_p),
        ("n_windows", c_int64),
        ("d_work", c_void_p),
        ("n_work", c_int64),
        ("d_out", c_void_p),
        ("grid", c_int32),
        ("n_lo", c_int32 * MAX_POPS),
        ("d_sched", c_void_p),
        ("tile", c_int32),
        ("group", c_int32),
    ]


SCHED_WORDS = 256  # PGLD_SCHED_WORDS

_SIGS = {
    "pgld_version": (c_int32, []),
    "pgld_last_error": (c_char_p, []),
    "pgld_panel_bytes": (c_int64, [POINTER(PanelDesc)]),
    "pgld_scan_sites": (c_int32, [c_void_p, c_int64, c_int64, c_int64, c_void_p,
                                  POINTER(c_int32), POINTER(PanelDesc), c_void_p, c_void_p,
                                  c_void_p]),
    "pgld_pack": (c_int32, [c_void_p, c_int64, c_int64, c_int64, c_void_p, POINTER(c_int32),
                            POINTER(PanelDesc), c_void_p, c_void_p]),
    "pgld_n_stats": (c_int32, [c_int32, c_int32]),
    "pgld_binned_fused": (c_int32, [POINTER(PanelDesc), POINTER(BinnedArgs), c_void_p]),
    "pgld_binned_plain": (c_int32, [POINTER(PanelDesc), POINTER(BinnedArgs), c_void_p]),
    "pgld_binned_fused_noalloc": (c_int32, [POINTER(PanelDesc), POINTER(BinnedArgs), c_void_p]),
    "pgld_count_band": (c_int32, [POINTER(PanelDesc), POINTER(BinnedArgs), c_void_p, c_void_p]),
    "pgld_windows_fused": (c_int32, [POINTER(PanelDesc), POINTER(WindowArgs), c_void_p]),
    "pgld_windows_plain": (c_int32, [POINTER(PanelDesc), POINTER(WindowArgs), c_void_p]),
    "pgld_count_block": (c_int32, [POINTER(PanelDesc), c_void_p, c_void_p, c_int32, c_int64,
                                   c_int64, c_int64, c_int64, c_void_p, c_void_p]),
    "pgld_pair_bounds": (c_int32, [c_void_p, c_int64, c_double, c_void_p, c_void_p]),
    "pgld_band_tiles": (c_int32, [c_void_p, c_int64, c_double, c_int32, c_void_p, c_int64,
                                  POINTER(c_int64), POINTER(c_int64)]),
    "pgld_emit_pairs": (c_int32, [c_void_p, c_int64, c_int64, c_int64, c_void_p, c_void_p,
                                  c_void_p]),
    "pgld_count_pairs": (c_int32, [POINTER(PanelDesc), c_void_p, c_void_p, c_int32, c_void_p,
                                   c_void_p, c_int64, c_void_p, c_void_p, c_void_p]),
    "pgld_hap_calls": (c_int32, [c_void_p, c_int64, c_int32, c_void_p, c_int64,
                                 POINTER(c_int32), c_int32, c_void_p, c_void_p]),
    "pgld_hap_stat_set": (c_int32, [c_void_p, c_int64, c_int32, c_void_p, c_int64, c_void_p,
                                    c_void_p]),
    "pgld_geno_calls": (c_int32, [c_void_p, c_int64, c_int32, c_void_p, c_int64,
                                  POINTER(c_int32), c_int32, c_void_p, c_void_p]),
    "pgld_geno_stat_set": (c_int32, [c_void_p, c_int64, c_int32, c_void_p, c_int64, c_void_p,
                                     c_void_p]),
    "pgld_sigma_d2_geno": (c_int32, [c_void_p, c_int64, c_int64, c_void_p, c_void_p]),
    "pgld_pair_block": (c_int32, [POINTER(PanelDesc), c_void_p, c_void_p, c_int32, c_int32,
                                  c_int64, c_int64, c_int64, c_int64, c_void_p, c_void_p]),
    "pgld_r_family": (c_int32, [c_void_p, c_int64, c_void_p, c_int64, c_int32, c_void_p,
                                c_void_p]),
    "pgld_heterozygosity": (c_int32, [POINTER(PanelDesc), c_void_p, c_void_p, c_void_p,
                                      c_void_p]),
    "pgld_biallelic_mask": (c_int32, [c_void_p, c_int64, c_int64, c_int64, c_void_p, c_int32,
                                      c_int32, c_void_p, c_void_p]),
    "pgld_hap_to_geno": (c_int32, [c_void_p, c_int64, c_int64, c_int64, c_void_p, c_int64,
                                   c_void_p]),
    "pgld_geno_to_hap": (c_int32, [c_void_p, c_int64, c_int64, c_int64, c_void_p, c_int64,
                                   c_void_p]),
}

_lib = None


class LibraryNotBuilt(RuntimeError):
    pass


def lib():
    """Load libpgld.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        path = os.environ.get("PGLD_LIBRARY", str(LIB_PATH))
        if not Path(path).exists():
            raise LibraryNotBuilt(
                f"{path} not found: build the HIP library first "
                "(python pg_gpu_amd/csrc/build.py)")
        L = ctypes.CDLL(path)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIGS)


def check(rc):
    """Map a pgld status code to the reference's exception types."""
    if rc == PGLD_OK:
        return
    msg = (lib().pgld_last_error() or b"").decode(errors="replace")
    if rc == PGLD_EINVAL:
        raise ValueError(msg)
    if rc == PGLD_EUNSUPPORTED:
        raise NotImplementedError(msg)
    raise RuntimeError(msg)


def ptr(t):
    """Device pointer of a torch tensor (None -> NULL)."""
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def stream_handle(stream=None):
    import torch

    s = torch.cuda.current_stream() if stream is None else stream
    return ctypes.c_void_p(s.cuda_stream)


def int32_array(values):
    arr = (c_int32 * len(values))(*[int(v) for v in values])
    return arr


byref = ctypes.byref
