"""Once-per-dataset passes of the containers on the GPU: the biallelic site
filters and the haplotype <-> genotype conversions (libpgld's
pgld_biallelic_mask, pgld_hap_to_geno, pgld_geno_to_hap).

Inputs may be NumPy arrays or torch tensors; the work runs on the GPU and
the result comes back on the input's device, as the reference keeps the
container's device (NumPy or CuPy).
"""
from __future__ import annotations

import numpy as np

from . import _device as dv
from . import _lib


def _gpu_int8(x):
    """(contiguous-row int8 CUDA tensor, was_on_gpu)."""
    t = dv.torch()
    on_gpu = dv.is_gpu_array(x)
    m = dv.as_tensor(x, t.int8)
    if m.dim() != 2:
        raise ValueError("expected a 2-D matrix")
    if m.stride(1) != 1:
        m = m.contiguous()
    return m, on_gpu


def biallelic_keep(mat, rule, rows=None):
    """Boolean CUDA tensor [n_sites]: the sites passing ``rule``
    (_lib.FILTER_HAP_BIALLELIC / FILTER_GENO_BIALLELIC) over ``rows``
    (default all rows)."""
    t = dv.torch()
    m, _ = _gpu_int8(mat)
    n_rows, V = m.shape
    keep = t.zeros(V, dtype=t.uint8, device=m.device)
    d_rows = None
    n_sel = n_rows
    if rows is not None:
        r = np.asarray(rows, np.int64).ravel()
        if r.size and (r.min() < 0 or r.max() >= n_rows):
            raise ValueError("row index out of range")
        d_rows = t.as_tensor(r.astype(np.int32), device=m.device)
        n_sel = int(r.size)
    _lib.check(_lib.lib().pgld_biallelic_mask(_lib.ptr(m), n_rows, V, m.stride(0) if n_rows else V,
                                              _lib.ptr(d_rows), n_sel, int(rule), _lib.ptr(keep),
                                              _lib.stream_handle()))
    return keep.bool()


def keep_index(keep):
    """int64 CUDA tensor of the kept column indices (ascending)."""
    return dv.torch().nonzero(keep).reshape(-1)


def take_columns(x, idx):
    """x[:, idx] (or x[idx] for 1-D) on x's device; ``idx`` a CUDA index tensor."""
    t = dv.torch()
    if isinstance(x, t.Tensor) and x.is_cuda:
        return x.index_select(x.dim() - 1, idx)
    i = idx.cpu().numpy()
    return np.asarray(x)[..., i]


def hap_to_geno(hap):
    """Genotypes of consecutive haplotype pairs (device of ``hap`` kept)."""
    t = dv.torch()
    m, on_gpu = _gpu_int8(hap)
    n_hap, V = m.shape
    if n_hap % 2:
        raise ValueError(f"Need even number of haplotypes for diploid conversion, got {n_hap}")
    out = t.empty((n_hap // 2, V), dtype=t.int8, device=m.device)
    _lib.check(_lib.lib().pgld_hap_to_geno(_lib.ptr(m), n_hap, V, m.stride(0), _lib.ptr(out), V,
                                           _lib.stream_handle()))
    return out if on_gpu else out.cpu().numpy()


def geno_to_hap(geno):
    """Two haplotypes per individual (device of ``geno`` kept)."""
    t = dv.torch()
    m, on_gpu = _gpu_int8(geno)
    n_ind, V = m.shape
    out = t.empty((2 * n_ind, V), dtype=t.int8, device=m.device)
    _lib.check(_lib.lib().pgld_geno_to_hap(_lib.ptr(m), n_ind, V, m.stride(0), _lib.ptr(out), V,
                                           _lib.stream_handle()))
    return out if on_gpu else out.cpu().numpy()
