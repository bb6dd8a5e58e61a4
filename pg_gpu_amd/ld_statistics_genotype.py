"""Genotype LD entry points of pg_gpu/ld_statistics_genotype.py that are not
covered by the batch dispatch in genotype_kernels."""
from __future__ import annotations

from . import _device as dv
from . import _lib


def sigma_d2_geno(matrix, idx_i=None, idx_j=None):
    """Per-pair sigma_D^2 = DD / pi2 of diploid dosages
    (pg_gpu/ld_statistics_genotype.py:1518-1598).  NaN where pi2 <= 0 or fewer
    than 4 individuals; missing calls raise ValueError."""
    from .genotype_matrix import GenotypeMatrix
    from .haplotype_matrix import HaplotypeMatrix
    from .ld_pipeline import compute_genotype_counts_for_pairs

    t = dv.torch()
    if isinstance(matrix, GenotypeMatrix):
        gm = matrix
    elif isinstance(matrix, HaplotypeMatrix):
        gm = GenotypeMatrix.from_haplotype_matrix(matrix)
    else:
        raise TypeError(f"matrix must be a GenotypeMatrix or HaplotypeMatrix, "
                        f"got {type(matrix).__name__}")
    if gm.device != "GPU":
        gm.transfer_to_gpu()
    g = gm.genotypes
    if bool((g < 0).any()):
        raise ValueError("sigma_d2_geno does not support missing values (-1); "
                         "filter or impute before calling")
    m = g.shape[1]
    if idx_i is None and idx_j is None:
        ii, jj = t.triu_indices(m, m, offset=1, device=g.device)
        idx_i, idx_j = ii.to(t.int32), jj.to(t.int32)
    elif idx_i is None or idx_j is None:
        raise ValueError("idx_i and idx_j must be provided together")
    else:
        idx_i = dv.as_tensor(idx_i, t.int32, g.device)
        idx_j = dv.as_tensor(idx_j, t.int32, g.device)
        if idx_i.shape != idx_j.shape:
            raise ValueError("idx_i and idx_j must have the same shape")
    counts, _ = compute_genotype_counts_for_pairs(g, idx_i, idx_j)
    c = counts.to(t.float64).contiguous()
    N = c.shape[0]
    out = t.empty(N, dtype=t.float64, device=g.device)
    _lib.check(_lib.lib().pgld_sigma_d2_geno(_lib.ptr(c), c.stride(0), N, _lib.ptr(out),
                                             _lib.stream_handle()))
    return out
