"""Batch dispatch of the genotype estimators (drop-in for
pg_gpu/genotype_kernels.py:561-936).

``pops`` are ``PopDataGeno`` objects of one pair batch (the nine counts in
the reference order n00 .. n22 and n_valid); call lists are evaluated by
libpgld (pgld_geno_calls) and the moments statistic set by
pgld_geno_stat_set.  Results are float64 CUDA tensors.
"""
from __future__ import annotations

from . import _device as dv
from . import _lib
from .haplotype_kernels import _flat_counts, _run_calls, weighted_specs


class PopDataGeno:
    """Nine-way counts (N, 9), n00 .. n22, and valid individuals per pair
    (the inputs of _PopDataGeno, pg_gpu/genotype_kernels.py:412-492)."""
    __slots__ = ("counts", "n")

    def __init__(self, counts, n_valid=None):
        t = dv.torch()
        self.counts = dv.as_tensor(counts, t.float64)
        if self.counts.dim() != 2 or self.counts.shape[1] != 9:
            raise ValueError("genotype counts must have shape (N, 9)")
        self.n = (dv.as_tensor(n_valid, t.float64, self.counts.device) if n_valid is not None
                  else self.counts.sum(dim=1))


def _pops(pops):
    return [p if isinstance(p, PopDataGeno) else PopDataGeno(*p) for p in pops]


def compute_all_dd_geno(pops, dd_calls):
    c, nv = _flat_counts(_pops(pops), 9)
    return _run_calls(_lib.lib().pgld_geno_calls, c, nv, len(pops),
                      [("dd", tuple(x)) for x in dd_calls])


def compute_all_dz_geno(pops, dz_calls):
    c, nv = _flat_counts(_pops(pops), 9)
    return _run_calls(_lib.lib().pgld_geno_calls, c, nv, len(pops),
                      [("dz", tuple(x)) for x in dz_calls])


def compute_all_pi2_geno(pops, pi2_calls):
    c, nv = _flat_counts(_pops(pops), 9)
    return _run_calls(_lib.lib().pgld_geno_calls, c, nv, len(pops),
                      [("pi2", tuple(x)) for x in pi2_calls])


def _geno_stat_set(pops):
    t = dv.torch()
    c, nv = _flat_counts(pops, 9)
    P = len(pops)
    N = c.shape[0]
    ns = _lib.lib().pgld_n_stats(_lib.MODE_GENO, P)
    out = t.empty((N, ns), dtype=t.float64, device=c.device)
    _lib.check(_lib.lib().pgld_geno_stat_set(_lib.ptr(c), c.stride(0), P, _lib.ptr(nv), N,
                                             _lib.ptr(out), _lib.stream_handle()))
    return out


def compute_multi_pop_statistics_batch_geno(counts_per_pop, n_valid_per_pop,
                                            ld_statistics_module=None, stat_specs=None):
    """(N, n_stats) genotype statistics for the given specs (default: the
    moments set in one fused launch)."""
    from .ld_pipeline import generate_stat_specs

    P = len(counts_per_pop)
    pops = [PopDataGeno(c, nv) for c, nv in zip(counts_per_pop, n_valid_per_pop)]
    if stat_specs is None or list(stat_specs) == generate_stat_specs(P):
        return _geno_stat_set(pops)
    c, nv = _flat_counts(pops, 9)
    uniq = sorted({(st, tuple(pp)) for _, calls in stat_specs for _, st, pp in calls})
    vals = _run_calls(_lib.lib().pgld_geno_calls, c, nv, P, uniq)
    return weighted_specs(dict(zip(uniq, vals)), stat_specs, c.shape[0], c.device)
