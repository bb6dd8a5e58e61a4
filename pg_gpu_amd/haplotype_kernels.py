"""Batch dispatch of the haplotype estimators (drop-in for
pg_gpu/haplotype_kernels.py:335-640).

``pops`` are ``ld_pipeline.PopData`` objects of one pair batch; every call
list is evaluated by libpgld (pgld_hap_calls: one launch per 64 calls, each
thread one pair) and the full moments statistic set by pgld_hap_stat_set.
Results are float64 CUDA tensors.
"""
from __future__ import annotations

import ctypes

from . import _device as dv
from . import _lib

_KIND = {"dd": 0, "dz": 1, "pi2": 2}


def _flat_counts(pops, width):
    t = dv.torch()
    if width == 4:
        cols = [x for p in pops for x in (p.c1, p.c2, p.c3, p.c4)]
    else:
        cols = [p.counts[:, m] for p in pops for m in range(9)]
    c = t.stack(cols, dim=1).to(t.float64).contiguous()
    nv = t.stack([p.n.to(t.float64) for p in pops], dim=1).contiguous()
    return c, nv


def _run_calls(fn, c, nv, P, calls):
    """calls: [(kind, pops tuple)] -> list of float64 [N] tensors."""
    t = dv.torch()
    N = c.shape[0]
    res = []
    for s0 in range(0, len(calls), 64):
        chunk = calls[s0:s0 + 64]
        out = t.empty((N, len(chunk)), dtype=t.float64, device=c.device)
        flat = []
        for kind, pp in chunk:
            flat += [_KIND[kind]] + list(pp) + [0] * (4 - len(pp))
        arr = (ctypes.c_int32 * len(flat))(*flat)
        _lib.check(fn(_lib.ptr(c), c.stride(0), P, _lib.ptr(nv), N, arr, len(chunk),
                      _lib.ptr(out), _lib.stream_handle()))
        res += [out[:, q] for q in range(len(chunk))]
    return res


def compute_all_dd_hap(pops, dd_calls):
    c, nv = _flat_counts(pops, 4)
    return _run_calls(_lib.lib().pgld_hap_calls, c, nv, len(pops),
                      [("dd", tuple(x)) for x in dd_calls])


def compute_all_dz_hap(pops, dz_calls):
    c, nv = _flat_counts(pops, 4)
    return _run_calls(_lib.lib().pgld_hap_calls, c, nv, len(pops),
                      [("dz", tuple(x)) for x in dz_calls])


def compute_all_pi2_hap(pops, pi2_calls):
    c, nv = _flat_counts(pops, 4)
    return _run_calls(_lib.lib().pgld_hap_calls, c, nv, len(pops),
                      [("pi2", tuple(x)) for x in pi2_calls])


def weighted_specs(raw, stat_specs, N, device):
    """Assemble (N, n_stats) from raw call values and weighted specs."""
    t = dv.torch()
    out = t.zeros((N, len(stat_specs)), dtype=t.float64, device=device)
    for s, (_, calls) in enumerate(stat_specs):
        if len(calls) == 1 and calls[0][0] == 1.0:
            out[:, s] = raw[(calls[0][1], tuple(calls[0][2]))]
        else:
            out[:, s] = sum(w * raw[(st, tuple(pp))] for w, st, pp in calls)
    return out


def compute_multi_pop_statistics_batch_hap(counts_per_pop, n_valid_per_pop,
                                           ld_statistics_module=None, stat_specs=None):
    """(N, n_stats) for the given specs (default: the moments set, evaluated
    by one fused launch)."""
    from .ld_pipeline import PopData, _stat_set, generate_stat_specs

    P = len(counts_per_pop)
    if stat_specs is None or list(stat_specs) == generate_stat_specs(P):
        return _stat_set(counts_per_pop, n_valid_per_pop)
    pops = [PopData(c, nv) for c, nv in zip(counts_per_pop, n_valid_per_pop)]
    c, nv = _flat_counts(pops, 4)
    uniq = sorted({(st, tuple(pp)) for _, calls in stat_specs for _, st, pp in calls})
    vals = _run_calls(_lib.lib().pgld_hap_calls, c, nv, P, uniq)
    return weighted_specs(dict(zip(uniq, vals)), stat_specs, c.shape[0], c.device)
