"""Per-pair LD statistics from haplotype counts (drop-in for the estimator part
of pg_gpu/ld_statistics.py).

Counts are (N, 4) arrays [n11, n10, n01, n00] for one population or (N, 4P)
concatenations for P populations; ``n_valid`` is None, a tuple of per-pop
(N,) arrays (entries may be None), an (N, P) array, or one (N,) array, with
the reference's meaning (pg_gpu/ld_statistics.py:1082-1113).  Inputs may be
torch tensors or NumPy arrays; results are float64 CUDA tensors computed by
libpgld (pgld_hap_calls / pgld_r_family).
"""
from __future__ import annotations

import ctypes

from . import _device as dv
from . import _lib

_KIND = {"dd": 0, "dz": 1, "pi2": 2}


def _prep(counts, n_valid, n_pops):
    """float64 counts [N, 4P] and an (N, P) n_valid matrix (-1 = use count sum)."""
    t = dv.torch()
    c = dv.as_tensor(counts, t.float64)
    if c.dim() != 2 or c.shape[1] < 4 * n_pops:
        raise ValueError(f"counts must have shape (N, {4 * n_pops}) or wider")
    N = c.shape[0]
    nv = None
    if n_valid is not None:
        nv = t.full((N, n_pops), -1.0, dtype=t.float64, device=c.device)
        if isinstance(n_valid, tuple):
            for p in range(min(n_pops, len(n_valid))):
                if n_valid[p] is not None:
                    nv[:, p] = dv.as_tensor(n_valid[p], t.float64, c.device).reshape(-1)
        else:
            arr = dv.as_tensor(n_valid, t.float64, c.device)
            if arr.dim() == 2:
                nv[:, :min(n_pops, arr.shape[1])] = arr[:, :n_pops]
            else:
                nv[:, :] = arr.reshape(-1, 1)   # one n_valid for every population
    return c, nv


def _calls(counts, n_valid, n_pops, calls):
    t = dv.torch()
    c, nv = _prep(counts, n_valid, n_pops)
    N = c.shape[0]
    out = t.empty((N, len(calls)), dtype=t.float64, device=c.device)
    flat = []
    for kind, pops in calls:
        pp = list(pops) + [0] * (4 - len(pops))
        flat += [_KIND[kind]] + pp
    arr = (ctypes.c_int32 * len(flat))(*flat)
    _lib.check(_lib.lib().pgld_hap_calls(_lib.ptr(c), c.stride(0), n_pops, _lib.ptr(nv), N, arr,
                                         len(calls), _lib.ptr(out), _lib.stream_handle()))
    return out


def _single(counts, n_valid, kind, pops):
    return _calls(counts, n_valid, 1, [(kind, pops)])[:, 0]


def _pop_nvalid(n_valid, p):
    """The n_valid entry of population p, reference rules (ld_statistics.py:58-63)."""
    if n_valid is None:
        return None
    if isinstance(n_valid, tuple):
        return n_valid[p] if p < len(n_valid) else None
    if getattr(n_valid, "ndim", 1) == 2:
        return n_valid[:, p]
    return n_valid


def dd(counts, populations=None, n_valid=None):
    """D^2 (pg_gpu/ld_statistics.py:13-65)."""
    if populations is None:
        if counts.shape[1] == 4:
            return _single(counts, n_valid, "dd", (0, 0))
        nv = n_valid[:, 0] if n_valid is not None and getattr(n_valid, "ndim", 1) == 2 else n_valid
        return _single(counts[:, :4], nv, "dd", (0, 0))
    p1, p2 = populations
    if p1 == p2:
        return _single(counts[:, 4 * p1:4 * p1 + 4], _pop_nvalid(n_valid, p1), "dd", (0, 0))
    P = max(p1, p2) + 1
    return _calls(counts, n_valid, P, [("dd", (p1, p2))])[:, 0]


def dz(counts, populations=None, n_valid=None):
    """Dz (pg_gpu/ld_statistics.py:68-97, :1172-1272)."""
    if populations is None:
        if counts.shape[1] == 4:
            return _single(counts, n_valid, "dz", (0, 0, 0))
        populations = (0, 0, 0)
    p1, p2, p3 = populations
    if p1 == p2 == p3:
        return _single(counts[:, 4 * p1:4 * p1 + 4], _pop_nvalid(n_valid, p1), "dz", (0, 0, 0))
    P = max(populations) + 1
    return _calls(counts, n_valid, P, [("dz", tuple(populations))])[:, 0]


def pi2(counts, populations=None, n_valid=None):
    """pi2 (pg_gpu/ld_statistics.py:100-129, :1275-1617)."""
    if populations is None:
        if counts.shape[1] == 4:
            return _single(counts, n_valid, "pi2", (0, 0, 0, 0))
        populations = (0, 0, 0, 0)
    i, j, k, l = populations
    if i == j == k == l:
        nv = n_valid[:, i] if n_valid is not None and getattr(n_valid, "ndim", 1) == 2 else n_valid
        return _single(counts[:, 4 * i:4 * i + 4], nv, "pi2", (0, 0, 0, 0))
    if i == j and k == l:
        # pi2(i,i,k,k) averages the two locus orientations (ld_statistics.py:1330-1341)
        P = max(i, k) + 1
        v = _calls(counts, n_valid, P, [("pi2", (i, i, k, k)), ("pi2", (k, k, i, i))])
        return 0.5 * (v[:, 0] + v[:, 1])
    P = max(populations) + 1
    return _calls(counts, n_valid, P, [("pi2", tuple(populations))])[:, 0]


def dd_within(counts, n_valid=None):
    return dd(counts, None, n_valid)


def dd_between(counts, pop1_idx, pop2_idx, n_valid=None):
    return dd(counts, (pop1_idx, pop2_idx), n_valid)


def _r_family(counts, n_valid, which):
    t = dv.torch()
    c = dv.as_tensor(counts, t.float64)
    if c.dim() != 2 or c.shape[1] < 4:
        raise ValueError("counts must have shape (N, 4)")
    nv = None if n_valid is None else dv.as_tensor(n_valid, t.float64, c.device).reshape(-1)
    out = t.empty(c.shape[0], dtype=t.float64, device=c.device)
    _lib.check(_lib.lib().pgld_r_family(_lib.ptr(c), c.stride(0), _lib.ptr(nv), c.shape[0], which,
                                        _lib.ptr(out), _lib.stream_handle()))
    return out


def r(counts, n_valid=None):
    """Pearson r from 4-way counts; NaN if monomorphic (ld_statistics.py:167-191)."""
    return _r_family(counts, n_valid, 0)


def r_squared(counts, n_valid=None):
    """r^2 from 4-way counts; NaN if monomorphic (ld_statistics.py:194-217)."""
    return _r_family(counts, n_valid, 1)


def d_prime(counts, n_valid=None):
    """Lewontin's D' (ld_statistics.py:220-246)."""
    return _r_family(counts, n_valid, 2)


def compute_ld_statistics(counts, statistics=("dd", "dz", "pi2"), populations=None, n_valid=None):
    """Several statistics in one call (ld_statistics.py:1030-1076)."""
    populations = populations or {}
    out = {}
    for s in statistics:
        if s == "dd":
            out["dd"] = dd(counts, populations.get("dd"), n_valid)
        elif s == "dz":
            out["dz"] = dz(counts, populations.get("dz"), n_valid)
        elif s == "pi2":
            out["pi2"] = pi2(counts, populations.get("pi2"), n_valid)
        elif s == "r":
            out["r"] = r(counts, n_valid)
        elif s == "r_squared":
            out["r_squared"] = r_squared(counts, n_valid)
        else:
            raise ValueError(f"Unknown statistic: {s}")
    return out


# Dense tile consumers (ld_statistics.py:534-1021), built on pgld_pair_block.
from .ld_tiles import omega, r2_matrix_diploid, zns  # noqa: E402,F401

