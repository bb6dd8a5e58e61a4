"""Pair generation, per-pair counting, batch statistics and moments stat specs
(drop-in for pg_gpu/ld_pipeline.py).

The device work goes through libpgld: pair bounds and emission
(pgld_pair_bounds / pgld_emit_pairs), counting from packed MFMA panels
(pgld_count_pairs) and the estimator batches (pgld_hap_stat_set).  The
binned hot path does not use these per-pair routines -- it runs the fused
kernel (see engine.LDPanels.binned) -- but they keep the reference's
per-pair API available.
"""
from __future__ import annotations

import weakref

import numpy as np

from . import _device as dv
from . import _lib
from .engine import LDPanels

# ---------------------------------------------------------------------------
# Memory estimation  (ld_pipeline.py:16-49)
# ---------------------------------------------------------------------------


def estimate_ld_chunk_size(n_haplotypes_per_pop, available_memory_bytes=None, num_pops=2):
    """Pairs per chunk from free device memory; same formula as the reference."""
    if available_memory_bytes is None:
        t = dv.torch()
        free, _ = t.cuda.mem_get_info()
        available_memory_bytes = int(free * 0.5)
    bytes_per_pair = (4 * n_haplotypes_per_pop * num_pops + 150 * num_pops) * 3
    return max(100_000, min(available_memory_bytes // bytes_per_pair, 10_000_000))


# ---------------------------------------------------------------------------
# Pair generation  (ld_pipeline.py:57-122)
# ---------------------------------------------------------------------------


def iter_pairs_within_distance(positions, max_dist, chunk_size):
    """Yield int32 CUDA tensors (idx_i, idx_j) of pairs i<j with
    pos[j] <= pos[i] + max_dist, in contiguous anchor blocks of about
    ``chunk_size`` pairs."""
    t = dv.torch()
    n = len(positions)
    if n < 2:
        return
    pos = dv.as_tensor(positions, t.float64)
    jmax = t.empty(n, dtype=t.int64, device=pos.device)
    h = _lib.stream_handle()
    _lib.check(_lib.lib().pgld_pair_bounds(_lib.ptr(pos), n, float(max_dist), _lib.ptr(jmax), h))
    per = np.maximum(0, jmax.cpu().numpy() - np.arange(n) - 1)
    cum = np.cumsum(per)
    total = int(cum[-1])
    if total == 0:
        return
    chunk_size = max(1, int(chunk_size))
    n_chunks = (total + chunk_size - 1) // chunk_size
    targets = np.minimum(np.arange(1, n_chunks + 1, dtype=np.int64) * chunk_size, total)
    v_his = np.minimum(np.searchsorted(cum, targets, side="left") + 1, n)
    v_lo = 0
    for v_hi in v_his:
        v_hi = int(v_hi)
        if v_hi <= v_lo:
            continue
        seg = per[v_lo:v_hi]
        m = int(seg.sum())
        if m:
            offs = np.concatenate([[0], np.cumsum(seg)]).astype(np.int64)
            d_offs = t.as_tensor(offs, device=pos.device)
            ii = t.empty(m, dtype=t.int32, device=pos.device)
            jj = t.empty(m, dtype=t.int32, device=pos.device)
            _lib.check(_lib.lib().pgld_emit_pairs(_lib.ptr(d_offs), v_lo, v_hi - v_lo, m,
                                                  _lib.ptr(ii), _lib.ptr(jj), h))
            yield ii, jj
        v_lo = v_hi


# ---------------------------------------------------------------------------
# Counting  (ld_pipeline.py:130-263)
# ---------------------------------------------------------------------------

_PANEL_CACHE = weakref.WeakKeyDictionary()
_PANELS_PER_MATRIX = 8   # populations kept packed per matrix (least recently used dropped)


def _panels_for(mat, pop_indices, mode):
    """Packed panels of ``mat`` restricted to one population, cached per
    (tensor, in-place version, rows): the reference's chunk loop calls the
    counting once per population and chunk (moments_ld.py:314-318), and each
    population is packed only once."""
    t = dv.torch()
    if not isinstance(mat, t.Tensor) or not mat.is_cuda:
        mat = dv.as_tensor(mat, t.int8)
    elif mat.dtype != t.int8:
        mat = mat.to(t.int8)
    rows = list(range(mat.shape[0])) if pop_indices is None else [int(x) for x in
                                                                   dv.to_numpy(pop_indices)]
    key = (mode, tuple(rows))
    per = _PANEL_CACHE.get(mat)
    if per is None or per[0] != mat._version:   # written in place: repack everything
        per = (mat._version, {})
        _PANEL_CACHE[mat] = per
    cache = per[1]
    if key in cache:
        cache[key] = cache.pop(key)   # most recently used last
        return cache[key]
    pan = LDPanels(mat, [rows], mode)
    cache[key] = pan
    while len(cache) > _PANELS_PER_MATRIX:
        cache.pop(next(iter(cache)))
    return pan


def _check_idx(idx, n):
    t = dv.torch()
    x = dv.as_tensor(idx, t.int32)
    if x.numel():
        lo, hi = t.aminmax(x)
        if int(lo) < 0 or int(hi) >= n:
            raise IndexError("pair index out of range")
    return x


def _count(mat, idx_i, idx_j, pop_indices, mode):
    t = dv.torch()
    pan = _panels_for(mat, pop_indices, mode)
    ii = _check_idx(idx_i, pan.V)
    jj = _check_idx(idx_j, pan.V)
    if ii.shape != jj.shape:
        raise ValueError("idx_i and idx_j must have the same shape")
    N = ii.numel()
    q = 4 if mode == "hap" else 9
    counts = t.empty((N, q), dtype=t.int32, device=pan.device)
    nvalid = t.empty(N, dtype=t.int32, device=pan.device)
    _lib.check(_lib.lib().pgld_count_pairs(_lib.byref(pan.desc), _lib.ptr(pan.panels),
                                           _lib.ptr(pan.site_stats), 0, _lib.ptr(ii),
                                           _lib.ptr(jj), N, _lib.ptr(counts), _lib.ptr(nvalid),
                                           _lib.stream_handle()))
    return counts, (nvalid if pan.has_missing[0] else None)


def compute_counts_for_pairs(haplotypes, idx_i, idx_j, pop_indices=None):
    """(int32 [N, 4] = n11, n10, n01, n00; int32 [N] n_valid or None)."""
    return _count(haplotypes, idx_i, idx_j, pop_indices, "hap")


def compute_genotype_counts_for_pairs(genotypes, idx_i, idx_j, pop_indices=None):
    """(int32 [N, 9] ordered n00, n01, ..., n22; int32 [N] n_valid or None)."""
    return _count(genotypes, idx_i, idx_j, pop_indices, "geno")


# ---------------------------------------------------------------------------
# Batch statistics  (ld_pipeline.py:271-349)
# ---------------------------------------------------------------------------


def _stat_set(counts_list, n_valid_list):
    t = dv.torch()
    P = len(counts_list)
    cs = [dv.as_tensor(c, t.float64) for c in counts_list]
    c = t.cat(cs, dim=1).contiguous()
    N = c.shape[0]
    nv = None
    if any(x is not None for x in n_valid_list):
        nv = t.full((N, P), -1.0, dtype=t.float64, device=c.device)
        for p, x in enumerate(n_valid_list):
            if x is not None:
                nv[:, p] = dv.as_tensor(x, t.float64, c.device).reshape(-1)
    ns = _lib.lib().pgld_n_stats(_lib.MODE_HAP, P)
    out = t.empty((N, ns), dtype=t.float64, device=c.device)
    _lib.check(_lib.lib().pgld_hap_stat_set(_lib.ptr(c), c.stride(0), P, _lib.ptr(nv), N,
                                            _lib.ptr(out), _lib.stream_handle()))
    return out


def compute_two_pop_statistics_batch(counts_pop1, counts_pop2, n_valid1, n_valid2,
                                     ld_statistics=None):
    """The 15 two-population statistics in ld_names(2) order, (N, 15)."""
    return _stat_set([counts_pop1, counts_pop2], [n_valid1, n_valid2])


def compute_single_pop_statistics_batch(counts, n_valid, ld_statistics=None):
    """(N, 3) columns DD, Dz, pi2."""
    return _stat_set([counts], [n_valid])


# ---------------------------------------------------------------------------
# Names and specs  (ld_pipeline.py:357-457)
# ---------------------------------------------------------------------------


def ld_names(num_pops):
    """LD statistic names in moments.LD.Util.ld_names order."""
    out = [f"DD_{a}_{b}" for a in range(num_pops) for b in range(a, num_pops)]
    out += [f"Dz_{a}_{b}_{c}" for a in range(num_pops) for b in range(num_pops)
            for c in range(b, num_pops)]
    for a in range(num_pops):
        for b in range(a, num_pops):
            for c in range(a, num_pops):
                for d in range(c, num_pops):
                    if c == a == d and b != a:
                        continue
                    if a == c and d < b:
                        continue
                    out.append(f"pi2_{a}_{b}_{c}_{d}")
    return out


def het_names(num_pops):
    return [f"H_{a}_{b}" for a in range(num_pops) for b in range(a, num_pops)]


def generate_stat_specs(num_pops):
    """(name, [(weight, type, pops), ...]) -- moments' _call_sgc averaging."""
    specs = []
    for name in ld_names(num_pops):
        kind, *pp = name.split("_")
        pp = tuple(int(x) for x in pp)
        if kind == "DD":
            specs.append((name, [(1.0, "dd", pp)]))
        elif kind == "Dz":
            a, b, c = pp
            specs.append((name, [(1.0, "dz", pp)] if b == c else
                          [(0.5, "dz", (a, b, c)), (0.5, "dz", (a, c, b))]))
        else:
            a, b, c, d = pp
            if a == b and c == d:
                calls = [(1.0, pp)] if a == c else [(0.5, (a, b, c, d)), (0.5, (c, d, a, b))]
            elif a == b:
                calls = [(0.25, x) for x in ((a, b, c, d), (a, b, d, c), (c, d, a, b), (d, c, a, b))]
            elif c == d:
                calls = [(0.25, x) for x in ((a, b, c, d), (b, a, c, d), (c, d, a, b), (c, d, b, a))]
            else:
                calls = [(0.125, x) for x in ((a, b, c, d), (a, b, d, c), (b, a, c, d),
                                              (b, a, d, c), (c, d, a, b), (d, c, a, b),
                                              (c, d, b, a), (d, c, b, a))]
            specs.append((name, [(w, "pi2", x) for w, x in calls]))
    return specs


class PopData:
    """Per-population float64 features of a pair batch (ld_pipeline.py:465-482)."""
    __slots__ = ("c1", "c2", "c3", "c4", "n", "D", "pA", "qA", "pB", "qB")

    def __init__(self, counts, n_valid):
        t = dv.torch()
        c = dv.as_tensor(counts, t.float64)
        self.c1, self.c2, self.c3, self.c4 = c[:, 0], c[:, 1], c[:, 2], c[:, 3]
        self.n = (dv.as_tensor(n_valid, t.float64, c.device) if n_valid is not None
                  else self.c1 + self.c2 + self.c3 + self.c4)
        self.D = self.c2 * self.c3 - self.c1 * self.c4
        self.pA = self.c1 + self.c2
        self.qA = self.c3 + self.c4
        self.pB = self.c1 + self.c3
        self.qB = self.c2 + self.c4
