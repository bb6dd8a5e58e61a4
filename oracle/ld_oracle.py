"""CPU oracle for pg_gpu's two-locus LD path (TEST INFRASTRUCTURE ONLY).

This module restates, in NumPy FP64 (plus exact summation), what the
reference computes on the binned-LD hot path.  It exists to *check* the HIP
implementation in ``pg_gpu_amd``; the product never imports it.  Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may use it.

Each function names the reference file:line whose behaviour it restates.
Sign and index conventions follow the reference exactly:

* haplotype counts per pair are ``[n11, n10, n01, n00]`` (c1..c4);
* ``D = c2*c3 - c1*c4`` in the DD/Dz/pi2 estimators (note the sign);
* ``pA = c1+c2``, ``qA = c3+c4``, ``pB = c1+c3``, ``qB = c2+c4`` are
  *counts*, not frequencies.

Pinning (what checks each part; see tests/test_oracle_*.py):

* pair enumeration: brute force i<j, d<=max_dist;
* counts: brute-force boolean counting and the reference's hand-VCF answers;
* haplotype estimators: two transcriptions of the reference (the
  ``ld_statistics.py`` formulas and the ``haplotype_kernels.py`` kernel
  formulas) must agree, every call must be an unbiased estimator of its
  population target by exact multinomial enumeration, and the end-to-end
  2-population result must match the moments golden fixture;
* r / r^2 / D': closed forms from the reference tests;
* names/specs: the reference's counts (3/15/45/105 names, 3/27/114/330
  unique raw calls).
"""
from __future__ import annotations

import math
from fractions import Fraction

import numpy as np

# ---------------------------------------------------------------------------
# L1 pair enumeration  (pg_gpu/ld_pipeline.py:57-122)
# ---------------------------------------------------------------------------


def pair_upper_bounds(positions, max_dist):
    """Exclusive right end ``j_max`` per anchor: count of positions <= pos+max.

    ld_pipeline.py:85-88 computes ``positions + max_dist`` in float64 and a
    right-sided searchsorted; we do the same so rounding of a fractional
    ``max_dist`` behaves identically.
    """
    pos = np.asarray(positions)
    ub = pos.astype(np.float64) + float(max_dist)
    return np.searchsorted(pos.astype(np.float64), ub, side="right")


def pairs_within_distance(positions, max_dist):
    """All (i, j), i<j, with pos[j] <= pos[i] + max_dist; ordered by i then j."""
    pos = np.asarray(positions)
    n = len(pos)
    if n < 2:
        return np.empty(0, np.int64), np.empty(0, np.int64)
    jmax = pair_upper_bounds(pos, max_dist)
    per = np.maximum(0, jmax - np.arange(n) - 1).astype(np.int64)
    total = int(per.sum())
    idx_i = np.repeat(np.arange(n, dtype=np.int64), per)
    start = np.concatenate([[0], np.cumsum(per)[:-1]])
    within = np.arange(total, dtype=np.int64) - np.repeat(start, per)
    idx_j = idx_i + 1 + within
    return idx_i, idx_j


def iter_pair_chunks(positions, max_dist, chunk_size):
    """Anchor-block chunking of the pair list (ld_pipeline.py:94-122)."""
    pos = np.asarray(positions)
    n = len(pos)
    if n < 2:
        return
    jmax = pair_upper_bounds(pos, max_dist)
    per = np.maximum(0, jmax - np.arange(n) - 1).astype(np.int64)
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
            ii = np.repeat(np.arange(v_lo, v_hi, dtype=np.int64), seg)
            start = np.concatenate([[0], np.cumsum(seg)[:-1]])
            jj = ii + 1 + (np.arange(m, dtype=np.int64) - np.repeat(start, seg))
            yield ii.astype(np.int32), jj.astype(np.int32)
        v_lo = v_hi


# ---------------------------------------------------------------------------
# L2 counting  (ld_pipeline.py:130-175 haplotype, :183-263 genotype)
# ---------------------------------------------------------------------------


def hap_counts(hap, idx_i, idx_j, rows=None):
    """4-way counts [n11, n10, n01, n00] and n_valid (None if no -1 in subset)."""
    h = np.asarray(hap)
    if rows is not None:
        h = h[np.asarray(rows, dtype=np.int64), :]
    a = h[:, np.asarray(idx_i)]
    b = h[:, np.asarray(idx_j)]
    has_missing = bool(np.any(h == -1))
    if has_missing:
        ok = (a >= 0) & (b >= 0)
        nv = ok.sum(axis=0, dtype=np.int32)
    else:
        ok = np.ones(a.shape, dtype=bool)
        nv = None
    cnt = np.stack([
        ((a == 1) & (b == 1) & ok).sum(axis=0, dtype=np.int32),
        ((a == 1) & (b == 0) & ok).sum(axis=0, dtype=np.int32),
        ((a == 0) & (b == 1) & ok).sum(axis=0, dtype=np.int32),
        ((a == 0) & (b == 0) & ok).sum(axis=0, dtype=np.int32),
    ], axis=1)
    return cnt, nv


def geno_counts(geno, idx_i, idx_j, rows=None):
    """9-way counts ordered (n00, n01, n02, n10, ..., n22) and n_valid."""
    g = np.asarray(geno)
    if rows is not None:
        g = g[np.asarray(rows, dtype=np.int64), :]
    a = g[:, np.asarray(idx_i)].astype(np.int32)
    b = g[:, np.asarray(idx_j)].astype(np.int32)
    ok = (a >= 0) & (b >= 0)
    combo = np.where(ok, a * 3 + b, -1)
    cnt = np.stack([(combo == c).sum(axis=0, dtype=np.int32) for c in range(9)], axis=1)
    nv = ok.sum(axis=0, dtype=np.int32)
    return cnt, (nv if bool(np.any(g < 0)) else None)


# ---------------------------------------------------------------------------
# Per-population features  (ld_pipeline.py:465-482 PopData)
# ---------------------------------------------------------------------------


class HapPop:
    """FP64 per-pair features of one population (counts, not frequencies)."""

    def __init__(self, counts, n_valid=None):
        c = np.asarray(counts, dtype=np.float64)
        self.c1, self.c2, self.c3, self.c4 = c[:, 0], c[:, 1], c[:, 2], c[:, 3]
        self.n = (np.asarray(n_valid, dtype=np.float64) if n_valid is not None
                  else self.c1 + self.c2 + self.c3 + self.c4)
        self.D = self.c2 * self.c3 - self.c1 * self.c4
        self.pA = self.c1 + self.c2
        self.qA = self.c3 + self.c4
        self.pB = self.c1 + self.c3
        self.qB = self.c2 + self.c4


def _ratio(num, den):
    """num/den where den > 0, else 0 (the reference's guard, e.g. haplotype_kernels.py:78)."""
    if isinstance(den, Fraction):
        return num / den if den > 0 else Fraction(0)
    num = np.asarray(num, dtype=np.float64)
    den = np.asarray(den, dtype=np.float64)
    out = np.zeros(np.broadcast(num, den).shape)
    ok = den > 0
    np.divide(num, den, out=out, where=ok)
    return out


def _ff(n, k):
    """Falling factorial n (n-1) ... (n-k+1) as float64 (exact below 2**53)."""
    if isinstance(n, Fraction):
        out = Fraction(1)
        for t in range(k):
            out *= (n - t)
        return out
    out = np.ones_like(n, dtype=np.float64)
    for t in range(k):
        out = out * (n - t)
    return out


# ---------------------------------------------------------------------------
# Haplotype estimators, kernel form  (haplotype_kernels.py:69-304)
# ---------------------------------------------------------------------------


def hap_dd(P, i, j):
    a = P[i]
    if i == j:  # haplotype_kernels.py:69-79
        num = (a.c1 * (a.c1 - 1) * a.c4 * (a.c4 - 1) + a.c2 * (a.c2 - 1) * a.c3 * (a.c3 - 1)
               - 2 * a.c1 * a.c2 * a.c3 * a.c4)
        return _ratio(num, _ff(a.n, 4))
    b = P[j]  # haplotype_kernels.py:81-89
    return _ratio(a.D * b.D, _ff(a.n, 2) * _ff(b.n, 2))


def hap_dz(P, i, j, k):
    if i == j == k:  # haplotype_kernels.py:95-107
        a = P[i]
        x = a.c1 * a.c4 - a.c2 * a.c3
        num = (x * (a.c3 + a.c4 - a.c1 - a.c2) * (a.c2 + a.c4 - a.c1 - a.c3)
               + x * (a.c2 + a.c3 - a.c1 - a.c4) + 2 * (a.c2 * a.c3 + a.c1 * a.c4))
        return _ratio(num, _ff(a.n, 4))
    if i == j:  # Dz(i,i,k): haplotype_kernels.py:109-120
        a, b = P[i], P[k]
        num = (a.qA - a.pA) * (-a.D) * (-b.c1 + b.c2 - b.c3 + b.c4)
        return _ratio(num, b.n * _ff(a.n, 3))
    if i == k:  # Dz(i,j,i): :122-133
        a, b = P[i], P[j]
        num = (-a.c1 + a.c2 - a.c3 + a.c4) * (-a.D) * (b.qA - b.pA)
        return _ratio(num, b.n * _ff(a.n, 3))
    if j == k:  # Dz(i,j,j): :135-147
        a, b = P[i], P[j]
        num = ((-a.D) * (-b.c1 + b.c2 + b.c3 - b.c4)
               + (-a.D) * (-b.c1 + b.c2 - b.c3 + b.c4) * (b.qA - b.pA))
        return _ratio(num, _ff(a.n, 2) * _ff(b.n, 2))
    a, b, c = P[i], P[j], P[k]  # :149-160
    num = -(a.D * (b.pA - b.qA) * (c.c1 - c.c2 + c.c3 - c.c4))
    return _ratio(num, _ff(a.n, 2) * b.n * c.n)


def pi2_case(i, j, k, l):
    """Case label used by the reference's pi2 dispatch (haplotype_kernels.py:463-484)."""
    cnt = {}
    for p in (i, j, k, l):
        cnt[p] = cnt.get(p, 0) + 1
    if len(cnt) == 1:
        return "same"
    if max(cnt.values()) == 3:
        return "triple"
    if i == j and k == l:
        return "iikk"
    if i == j:
        return "iikl"
    if k == l:
        return "ijkk"
    if (i == k and j == l) or (i == l and j == k):
        return "ijij"
    if len(cnt) == 3:
        return "shared"
    return "alldiff"


def _pi2_triple(T, S):
    """pi2 with pop T three times and pop S once (haplotype_kernels.py:268-289)."""
    tpA, tqA = T.c1 + T.c2, T.c3 + T.c4
    num = (-(tpA * T.c4 * S.pB) - (T.c2 * tqA * S.pB) + tpA * (T.c2 + T.c4) * tqA * S.pB
           + tpA * tqA * (-2 * S.c2 - 2 * S.c4) + tpA * T.c4 * S.qB + T.c2 * tqA * S.qB
           + tpA * (T.c1 + T.c3) * tqA * S.qB) / 2
    return _ratio(num, S.n * _ff(T.n, 3))


def _pi2_ijij(A, B):
    """pi2(i,j,i,j) / pi2(i,j,j,i) (haplotype_kernels.py:243-266)."""
    a1, a2, a3, a4 = A.c1, A.c2, A.c3, A.c4
    b1, b2, b3, b4 = B.c1, B.c2, B.c3, B.c4
    quartic = ((a2 + a4) * (a3 + a4) * (b1 + b2) * (b1 + b3)
               + (a1 + a3) * (a3 + a4) * (b1 + b2) * (b2 + b4)
               + (a1 + a2) * (a2 + a4) * (b1 + b3) * (b3 + b4)
               + (a1 + a2) * (a1 + a3) * (b2 + b4) * (b3 + b4))
    # Cubic correction, grouped by the b-count it multiplies.
    cubic = (b1 * (a4 - a2 * a3 - a2 * a4 - a3 * a4 - a4 * a4 - a4 * b1)
             + b2 * (a3 - a1 * a3 - a3 * a3 - a1 * a4 - a3 * a4 - a3 * b1 - a4 * b1 - a3 * b2)
             + b3 * (a2 - a1 * a2 - a2 * a2 - a1 * a4 - a2 * a4 - a2 * b1 - a4 * b1
                     - a1 * b2 - a4 * b2 - a2 * b3)
             + b4 * (a1 - a1 * a1 - a1 * a2 - a1 * a3 - a2 * a3 - a2 * b1 - a3 * b1
                     - a1 * b2 - a3 * b2 - a1 * b3 - a2 * b3 - a1 * b4))
    return _ratio((quartic + cubic) / 4, _ff(A.n, 2) * _ff(B.n, 2))


def _pi2_shared(S, A, B):
    """One pop shared across the two locus pairs (haplotype_kernels.py:218-241)."""
    s1, s2, s3, s4 = S.c1, S.c2, S.c3, S.c4
    inner = (s3 - 1) * B.c1 + s3 * B.c2 - B.c3 + s3 * B.c3 + s3 * B.c4 + s1 * B.qB
    num = (s4 * s4 * A.pA * B.pB
           + s2 * s2 * A.qA * B.pB
           + (s1 + s3 - 1) * (s3 * A.pA + s1 * A.qA) * B.qB
           + s4 * (s1 * A.qA * B.pB + A.c1 * inner + A.c2 * inner)
           + s2 * (s4 * (A.pA + A.qA) * B.pB
                   + s3 * (A.c1 * B.pB + A.c2 * B.pB + A.qA * B.qB)
                   + A.qA * ((s1 - 1) * B.c1 - B.c3 + s1 * (B.c2 + B.c3 + B.c4)))) / 4
    return _ratio(num, _ff(S.n, 2) * A.n * B.n)


def hap_pi2(P, i, j, k, l):
    case = pi2_case(i, j, k, l)
    if case == "same":  # haplotype_kernels.py:291-304
        a = P[i]
        num = (a.pA * a.pB * a.qB * a.qA
               - a.c1 * a.c4 * (-1 + a.c1 + 3 * a.c2 + 3 * a.c3 + a.c4)
               - a.c2 * a.c3 * (-1 + 3 * a.c1 + a.c2 + a.c3 + 3 * a.c4))
        return _ratio(num, _ff(a.n, 4))
    if case == "triple":  # normalised to (single, t, t, t): :549-561
        cnt = {}
        for p in (i, j, k, l):
            cnt[p] = cnt.get(p, 0) + 1
        t = [p for p, c in cnt.items() if c == 3][0]
        s = [p for p, c in cnt.items() if c == 1][0]
        return _pi2_triple(P[t], P[s])
    if case == "iikk":  # :180-190
        a, b = P[i], P[k]
        return _ratio(a.pA * a.qA * b.pB * b.qB, _ff(a.n, 2) * _ff(b.n, 2))
    if case == "iikl":  # :192-203
        a, b, c = P[i], P[k], P[l]
        return _ratio(a.pA * a.qA * (b.pB * c.qB + b.qB * c.pB), 2 * _ff(a.n, 2) * b.n * c.n)
    if case == "ijkk":  # :205-216
        a, b, c = P[i], P[j], P[k]
        return _ratio(c.pB * c.qB * (a.pA * b.qA + a.qA * b.pA), 2 * _ff(c.n, 2) * a.n * b.n)
    if case == "ijij":  # :243-266, I = first index, J = second
        return _pi2_ijij(P[i], P[j])
    if case == "shared":  # :526-547
        if i == k:
            s, a, b = i, j, l
        elif i == l:
            s, a, b = i, j, k
        elif j == k:
            s, a, b = j, i, l
        else:
            s, a, b = j, i, k
        return _pi2_shared(P[s], P[a], P[b])
    a, b, c, d = P[i], P[j], P[k], P[l]  # :167-178
    num = (a.pA * b.qA + a.qA * b.pA) * (c.pB * d.qB + c.qB * d.pB)
    return _ratio(num, 4 * a.n * b.n * c.n * d.n)


def hap_call(P, stat_type, pops):
    if stat_type == "dd":
        return hap_dd(P, *pops)
    if stat_type == "dz":
        return hap_dz(P, *pops)
    return hap_pi2(P, *pops)


# ---------------------------------------------------------------------------
# Haplotype estimators, ld_statistics form  (ld_statistics.py:1116-1617)
# This is the 2-population fast path's arithmetic; kept as a second
# transcription so that the two reference forms can be checked against each
# other.
# ---------------------------------------------------------------------------


def _split(counts):
    c = np.asarray(counts, dtype=np.float64)
    return c[:, 0], c[:, 1], c[:, 2], c[:, 3]


def _masked(num, den, ok):
    out = np.zeros(np.shape(num))
    out[ok] = num[ok] / den[ok]
    return out


def lds_dd_single(counts, n=None):
    c1, c2, c3, c4 = _split(counts)
    n = c1 + c2 + c3 + c4 if n is None else np.asarray(n, dtype=np.float64)
    num = c1 * (c1 - 1) * c4 * (c4 - 1) + c2 * (c2 - 1) * c3 * (c3 - 1) - 2 * c1 * c2 * c3 * c4
    return _masked(num, n * (n - 1) * (n - 2) * (n - 3), n >= 4)


def lds_dz_single(counts, n=None):
    c1, c2, c3, c4 = _split(counts)
    n = c1 + c2 + c3 + c4 if n is None else np.asarray(n, dtype=np.float64)
    x = c1 * c4 - c2 * c3
    num = (x * ((c3 + c4) - (c1 + c2)) * ((c2 + c4) - (c1 + c3)) + x * ((c2 + c3) - (c1 + c4))
           + 2 * (c2 * c3 + c1 * c4))
    return _masked(num, n * (n - 1) * (n - 2) * (n - 3), n >= 4)


def lds_pi2_single(counts, n=None):
    c1, c2, c3, c4 = _split(counts)
    n = c1 + c2 + c3 + c4 if n is None else np.asarray(n, dtype=np.float64)
    num = ((c1 + c2) * (c1 + c3) * (c2 + c4) * (c3 + c4)
           - c1 * c4 * (-1 + c1 + 3 * c2 + 3 * c3 + c4)
           - c2 * c3 * (-1 + 3 * c1 + c2 + c3 + 3 * c4))
    return _masked(num, n * (n - 1) * (n - 2) * (n - 3), n >= 4)


def lds_two_pop_15(counts0, counts1, n0=None, n1=None):
    """ld_pipeline.py:271-322 with ld_statistics.py's multi-pop formulas."""
    a1, a2, a3, a4 = _split(counts0)
    b1, b2, b3, b4 = _split(counts1)
    na = a1 + a2 + a3 + a4 if n0 is None else np.asarray(n0, np.float64)
    nb = b1 + b2 + b3 + b4 if n1 is None else np.asarray(n1, np.float64)
    pops = {0: (a1, a2, a3, a4, na), 1: (b1, b2, b3, b4, nb)}

    def dz3(p1, p2, p3):
        if p1 == p2:  # ld_statistics.py:1211-1223
            x1, x2, x3, x4, n1_ = pops[p1]
            y1, y2, y3, y4, n2_ = pops[p3]
            num = (-x1 - x2 + x3 + x4) * (-(x2 * x3) + x1 * x4) * (-y1 + y2 - y3 + y4)
            return _masked(num, n2_ * n1_ * (n1_ - 1) * (n1_ - 2), (n1_ >= 3) & (n2_ >= 1))
        if p1 == p3:  # :1225-1237
            x1, x2, x3, x4, n1_ = pops[p1]
            y1, y2, y3, y4, n2_ = pops[p2]
            num = (-x1 + x2 - x3 + x4) * (-(x2 * x3) + x1 * x4) * (-y1 - y2 + y3 + y4)
            return _masked(num, n2_ * n1_ * (n1_ - 1) * (n1_ - 2), (n1_ >= 3) & (n2_ >= 1))
        # p2 == p3: :1239-1250
        x1, x2, x3, x4, n1_ = pops[p1]
        y1, y2, y3, y4, n2_ = pops[p2]
        w = -(x2 * x3) + x1 * x4
        num = w * (-y1 + y2 + y3 - y4) + w * (-y1 + y2 - y3 + y4) * (-y1 - y2 + y3 + y4)
        return _masked(num, n1_ * (n1_ - 1) * n2_ * (n2_ - 1), (n1_ >= 2) & (n2_ >= 2))

    def iiij(single, triple):  # ld_statistics.py:1415-1445
        x1, x2, x3, x4, n1_ = pops[triple]
        y1, y2, y3, y4, n2_ = pops[single]
        num = (-((x1 + x2) * x4 * (y1 + y3)) - (x2 * (x3 + x4) * (y1 + y3))
               + ((x1 + x2) * (x2 + x4) * (x3 + x4) * (y1 + y3))
               + ((x1 + x2) * (x3 + x4) * (-2 * y2 - 2 * y4))
               + ((x1 + x2) * x4 * (y2 + y4)) + (x2 * (x3 + x4) * (y2 + y4))
               + ((x1 + x2) * (x1 + x3) * (x3 + x4) * (y2 + y4))) / 2.0
        return _masked(num, n2_ * n1_ * (n1_ - 1) * (n1_ - 2), (n1_ >= 3) & (n2_ >= 1))

    D0 = a2 * a3 - a1 * a4
    D1 = b2 * b3 - b1 * b4
    dd01 = _masked(D0 * D1, na * (na - 1) * nb * (nb - 1), (na >= 2) & (nb >= 2))
    # pi2(0,0,1,1): ld_statistics.py:1330-1341
    num1 = (a1 + a2) * (a3 + a4) * (b1 + b3) * (b2 + b4)
    num2 = (b1 + b2) * (b3 + b4) * (a1 + a3) * (a2 + a4)
    den = na * (na - 1) * nb * (nb - 1)
    ok = (na >= 2) & (nb >= 2)
    p0011 = np.zeros_like(na)
    p0011[ok] = 0.5 * (num1[ok] + num2[ok]) / den[ok]
    # pi2(0,1,0,1): ld_statistics.py:1351-1407 (same polynomial as the kernel form)
    p0101 = _pi2_ijij(HapPop(np.stack([a1, a2, a3, a4], 1), na),
                      HapPop(np.stack([b1, b2, b3, b4], 1), nb))
    return np.stack([
        lds_dd_single(counts0, na), dd01, lds_dd_single(counts1, nb),
        lds_dz_single(counts0, na), 0.5 * (dz3(0, 0, 1) + dz3(0, 1, 0)), dz3(0, 1, 1),
        dz3(1, 0, 0), 0.5 * (dz3(1, 0, 1) + dz3(1, 1, 0)), lds_dz_single(counts1, nb),
        lds_pi2_single(counts0, na), iiij(1, 0), p0011, p0101, iiij(0, 1),
        lds_pi2_single(counts1, nb),
    ], axis=1)


# ---------------------------------------------------------------------------
# L5 stat names / specs  (ld_pipeline.py:357-457)
# ---------------------------------------------------------------------------


def ld_names(num_pops):
    out = []
    P = range(num_pops)
    out += [f"DD_{a}_{b}" for a in P for b in P if b >= a]
    out += [f"Dz_{a}_{b}_{c}" for a in P for b in P for c in P if c >= b]
    for a in P:
        for b in P:
            if b < a:
                continue
            for c in P:
                if c < a:
                    continue
                for d in P:
                    if d < c:
                        continue
                    if c == a == d and b != a:
                        continue
                    if a == c and d < b:
                        continue
                    out.append(f"pi2_{a}_{b}_{c}_{d}")
    return out


def het_names(num_pops):
    return [f"H_{a}_{b}" for a in range(num_pops) for b in range(a, num_pops)]


def stat_specs(num_pops):
    """(name, [(weight, type, pops), ...]) -- moments' _call_sgc averaging."""
    specs = []
    for name in ld_names(num_pops):
        parts = name.split("_")
        kind, pp = parts[0], tuple(int(x) for x in parts[1:])
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
                calls = [(0.125, x) for x in ((a, b, c, d), (a, b, d, c), (b, a, c, d), (b, a, d, c),
                                              (c, d, a, b), (d, c, a, b), (c, d, b, a), (d, c, b, a))]
            specs.append((name, [(w, "pi2", x) for w, x in calls]))
    return specs


def multi_pop_stats_hap(counts_list, nvalid_list, specs=None):
    """(N, n_stats) via unique raw calls + weighted sums (haplotype_kernels.py:584-640)."""
    P = [HapPop(c, nv) for c, nv in zip(counts_list, nvalid_list)]
    specs = stat_specs(len(P)) if specs is None else specs
    cache = {}
    out = np.zeros((P[0].n.shape[0], len(specs)))
    for s, (_, calls) in enumerate(specs):
        acc = None
        for w, t, pp in calls:
            key = (t, pp)
            if key not in cache:
                cache[key] = hap_call(P, t, pp)
            term = cache[key] if (w == 1.0 and len(calls) == 1) else w * cache[key]
            acc = term if acc is None else acc + term
        out[:, s] = acc
    return out


def hap_stats_for_pairs(counts_list, nvalid_list):
    """Per-pair statistic matrix the way moments_ld._compute_ld_sums builds it:
    the 2-pop fast path for P == 2 (moments_ld.py:324-327), specs otherwise."""
    if len(counts_list) == 2:
        return lds_two_pop_15(counts_list[0], counts_list[1], nvalid_list[0], nvalid_list[1])
    return multi_pop_stats_hap(counts_list, nvalid_list)


# ---------------------------------------------------------------------------
# r, r^2, D'  (haplotype_kernels.py:20-62; ld_statistics.py:153-246)
# ---------------------------------------------------------------------------


def r_family(counts, n_valid=None):
    a, b, c, d = _split(counts)
    n = a + b + c + d if n_valid is None else np.asarray(n_valid, np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        pA, pB = (a + b) / n, (a + c) / n
        D = (a * d - b * c) / (n * n)
        den = pA * (1.0 - pA) * pB * (1.0 - pB)
        r = np.where(den > 0, D / np.sqrt(np.where(den > 0, den, 1.0)), np.nan)
        r2 = np.where(den > 0, D * D / np.where(den > 0, den, 1.0), np.nan)
        qA, qB = 1.0 - pA, 1.0 - pB
        dmax = np.where(D >= 0, np.minimum(pA * qB, qA * pB), np.minimum(pA * pB, qA * qB))
        dp = np.where(dmax > 0, D / np.where(dmax > 0, dmax, 1.0), np.nan)
    return r, r2, dp


# ---------------------------------------------------------------------------
# L6 binning  (moments_ld.py:248-343; haplotype_matrix.py:1592-1866)
# ---------------------------------------------------------------------------


def digitize_bins(dist, bins):
    """np.digitize(d, bins) - 1; entries outside [bins[0], bins[-1]) give -1 / n_bins."""
    return np.digitize(dist, np.asarray(bins, np.float64)) - 1


def binned_sums_hap(hap, positions, pop_rows, bins, **kw):
    """Binned LD sums for haplotype data (moments_ld.py:248-343, haplotype branch)."""
    return binned_sums(hap, positions, pop_rows, bins, mode="hap", **kw)


def binned_sums_geno(geno, positions, pop_rows, bins, **kw):
    """Binned LD sums for genotype data (moments_ld.py:248-343, genotype branch:
    9-way counts and the kernel-form estimators of every population count)."""
    return binned_sums(geno, positions, pop_rows, bins, mode="geno", **kw)


def binned_sums(mat, positions, pop_rows, bins, max_dist=None, coords=None,
                chunk=200_000, exact=True, with_abs=False, anchors=None, mode="hap", r2=False):
    """Binned LD sums (moments_ld.py:248-343).

    With ``r2`` two columns follow the LD statistics: the per-bin sum of r^2
    of the first population over pairs where it is defined (haplotypes:
    ld_statistics.r_squared, ld_statistics.py:190-217; genotypes: the
    pairwise-complete dosage r^2) and the number of such pairs.

    Returns (sums[n_bins, n_stats], pair_counts[n_bins], n_enumerated) and,
    with ``with_abs``, also the per-bin sums of |value| (the scale against which
    rounding of individual terms is judged).
    With ``exact`` the per-bin sums of the FP64 per-pair values are formed
    with math.fsum (correctly rounded); otherwise with np.add.at in pair order.
    ``anchors=(lo, hi)`` restricts the work to pairs whose first site i lies
    in [lo, hi) (a deterministic prefix / block of the pair list).
    """
    bins = np.asarray(bins, np.float64)
    nb = len(bins) - 1
    max_dist = float(bins[-1]) if max_dist is None else float(max_dist)
    pos = np.asarray(positions)
    co = pos.astype(np.float64) if coords is None else np.asarray(coords, np.float64)
    P = len(pop_rows)
    n_stats = len(ld_names(P)) + (2 if r2 else 0)
    parts = [[[] for _ in range(n_stats)] for _ in range(nb)]
    sums = np.zeros((nb, n_stats))
    abs_sums = np.zeros((nb, n_stats))
    counts = np.zeros(nb, np.int64)
    n_enum = 0
    for ii, jj in iter_pair_chunks(pos, max_dist, chunk):
        if anchors is not None:
            keep = (ii >= anchors[0]) & (ii < anchors[1])
            if not keep.any():
                if ii[0] >= anchors[1]:
                    break
                continue
            ii, jj = ii[keep], jj[keep]
        n_enum += len(ii)
        d = np.abs(co[jj] - co[ii]) if coords is not None else (pos[jj] - pos[ii]).astype(np.float64)
        b = digitize_bins(d, bins)
        cl, nl = [], []
        for rows in pop_rows:
            c, nv = (hap_counts if mode == "hap" else geno_counts)(mat, ii, jj, rows)
            cl.append(c)
            nl.append(nv)
        if mode == "hap":
            st = hap_stats_for_pairs(cl, nl)
        else:
            from .geno_oracle import multi_pop_stats_geno

            st = multi_pop_stats_geno(cl, nl)
        if r2:
            if mode == "hap":
                rr = r_family(cl[0], nl[0])[1]
            else:
                from .geno_oracle import r2_dosage

                rr = r2_dosage(cl[0])
            fin = np.isfinite(rr)
            st = np.concatenate([st, np.where(fin, rr, 0.0)[:, None],
                                 fin.astype(np.float64)[:, None]], axis=1)
        ok = (b >= 0) & (b < nb)
        counts += np.bincount(b[ok], minlength=nb)[:nb]
        np.add.at(abs_sums, b[ok], np.abs(st[ok]))
        if exact:
            order = np.argsort(b[ok], kind="stable")
            bs = b[ok][order]
            vs = st[ok][order]
            cuts = np.searchsorted(bs, np.arange(nb + 1))
            for k in range(nb):
                seg = vs[cuts[k]:cuts[k + 1]]
                if len(seg):
                    for s in range(n_stats):
                        parts[k][s].append(math.fsum(seg[:, s]))
        else:
            np.add.at(sums, b[ok], st[ok])
    if exact:
        for k in range(nb):
            for s in range(n_stats):
                sums[k, s] = math.fsum(parts[k][s]) if parts[k][s] else 0.0
    if with_abs:
        return sums, counts, n_enum, abs_sums
    return sums, counts, n_enum


def window_sums(mat, windows, rows=None, mode="hap", exact=True):
    """Per-window sums over every i < j pair inside each site range [lo, hi):
    columns DD, Dz, pi2 (single population), r^2 sum where defined, number of
    pairs with r^2 defined; plus the pair count per window.  The reference
    form: pairs of a window (ld_pipeline.py:57-122 with the cutoff spanning
    the window), counts (:130-263), ld_statistics dd/dz/pi2/r_squared."""
    windows = np.asarray(windows, np.int64).reshape(-1, 2)
    sums = np.zeros((len(windows), 5))
    counts = np.zeros(len(windows), np.int64)
    for w, (lo, hi) in enumerate(windows):
        k = hi - lo
        if k < 2:
            continue
        ii, jj = np.triu_indices(k, 1)
        ii, jj = ii + lo, jj + lo
        if mode == "hap":
            c, nv = hap_counts(mat, ii, jj, rows)
            st = hap_stats_for_pairs([c], [nv])
            rr = r_family(c, nv)[1]
        else:
            from .geno_oracle import multi_pop_stats_geno, r2_dosage

            c, nv = geno_counts(mat, ii, jj, rows)
            st = multi_pop_stats_geno([c], [nv])
            rr = r2_dosage(c)
        fin = np.isfinite(rr)
        cols = [st[:, 0], st[:, 1], st[:, 2], np.where(fin, rr, 0.0), fin.astype(np.float64)]
        for q, v in enumerate(cols):
            sums[w, q] = math.fsum(v) if exact else float(np.sum(v))
        counts[w] = len(ii)
    return sums, counts


# ---------------------------------------------------------------------------
# Dense tile consumers (ld_statistics.py:249-1021, haplotype_matrix.py:1092-1170)
# ---------------------------------------------------------------------------


def segregating(h, missing_data="include"):
    """Columns kept by _prepare_segregating (ld_statistics.py:249-278)."""
    h = np.asarray(h)
    if missing_data == "exclude":
        h = h[:, (h < 0).sum(axis=0) == 0]
    valid = (h >= 0).sum(axis=0)
    dac = np.maximum(h, 0).astype(np.int64).sum(axis=0)
    return h[:, (dac > 0) & (dac < valid)]


def _dense_counts(h):
    hc = np.where(h >= 0, h, 0).astype(np.float64)
    vm = (h >= 0).astype(np.float64)
    c1 = hc.T @ hc
    s12 = hc.T @ vm
    s13 = vm.T @ hc
    n = vm.T @ vm
    return c1, s12 - c1, s13 - c1, n - s12 - s13 + c1, n, hc, vm


def sigma_d2_matrix(h):
    """(ratio, n >= 4 flag) matrices of _tile_sigma_d2 (ld_statistics.py:311-341),
    diagonal zero."""
    c1, c2, c3, c4, n, _, _ = _dense_counts(h)
    dd = c1 * (c1 - 1) * c4 * (c4 - 1) + c2 * (c2 - 1) * c3 * (c3 - 1) - 2 * c1 * c2 * c3 * c4
    s12, s13, s24, s34 = c1 + c2, c1 + c3, c2 + c4, c3 + c4
    pi2 = (s12 * s13 * s24 * s34 - c1 * c4 * (-1 + c1 + 3 * c2 + 3 * c3 + c4)
           - c2 * c3 * (-1 + 3 * c1 + c2 + c3 + 3 * c4))
    valid = n >= 4
    with np.errstate(divide="ignore", invalid="ignore"):
        val = np.where(valid & (pi2 != 0), dd / np.where(pi2 != 0, pi2, 1.0), 0.0)
    np.fill_diagonal(val, 0.0)
    flag = valid.astype(np.float64)
    np.fill_diagonal(flag, 0.0)
    return val, flag


def r2_naive_matrix(h):
    """HaplotypeMatrix.pairwise_r2 / _tile_r2_naive (haplotype_matrix.py:1092-1170)."""
    c1, _, _, _, n, hc, vm = _dense_counts(h)
    nv = vm.sum(axis=0)
    with np.errstate(divide="ignore", invalid="ignore"):
        p = np.where(nv > 0, hc.sum(axis=0) / np.where(nv > 0, nv, 1), 0.0)
        pab = np.where(n > 0, c1 / np.where(n > 0, n, 1), 0.0)
        D = pab - np.outer(p, p)
        den = np.outer(p * (1 - p), p * (1 - p))
        r2 = np.where(den > 0, D ** 2 / np.where(den > 0, den, 1), 0.0)
    np.fill_diagonal(r2, 0.0)
    return r2


def r2_dosage_matrix(g):
    """_r2_matrix_diploid (ld_statistics.py:973-1021): per-site means over
    valid individuals, missing entries zeroed after centring."""
    g = np.asarray(g)
    vm = (g >= 0).astype(np.float64)
    gc = np.where(g >= 0, g, 0).astype(np.float64)
    nv = vm.sum(axis=0)
    with np.errstate(divide="ignore", invalid="ignore"):
        mean = np.where(nv > 0, gc.sum(axis=0) / np.where(nv > 0, nv, 1), 0.0)
    gn = (gc - mean[None, :]) * vm
    var = (gn ** 2).sum(axis=0)
    cov = gn.T @ gn
    den = np.sqrt(np.outer(var, var))
    with np.errstate(divide="ignore", invalid="ignore"):
        r2 = np.where(den > 0, (cov / np.where(den > 0, den, 1)) ** 2, 0.0)
    np.fill_diagonal(r2, 0.0)
    return r2


def zns_value(r2, flag=None):
    """ZnS from a full r^2 (or sigma_D^2) matrix (ld_statistics.py:534-731)."""
    m = r2.shape[0]
    if m < 2:
        return 0.0
    if flag is not None:
        n = flag.sum()
        return float(r2.sum() / n) if n > 0 else 0.0
    return float((r2.sum() - np.trace(r2)) / (m * (m - 1)))


def omega_value(r2):
    """Omega by direct summation over split points l = 3..m-2, upper-triangle
    pairs (ld_statistics.py:761-862 computes the same with prefix sums)."""
    m = r2.shape[0]
    if m < 5:
        return 0.0
    u = np.triu(r2, 1)
    best = 0.0
    for l in range(3, m - 1):
        left = u[:l, :l].sum()
        right = u[l:, l:].sum()
        cross = u[:l, l:].sum()
        nw = l * (l - 1) // 2 + (m - l) * (m - l - 1) // 2
        nc = l * (m - l)
        if nw > 0 and nc > 0 and cross > 0:
            best = max(best, ((left + right) / nw) / (cross / nc))
    return float(best)


def heterozygosity_hap(hap, pop_rows):
    """H_i_j sums over sites (moments_ld.py:346-388, haplotype branch).

    Missing (-1) calls count as reference alleles and the sample size is the
    number of rows in the population, as in the reference (:367-368).
    """
    alt, ref, nh = [], [], []
    for rows in pop_rows:
        h = np.asarray(hap)[np.asarray(rows, np.int64), :]
        a = np.maximum(h, 0).astype(np.int64).sum(axis=0).astype(np.float64)
        n = float(len(rows))
        alt.append(a)
        ref.append(n - a)
        nh.append(n)
    out = {}
    P = len(pop_rows)
    for i in range(P):
        for j in range(i, P):
            if i == j:
                v = np.sum(2.0 * ref[i] * alt[i] / (nh[i] * (nh[i] - 1)))
            else:
                v = np.sum((ref[i] * alt[j] + alt[i] * ref[j]) / (nh[i] * nh[j]))
            out[f"H_{i}_{j}"] = float(v)
    return out


def biallelic_filter_hap(hap):
    """Keep sites with a 0 and a 1 and no allele >= 2 (haplotype_matrix.py:724-789)."""
    h = np.asarray(hap)
    keep = (np.sum(h >= 2, axis=0) == 0) & (np.sum(h == 0, axis=0) > 0) & (np.sum(h == 1, axis=0) > 0)
    return np.nonzero(keep)[0]


# ---------------------------------------------------------------------------
# Exact (rational) evaluation used by the unbiasedness checks
# ---------------------------------------------------------------------------


class _FracPop:
    def __init__(self, c):
        self.c1, self.c2, self.c3, self.c4 = (Fraction(int(x)) for x in c)
        self.n = self.c1 + self.c2 + self.c3 + self.c4
        self.D = self.c2 * self.c3 - self.c1 * self.c4
        self.pA, self.qA = self.c1 + self.c2, self.c3 + self.c4
        self.pB, self.qB = self.c1 + self.c3, self.c2 + self.c4


def exact_hap_call(counts_per_pop, stat_type, pops):
    """Rational value of one raw call; same formulas as hap_call, in Fractions."""
    return hap_call([_FracPop(c) for c in counts_per_pop], stat_type, pops)


def windowed_r_squared(hap, positions, bp_bins, percentile=50, rows=None, estimator="r2"):
    """Percentiles of r^2 per distance bin over all i < j pairs
    (haplotype_matrix.py:1235-1317): 'r2' from each pair's counts,
    'rogers_huff' from the dosages of adjacent haplotype pairs (NaN where a
    site is monomorphic, ld_statistics.py:446-466)."""
    h = np.asarray(hap)
    if rows is not None:
        h = h[np.asarray(rows)]
    m = h.shape[1]
    ii, jj = np.triu_indices(m, 1)
    if estimator == "rogers_huff":
        g = h[0::2].astype(np.int64) + h[1::2]
        poly = g.var(axis=0) > 0
        r2 = np.where(poly[ii] & poly[jj], r2_dosage_matrix(g)[ii, jj], np.nan)
    else:
        c, nv = hap_counts(h, ii, jj)
        r2 = r_family(c, nv)[1]
    pos = np.asarray(positions, np.float64)
    b = np.digitize(pos[jj] - pos[ii], np.asarray(bp_bins, np.float64)) - 1
    nb = len(bp_bins) - 1
    ok = (b >= 0) & (b < nb) & ~np.isnan(r2)
    pct = np.atleast_1d(percentile)
    res = np.full((nb, len(pct)), np.nan)
    cnt = np.zeros(nb, dtype=int)
    for k in range(nb):
        sel = r2[ok & (b == k)]
        cnt[k] = len(sel)
        if len(sel):
            res[k] = np.percentile(sel, pct)
    return (res[:, 0] if res.shape[1] == 1 else res), cnt


# ---------------------------------------------------------------------------
# Container passes and the moments driver's host steps
# (genotype_matrix.py:247-340, :506-539; haplotype_matrix.py:1320-1589;
#  moments_ld.py:166-245, :256-279, :346-388)
# ---------------------------------------------------------------------------


def hap_to_geno(hap):
    """GenotypeMatrix.from_haplotype_matrix (genotype_matrix.py:263-291):
    rows (2i, 2i+1) summed after max(., 0); -1 where either is negative."""
    h = np.asarray(hap)
    h1, h2 = h[0::2], h[1::2]
    g = (np.maximum(h1, 0).astype(np.int16) + np.maximum(h2, 0)).astype(np.int8)
    g[(h1 < 0) | (h2 < 0)] = -1
    return g


def geno_to_hap(geno):
    """GenotypeMatrix.to_haplotype_matrix (genotype_matrix.py:319-336)."""
    g = np.asarray(geno)
    gg = np.maximum(g, 0)
    hap = np.zeros((2 * g.shape[0], g.shape[1]), np.int8)
    hap[0::2] = (gg >= 1)
    hap[1::2] = (gg >= 2)
    hap[0::2][g < 0] = -1
    hap[1::2][g < 0] = -1
    return hap


def biallelic_filter_geno(geno, rows=None):
    """Kept columns of GenotypeMatrix.apply_biallelic_filter
    (genotype_matrix.py:516-523); with ``rows`` the union filter of
    moments_ld._compute_ld_sums over those individuals, each counted once
    (moments_ld.py:256-276)."""
    g = np.asarray(geno)
    if rows is not None:
        g = g[np.asarray(sorted(set(int(r) for r in rows)), np.int64)]
    valid = g >= 0
    alt = np.where(valid, g, 0).astype(np.int64).sum(axis=0)
    nv = valid.sum(axis=0)
    return np.nonzero((alt > 0) & (alt < 2 * nv) & (nv >= 2))[0]


def heterozygosity_geno(geno, pop_rows):
    """H_i_j sums over sites for genotype data (moments_ld.py:356-388,
    genotype branch): alt = sum of valid genotypes, n = 2 x valid
    individuals, ref = n - alt.  Sites where a term is not finite (a
    population with fewer than one valid individual, or one for H_i_i)
    make the sum NaN, as the reference's cp.sum does."""
    alt, ref, nh = [], [], []
    for rows in pop_rows:
        g = np.asarray(geno)[np.asarray(rows, np.int64), :]
        v = g >= 0
        a = np.where(v, g, 0).astype(np.int64).sum(axis=0).astype(np.float64)
        n = 2.0 * v.sum(axis=0).astype(np.float64)
        alt.append(a)
        ref.append(n - a)
        nh.append(n)
    out = {}
    P = len(pop_rows)
    with np.errstate(divide="ignore", invalid="ignore"):
        for i in range(P):
            for j in range(i, P):
                if i == j:
                    v = np.sum(2.0 * ref[i] * alt[i] / (nh[i] * (nh[i] - 1)))
                else:
                    v = np.sum((ref[i] * alt[j] + alt[i] * ref[j]) / (nh[i] * nh[j]))
                out[f"H_{i}_{j}"] = float(v)
    return out


def read_rec_map(path):
    """(positions, Morgans) from a whitespace map file of (pos, cM) rows;
    lines that do not parse are skipped (moments_ld.py:211-223)."""
    mp, mv = [], []
    with open(path) as fh:
        for line in fh:
            parts = line.strip().split()
            if len(parts) >= 2:
                try:
                    mp.append(float(parts[0]))
                    mv.append(float(parts[1]) / 100.0)
                except ValueError:
                    continue
    return np.array(mp), np.array(mv)


def interpolate_genetic_distances(positions, rec_map_file):
    """Per-variant map position in Morgans (moments_ld.py:211-223)."""
    mp, mv = read_rec_map(rec_map_file)
    return np.interp(positions, mp, mv)


def max_bp_for_r_dist(positions, gen_dists, max_r):
    """Conservative bp cutoff for a Morgan cutoff (moments_ld.py:226-245):
    max_r / (smallest positive local rate) x 1.1, capped at the span."""
    chrom_len = float(positions[-1] - positions[0])
    if len(positions) < 2:
        return chrom_len
    bp = np.diff(positions).astype(np.float64)
    rd = np.diff(gen_dists)
    ok = bp > 0
    if not np.any(ok):
        return chrom_len
    rates = rd[ok] / bp[ok]
    min_rate = np.min(rates[rates > 0]) if np.any(rates > 0) else max_r / chrom_len
    return min(max_r / min_rate * 1.1, chrom_len)


def moments_ld_stats(mat, positions, pop_rows, mode="hap", bp_bins=None, r_bins=None,
                     rec_map_file=None):
    """The {'bins', 'sums', 'stats'} content of moments_ld.compute_ld_statistics
    (moments_ld.py:148-203) for an already loaded matrix: bp or Morgan bins,
    the genotype union filter before the pair loop, heterozygosity over the
    unfiltered matrix.  Also returned: per-bin pair counts, per-bin sums of
    |value| ("abs", the scale for rounding checks) and the bp cutoff."""
    pos = np.asarray(positions, np.int64)
    if r_bins is not None:
        bins = np.asarray(r_bins, np.float64)
        gen = interpolate_genetic_distances(pos, rec_map_file)
        max_bp = max_bp_for_r_dist(pos, gen, float(bins[-1]))
    else:
        bins = np.asarray(bp_bins, np.float64)
        gen = None
        max_bp = float(bins[-1])
    P = len(pop_rows)
    if mode == "geno":
        keep = biallelic_filter_geno(mat, [r for pr in pop_rows for r in pr])
        sums, counts, _, absr = binned_sums(np.asarray(mat)[:, keep], pos[keep], pop_rows, bins,
                                            max_dist=max_bp,
                                            coords=None if gen is None else gen[keep],
                                            mode="geno", with_abs=True)
        het = heterozygosity_geno(mat, pop_rows)
    else:
        sums, counts, _, absr = binned_sums(mat, pos, pop_rows, bins, max_dist=max_bp,
                                            coords=gen, mode="hap", with_abs=True)
        het = heterozygosity_hap(mat, pop_rows)
    hn = het_names(P)
    return {"bins": [(float(bins[i]), float(bins[i + 1])) for i in range(len(bins) - 1)],
            "sums": [sums[i] for i in range(len(bins) - 1)] + [np.array([het[h] for h in hn])],
            "stats": (ld_names(P), hn), "counts": counts, "abs": absr, "max_bp": max_bp}


def tally_upper(hap, rows=None):
    """(counts [V(V-1)/2, 4], n_valid or None) of every variant pair i < j in
    np.triu_indices order (haplotype_matrix.py:1320-1433), counting only
    haplotypes valid at both variants; n_valid is None when the rows hold no
    missing call.  (The reference's no-missing branch uses an int8 matmul
    whose n11 wraps above 127; this restatement keeps exact counts.)"""
    h = np.asarray(hap)
    if rows is not None:
        h = h[np.asarray(rows, np.int64)]
    ii, jj = np.triu_indices(h.shape[1], 1)
    return hap_counts(h, ii, jj)


# ---------------------------------------------------------------------------
# Counting-stage checksum over whole 32 x 32 site tiles
# ---------------------------------------------------------------------------


def band_count_checksum(hap, pop_rows, tiles, marginals):
    """Per population, the sums over every pair (i, j) of every listed tile
    (rows 32 rb .. 32 rb + 31, columns 32 cb .. 32 cb + 31, sites past the end
    contributing zero samples) of the four count fields the counting stage
    produces: with ``marginals`` (n11, pA, pB, n), else (n11, n10, n01, n00);
    counts over samples valid at both sites when the population has a missing
    call, otherwise pA / pB are the sites' allele tallies and n the population
    size for every slot.  Evaluated as dot products of per-block sample
    tallies: sum_{i in R, j in C} sum_s a_s(i) b_s(j) = sum_s A_s(R) B_s(C)."""
    h = np.asarray(hap)
    V = h.shape[1]
    nb = (V + 31) // 32
    tiles = np.asarray(tiles, np.int64).reshape(-1, 2)
    out = np.zeros((len(pop_rows), 4), np.int64)

    def blocks(x):   # [K, V] 0/1 -> [nb, K] per-block column sums
        xp = np.zeros((x.shape[0], nb * 32), np.int64)
        xp[:, :V] = x
        return xp.reshape(x.shape[0], nb, 32).sum(axis=2).T.copy()

    def tile_dots(A, B):
        tot = 0
        for lo in range(0, len(tiles), 20_000):
            t = tiles[lo:lo + 20_000]
            tot += int(np.einsum("ts,ts->", A[t[:, 0]], B[t[:, 1]]))
        return tot

    for p, rows in enumerate(pop_rows):
        x = h[np.asarray(rows, np.int64)]
        K = x.shape[0]
        H = blocks(x == 1)
        hh = tile_dots(H, H)
        if np.any(x < 0):
            Vb = blocks(x >= 0)
            hv, vh, vv = tile_dots(H, Vb), tile_dots(Vb, H), tile_dots(Vb, Vb)
        else:
            ks = H.sum(axis=1)                 # per block: sum of the sites' tallies
            hv = 32 * int(ks[tiles[:, 0]].sum())
            vh = 32 * int(ks[tiles[:, 1]].sum())
            vv = 1024 * K * len(tiles)
        if marginals:
            out[p] = (hh, hv, vh, vv)
        else:
            out[p] = (hh, hv - hh, vh - hh, vv - hv - vh + hh)
    return out
