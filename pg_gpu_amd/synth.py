"""Seeded synthetic haplotype / genotype data for the LD benchmark workloads.

Founder-copying model: each population has F founder haplotypes whose
alleles are Bernoulli(p) per site; every sampled haplotype copies a uniformly
chosen founder and, between consecutive sites, switches to a fresh uniformly
chosen founder with probability 1 - exp(-rate * gap_bp).  With several
populations the per-population founder frequencies are Balding-Nichols draws
around a shared p ~ Beta(a, a).  Missing calls are i.i.d. -1.

Workloads (name: description):
  c0  1 pop x 200 haps, 5,000 SNPs on 1 Mb, no missing, 20 founders
  c1  2 pops x 500 haps, 200,000 SNPs on 20 Mb, 50 founders, BN F=0.05, 5% missing
  c2  1,000 diploids (pairs of founder-copying haplotypes), 100,000 SNPs on
      10 Mb, 50 founders, 2% missing genotypes
  c3  300 haps on a 250 Mb chromosome cut in 5,000 windows of 50 kb; SNPs per
      window clip(round(Gamma(0.8, scale=500)), 20, 4000), 50 founders
  c4  4 pops of 40/120/300/800 haps, 500,000 SNPs on 50 Mb, 50 founders,
      BN F=0.1, 3% missing; at a random 30% of sites the allele is replaced by
      a rare allele carried by 1..12 haplotypes in total (MAF < 0.01)
  g4  the c4 haplotypes paired into 4 pops of 20/60/150/400 diploid genotypes
      (a genotype is missing where either haplotype is); same sites, bins
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

SEED_BASE = 20261203


@dataclass
class Workload:
    name: str
    mat: np.ndarray               # int8 [n_rows, n_sites]
    positions: np.ndarray         # int64 [n_sites], sorted, unique
    pop_rows: list                # list of row-index lists
    mode: str                     # "hap" or "geno"
    bins: np.ndarray
    max_dist: float
    windows: np.ndarray = None    # [n_windows, 2] site ranges (c3)
    params: dict = field(default_factory=dict)


def _positions(rng, n, span):
    pos = rng.choice(np.arange(1, span + 1, dtype=np.int64), size=n, replace=False)
    return np.sort(pos)


def _founders(rng, n_founders, freqs):
    return (rng.random((n_founders, len(freqs))) < freqs[None, :]).astype(np.int8)


def _copy(rng, founders, positions, n_hap, rate, chunk=256):
    """Founder-copying haplotypes [n_hap, V]."""
    F, V = founders.shape
    gaps = np.diff(positions).astype(np.float64)
    p_switch = -np.expm1(-rate * gaps)                     # P(at least one switch)
    out = np.empty((n_hap, V), np.int8)
    cols = np.arange(V)
    for lo in range(0, n_hap, chunk):
        hi = min(n_hap, lo + chunk)
        m = hi - lo
        sw = np.zeros((m, V), bool)
        sw[:, 0] = True
        sw[:, 1:] = rng.random((m, V - 1)) < p_switch[None, :]
        new = rng.integers(0, F, size=(m, V))
        last = np.where(sw, cols[None, :], 0)
        np.maximum.accumulate(last, axis=1, out=last)
        who = np.take_along_axis(new, last, axis=1)
        out[lo:hi] = founders[who, cols[None, :]]
    return out


def _bn(rng, p, F):
    """Balding-Nichols population frequencies around p."""
    a = p * (1 - F) / F
    b = (1 - p) * (1 - F) / F
    return rng.beta(np.maximum(a, 1e-9), np.maximum(b, 1e-9))


def _missing(rng, mat, rate):
    if rate > 0:
        mat[rng.random(mat.shape) < rate] = -1
    return mat


def log_bins(lo, hi, n):
    return np.logspace(np.log10(lo), np.log10(hi), n + 1)


def make(name, scale=1.0, seed=None):
    """Build workload ``name`` (c0..c4, g4).  ``scale`` < 1 shrinks the number of
    sites (and the span, keeping the SNP density) for quick tests."""
    idx = int(name[1])
    rng = np.random.default_rng(SEED_BASE + idx if seed is None else seed)
    rate = 1e-5
    if name == "c0":
        V, span = int(5000 * scale), int(1_000_000 * scale)
        pos = _positions(rng, V, span)
        fr = rng.beta(0.3, 0.3, V)
        hap = _copy(rng, _founders(rng, 20, fr), pos, 200, rate)
        return Workload(name, hap, pos, [list(range(200))], "hap", log_bins(100, 100_000, 10),
                        100_000.0, params=dict(founders=20, V=V, span=span, missing=0.0))
    if name == "c1":
        V, span = int(200_000 * scale), int(20_000_000 * scale)
        pos = _positions(rng, V, span)
        p = rng.beta(0.3, 0.3, V)
        haps = [_copy(rng, _founders(rng, 50, _bn(rng, p, 0.05)), pos, 500, rate) for _ in range(2)]
        hap = _missing(rng, np.concatenate(haps, axis=0), 0.05)
        return Workload(name, hap, pos, [list(range(500)), list(range(500, 1000))], "hap",
                        log_bins(100, 200_000, 20), 200_000.0,
                        params=dict(founders=50, V=V, span=span, missing=0.05, bn_F=0.05))
    if name == "c2":
        V, span = int(100_000 * scale), int(10_000_000 * scale)
        pos = _positions(rng, V, span)
        fr = rng.beta(0.3, 0.3, V)
        hap = _copy(rng, _founders(rng, 50, fr), pos, 2000, rate)
        geno = (hap[0::2].astype(np.int16) + hap[1::2]).astype(np.int8)
        geno = _missing(rng, geno, 0.02)
        return Workload(name, geno, pos, [list(range(1000))], "geno", log_bins(100, 100_000, 15),
                        100_000.0, params=dict(founders=50, V=V, span=span, missing=0.02))
    if name == "c3":
        n_win = max(1, int(5000 * scale))
        wlen = 50_000
        k = np.clip(np.rint(rng.gamma(0.8, 500.0, n_win)), 20, 4000).astype(np.int64)
        pos_l, win = [], []
        off = 0
        for w in range(n_win):
            pw = np.sort(rng.choice(np.arange(w * wlen + 1, (w + 1) * wlen + 1), size=k[w],
                                    replace=False))
            pos_l.append(pw)
            win.append((off, off + k[w]))
            off += k[w]
        pos = np.concatenate(pos_l)
        fr = rng.beta(0.3, 0.3, len(pos))
        hap = _copy(rng, _founders(rng, 50, fr), pos, 300, rate)
        return Workload(name, hap, pos, [list(range(300))], "hap", np.array([0.0, wlen]),
                        float(wlen), windows=np.asarray(win, np.int64),
                        params=dict(founders=50, n_windows=n_win, gamma_shape=0.8,
                                    gamma_scale=500.0, clip=(20, 4000)))
    if name == "c4":
        V, span = int(500_000 * scale), int(50_000_000 * scale)
        pos = _positions(rng, V, span)
        p = rng.beta(0.3, 0.3, V)
        sizes = [40, 120, 300, 800]
        haps = [_copy(rng, _founders(rng, 50, _bn(rng, p, 0.1)), pos, n, rate) for n in sizes]
        hap = np.concatenate(haps, axis=0)
        n_tot = hap.shape[0]
        rare = rng.random(V) < 0.30
        idx = np.nonzero(rare)[0]
        hap[:, idx] = 0
        carriers = rng.integers(1, int(0.01 * n_tot) + 1, size=len(idx))
        for site, c in zip(idx, carriers):
            hap[rng.choice(n_tot, size=c, replace=False), site] = 1
        hap = _missing(rng, hap, 0.03)
        offs = np.concatenate([[0], np.cumsum(sizes)])
        return Workload(name, hap, pos, [list(range(offs[q], offs[q + 1])) for q in range(4)],
                        "hap", log_bins(100, 500_000, 25), 500_000.0,
                        params=dict(founders=50, V=V, span=span, missing=0.03, bn_F=0.1,
                                    rare_fraction=0.3))
    if name == "g4":
        h = make("c4", scale=scale, seed=seed)
        parts, rows, off = [], [], 0
        for r in h.pop_rows:
            a, b = h.mat[r[0::2]], h.mat[r[1::2]]
            g = (a.astype(np.int16) + b).astype(np.int8)
            g[(a < 0) | (b < 0)] = -1
            parts.append(g)
            rows.append(list(range(off, off + len(g))))
            off += len(g)
        return Workload(name, np.concatenate(parts, axis=0), h.positions, rows, "geno", h.bins,
                        h.max_dist, params=dict(h.params, diploid_of="c4"))
    raise ValueError(f"unknown workload {name}")
