"""Accessible-region masks from BED files (the part of pg_gpu/accessible.py
the LD drivers use: dropping variants outside callable regions)."""
from __future__ import annotations

import numpy as np


def read_bed_mask(path, chrom=None):
    """Return a function pos -> bool: True where a 1-based position lies in a
    BED interval (0-based, half-open) of ``chrom`` (all chromosomes if None)."""
    starts, ends = [], []
    with open(path) as fh:
        for line in fh:
            if not line.strip() or line.startswith(("#", "track", "browser")):
                continue
            f = line.split()
            if chrom is not None and f[0] != str(chrom):
                continue
            starts.append(int(f[1]))
            ends.append(int(f[2]))
    order = np.argsort(starts)
    s = np.asarray(starts, np.int64)[order]
    e = np.asarray(ends, np.int64)[order]
    if len(e):
        e = np.maximum.accumulate(e)

    def mask(pos):
        p0 = np.asarray(pos, np.int64) - 1
        k = np.searchsorted(s, p0, side="right") - 1
        ok = k >= 0
        out = np.zeros(p0.shape, bool)
        out[ok] = p0[ok] < e[k[ok]]
        return out

    return mask
