"""Per-window all-pairs LD over ragged genomic windows on one GPU.

Every i < j pair of sites inside a window contributes (no distance cutoff);
each window gets the single-population DD, Dz and pi2 sums (the estimators of
pg_gpu/ld_statistics.py:13-129), the sum of r^2 over pairs where it is
defined and the number of such pairs (ld_statistics.r_squared, :190-217;
for genotypes the pairwise-complete dosage r^2), and its pair count.  All
windows run in one persistent launch over a (window, tile) work list, so
windows of 20 and of 4,000 sites share the machine without per-window
launches.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from .engine import TILE, LDPanels, decode_limbs


def windows_from_positions(positions, window_len, start=0.0, end=None):
    """Site ranges [lo, hi) of the windows [start + w*len, start + (w+1)*len)
    over sorted positions; windows without sites are kept (empty ranges)."""
    pos = np.asarray(positions, np.float64)
    if end is None:
        end = float(pos[-1]) + 1.0 if len(pos) else start
    n = max(0, int(np.ceil((end - start) / window_len)))
    edges = start + window_len * np.arange(n + 1)
    idx = np.searchsorted(pos, edges, side="left")
    return np.stack([idx[:-1], idx[1:]], axis=1).astype(np.int64)


def window_work(windows, tile=TILE):
    """(window, row block, col block) triples covering each window's upper
    triangle with ``tile``-site blocks (32 or 16), window by window."""
    windows = np.asarray(windows, np.int64).reshape(-1, 2)
    parts = []
    for w, (lo, hi) in enumerate(windows):
        if hi - lo < 2:
            continue
        b0, b1 = lo // tile, (hi - 1) // tile
        rb, cb = np.triu_indices(b1 - b0 + 1)
        parts.append(np.stack([np.full(len(rb), w), rb + b0, cb + b0], axis=1))
    if not parts:
        return np.zeros((0, 3), np.int32)
    return np.concatenate(parts).astype(np.int32)


@dataclass
class WindowResult:
    dd: np.ndarray         # [n_windows] sums
    dz: np.ndarray
    pi2: np.ndarray
    r2_sum: np.ndarray     # sum of r^2 over pairs where it is defined
    r2_count: np.ndarray   # pairs with r^2 defined
    n_pairs: np.ndarray    # i < j pairs in the window

    @property
    def mean_r2(self):
        with np.errstate(invalid="ignore", divide="ignore"):
            return np.where(self.r2_count > 0, self.r2_sum / np.maximum(self.r2_count, 1), np.nan)


class WindowPlan:
    """Device copies of the windows and work list of one job.  ``tile`` is
    the side of the work items: 32 (32 x 32 MFMA tiles) or 16 (16 x 16,
    haplotypes); autotune.tune_windows picks the faster for a job."""

    def __init__(self, pan: LDPanels, windows, tile=TILE, group=0):
        torch = _torch()
        if pan.P != 1:
            raise NotImplementedError("window sums are built for one population")
        win = np.asarray(windows, np.int64).reshape(-1, 2)
        if len(win) and (np.any(win[:, 0] > win[:, 1]) or np.any(win[:, 0] < 0) or
                         np.any(win[:, 1] > pan.V)):
            raise ValueError("windows must be site ranges [lo, hi) inside the matrix")
        if len(win) > 1 and np.any(win[1:, 0] < win[:-1, 1]):
            raise ValueError("windows must be sorted and non-overlapping")
        if tile not in (16, 32):
            raise ValueError("tile must be 16 or 32")
        if tile == 16 and pan.mode != _lib.MODE_HAP:
            raise NotImplementedError("16-site window tiles are built for haplotypes")
        self.pan = pan
        self.tile = int(tile)
        if not 0 <= int(group) <= 4096:
            raise ValueError("group must be in [0, 4096]")
        self.group = int(group)
        self.windows = win
        self.work = window_work(win, self.tile)
        k = win[:, 1] - win[:, 0]
        self.n_pairs = (k * (k - 1) // 2).astype(np.int64)
        self.d_win = torch.as_tensor(win.reshape(-1), device=pan.device)
        self.d_work = torch.as_tensor(self.work.reshape(-1), dtype=torch.int32, device=pan.device)
        self.ns = pan.n_stats

    def args(self, out, grid=0):
        pan = self.pan
        nw = len(self.windows)
        a = _lib.WindowArgs()
        a.d_panels = _lib.ptr(pan.panels).value
        a.d_site_stats = _lib.ptr(pan.site_stats).value
        a.d_mat = _lib.ptr(pan.mat).value
        a.ld = pan.ld
        a.d_rows = _lib.ptr(pan.d_rows).value
        a.d_rtab = _lib.ptr(pan.rtab).value
        a.d_windows = _lib.ptr(self.d_win).value if nw else None
        a.n_windows = nw
        a.d_work = _lib.ptr(self.d_work).value if len(self.work) else None
        a.n_work = len(self.work)
        a.d_out = _lib.ptr(out).value
        a.grid = int(grid)
        a.n_lo[0] = int(pan.n_lo[0])
        a.tile = self.tile
        a.group = self.group
        return a

    def run(self, kernel="fused", grid=0):
        torch = _torch()
        pan = self.pan
        nw = len(self.windows)
        out = torch.zeros((max(nw, 1), self.ns + 3, 3), dtype=torch.int64, device=pan.device)
        a = self.args(out, grid)
        # Window items are short (one population, often no missing calls): a
        # shared atomic counter per XCD would be claimed tens of millions of
        # times per second, so the window kernel keeps the static schedule,
        # under which the items of one window already run side by side.
        sched = None
        fn = _lib.lib().pgld_windows_fused if kernel == "fused" else _lib.lib().pgld_windows_plain
        _lib.check(fn(ctypes.byref(pan.desc), ctypes.byref(a), _lib.stream_handle()))
        res = _PendingWindows(out, self.ns, nw)
        res._sched = sched
        return res


    def capture(self, kernel="fused", grid=0, repack=False):
        """Record the whole job (optionally re-tallying and re-packing the
        panels first) as a HIP graph: one replay = clear + one persistent
        launch over every window's work items.  The output is rewritten by
        each replay."""
        return WindowGraph(self, kernel, grid, repack)


class WindowGraph:
    def __init__(self, plan, kernel, grid, repack):
        torch = _torch()
        pan = plan.pan
        self.plan, self.repack = plan, bool(repack)
        self.out = torch.zeros((max(len(plan.windows), 1), plan.ns + 3, 3), dtype=torch.int64,
                               device=pan.device)
        self.args = plan.args(self.out, grid)
        self.fn = (_lib.lib().pgld_windows_fused if kernel == "fused"
                   else _lib.lib().pgld_windows_plain)
        side = torch.cuda.Stream(device=pan.device)   # default priority
        side.wait_stream(torch.cuda.current_stream(pan.device))
        with torch.cuda.stream(side):
            self._body()
        torch.cuda.current_stream(pan.device).wait_stream(side)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self._body()

    def _body(self):
        pan = self.plan.pan
        if self.repack:
            pan.rescan()
            pan.repack()
        self.out.zero_()
        _lib.check(self.fn(ctypes.byref(pan.desc), ctypes.byref(self.args), _lib.stream_handle()))

    def replay(self):
        self.graph.replay()
        return _PendingWindows(self.out, self.plan.ns, len(self.plan.windows))


class _PendingWindows:
    def __init__(self, out, ns, nw):
        self.out, self.ns, self.nw = out, ns, nw

    def finish(self):
        sums, counts = decode_limbs(self.out.cpu().numpy(), self.ns + 2)
        sums, counts = sums[:self.nw], counts[:self.nw]
        return WindowResult(sums[:, 0].copy(), sums[:, 1].copy(), sums[:, 2].copy(),
                            sums[:, self.ns].copy(),
                            np.rint(sums[:, self.ns + 1]).astype(np.int64), counts.astype(np.int64))


def window_ld(mat, windows, pop_rows=None, mode="hap", kernel="fused", grid=0):
    """Per-window sums for one population.

    mat:      int8 CUDA tensor [n_rows, n_sites] (haplotypes 0/1/-1 or
              genotypes 0/1/2/-1), sites sorted by position
    windows:  [n_windows][2] site ranges [lo, hi), sorted, non-overlapping
              (see windows_from_positions)
    pop_rows: rows of the population (default: all rows)
    """
    if pop_rows is None:
        pop_rows = list(range(mat.shape[0]))
    pan = LDPanels(mat, [list(pop_rows)], mode)
    return WindowPlan(pan, windows).run(kernel, grid).finish()


def _torch():
    import torch

    return torch
