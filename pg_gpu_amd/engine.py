"""Host orchestration of the binned two-locus LD path on one GPU.

``LDPanels`` owns the per-matrix state: the site tallies and the packed MFMA
operand panels of the selected populations (built once, reused for every
bin set).  ``LDPanels.binned`` runs the fused kernel (or the plain
comparator) for one bin set and returns exact bin sums.

The device work is all in libpgld.so; this module only sizes buffers, builds
the band's tile list and decodes the fixed-point result.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import MODE_GENO, MODE_HAP

TILE = 32  # sites per tile side: one 32x32 MFMA block per wave
SUPER = 16  # tiles per side of the L2-sharing groups of band_tiles


def _torch():
    import torch

    return torch


def reciprocal_table(kmax):
    """4 doubles per n in [0, kmax]: 1/(n)_1 .. 1/(n)_4, 0 where (n)_k <= 0."""
    tab = np.zeros((kmax + 1, 4), np.float64)
    for n in range(kmax + 1):
        ff = 1
        for k in range(4):
            ff *= (n - k)
            tab[n, k] = 1.0 / ff if ff > 0 else 0.0
    return tab.ravel()


def band_tiles(pos, max_dist, square=None, with_pairs=False):
    """(row tile, col tile) pairs of TILE sites covering every i<j pair with
    pos[j] <= pos[i] + max_dist (pg_gpu/ld_pipeline.py:85-88), in squares of
    ``square`` x ``square`` tiles (default SUPER): squares in row-major order,
    tiles row-major inside a square.  The waves running at the same time on
    one XCD then share ~2 * square operand panels in that XCD's L2.  Built by
    the library's host planner (pgld_band_tiles, one sweep over the sites);
    ``with_pairs`` also returns the number of pairs within the cutoff."""
    import ctypes

    sq = SUPER if square is None else int(square)
    if sq < 1:
        raise ValueError("square must be positive")
    pos = np.ascontiguousarray(pos, np.float64)
    L = _lib.lib()
    n_tiles, n_pairs = ctypes.c_int64(0), ctypes.c_int64(0)
    pp = pos.ctypes.data_as(ctypes.c_void_p)
    _lib.check(L.pgld_band_tiles(pp, len(pos), float(max_dist), sq, None, 0,
                                 ctypes.byref(n_tiles), None))
    out = np.empty((n_tiles.value, 2), np.int32)
    _lib.check(L.pgld_band_tiles(pp, len(pos), float(max_dist), sq,
                                 out.ctypes.data_as(ctypes.c_void_p), n_tiles.value,
                                 ctypes.byref(n_tiles), ctypes.byref(n_pairs) if with_pairs
                                 else None))
    return (out, n_pairs.value) if with_pairs else out


def decode_limbs(raw, n_stats):
    """int64 [n_bins, n_stats+1, 3] limbs -> (float64 sums [n_bins, n_stats], int64 counts)."""
    raw = np.asarray(raw, np.int64).reshape(-1, n_stats + 1, 3)
    nb = raw.shape[0]
    # a producer/consumer launch whose mailbox handshake timed out adds 2^62
    # to bin 0's pair count: no real count comes near it
    if nb and raw[0, n_stats, 0] >= (1 << 60):
        raise RuntimeError("fused kernel: a mailbox wait timed out; the result is invalid")
    # exact integer sum of the limbs, one correctly rounded division each
    # (Python integers in an object array: the loop runs in C)
    r = raw[:, :n_stats, :].astype(object)
    v = (r[..., 2] << 64) + (r[..., 1] << 32) + r[..., 0]
    sums = (v / (1 << 72)).astype(np.float64) if v.size else np.zeros((nb, n_stats))
    return sums, raw[:, n_stats, 0].copy()


class PendingResult:
    """A launched binning job; ``finish()`` waits for it and decodes the limbs."""

    def __init__(self, out, n_stats, n_pairs, n_tiles, r2=False):
        self.out, self.n_stats, self.n_pairs, self.n_tiles = out, n_stats, n_pairs, n_tiles
        self.r2 = r2
        self._res = None

    def finish(self):
        if self._res is None:
            width = self.n_stats + (2 if self.r2 else 0)
            sums, counts = decode_limbs(self.out.cpu().numpy(), width)
            res = BinnedResult(sums[:, :self.n_stats].copy(), counts, self.n_pairs, self.n_tiles)
            if self.r2:
                res.r2_sum = sums[:, self.n_stats].copy()
                res.r2_count = np.rint(sums[:, self.n_stats + 1]).astype(np.int64)
            self._res = res
        return self._res


class BinnedPlan:
    """Device copies of the positions, bin edges and band tile list of one job."""

    def __init__(self, pan, positions, bins, max_dist=None, coords=None, tiles=None,
                 square=None):
        torch = _torch()
        self.pan = pan
        self.square = SUPER if square is None else int(square)
        bins = np.asarray(bins, np.float64)
        nb = len(bins) - 1
        if nb < 1 or nb > _lib.MAX_BINS:
            raise ValueError("need between 2 and 65 bin edges")
        if np.any(np.diff(bins) < 0) or bins[-1] <= bins[0]:
            raise ValueError("bins must be increasing")
        pos_np = np.asarray(positions.cpu() if hasattr(positions, "cpu") else positions,
                            np.float64)
        if len(pos_np) != pan.V:
            raise ValueError("positions length differs from the number of sites")
        if pan.V > 1 and np.any(np.diff(pos_np) < 0):
            raise ValueError("positions must be sorted")
        self.max_dist = float(bins[-1]) if max_dist is None else float(max_dist)
        self.nb = nb
        self.d_pos = torch.as_tensor(pos_np, device=pan.device)
        self.coords = coords
        if coords is not None:
            co_np = np.asarray(coords.cpu() if hasattr(coords, "cpu") else coords, np.float64)
            if len(co_np) != pan.V:
                raise ValueError("coords length differs from the number of sites")
            # tiles are classified by their corner pairs, which bound every
            # pair's distance only when the coordinate grows with position
            if pan.V > 1 and np.any(np.diff(co_np) < 0):
                raise ValueError("coords must be non-decreasing along the sorted positions")
        self.d_coord = (torch.as_tensor(co_np, device=pan.device)
                        if coords is not None else self.d_pos)
        self.d_edges = torch.as_tensor(bins, device=pan.device)
        self._pos_np = np.ascontiguousarray(pos_np)
        self._n_pairs = None
        self._full_band = tiles is None
        self.tiles = (band_tiles(pos_np, self.max_dist, self.square) if tiles is None
                      else np.asarray(tiles))
        self.d_tiles = torch.as_tensor(self.tiles.reshape(-1), dtype=torch.int32,
                                       device=pan.device)
        self.ns = pan.n_stats
        if self.ns < 0:
            raise NotImplementedError("statistic set not available for this mode / n_pops")

    @property
    def n_pairs(self):
        """Pairs within the cutoff (with an explicit tile list: those whose
        anchor lies in one of its row tiles); counted on first use."""
        if self._n_pairs is None:
            pos_np, V = self._pos_np, self.pan.V
            if self._full_band:
                import ctypes

                n_t, n_p = ctypes.c_int64(0), ctypes.c_int64(0)
                _lib.check(_lib.lib().pgld_band_tiles(
                    pos_np.ctypes.data_as(ctypes.c_void_p), V, self.max_dist, 1, None, 0,
                    ctypes.byref(n_t), ctypes.byref(n_p)))
                self._n_pairs = int(n_p.value)
            else:
                jmax = np.searchsorted(pos_np, pos_np + self.max_dist, side="right")
                per = np.maximum(0, jmax - np.arange(V) - 1)
                rows = (np.unique(self.tiles[:, 0]) if len(self.tiles)
                        else np.zeros(0, np.int64))
                sel = np.zeros((V + TILE - 1) // TILE, bool)
                sel[rows] = True
                self._n_pairs = int(per[np.repeat(sel, TILE)[:V]].sum())
        return self._n_pairs

    def args(self, out, grid=0, r2=False, form=0):
        pan = self.pan
        a = _lib.BinnedArgs()
        a.d_panels = _lib.ptr(pan.panels).value if pan.panels is not None else None
        a.d_site_stats = _lib.ptr(pan.site_stats).value
        a.d_mat = _lib.ptr(pan.mat).value
        a.ld = pan.ld
        a.d_rows = _lib.ptr(pan.d_rows).value
        for p in range(pan.P + 1):
            a.row_offset[p] = int(pan.row_offsets[p])
        a.d_pos = _lib.ptr(self.d_pos).value
        a.d_coord = _lib.ptr(self.d_coord).value
        a.coord_is_pos = int(self.coords is None)
        a.max_dist = self.max_dist
        a.d_edges = _lib.ptr(self.d_edges).value
        a.n_bins = self.nb
        a.d_rtab = _lib.ptr(pan.rtab).value
        a.d_tiles = _lib.ptr(self.d_tiles).value if len(self.tiles) else None
        a.n_tiles = len(self.tiles)
        a.d_out = _lib.ptr(out).value
        a.grid = int(grid)
        for p in range(pan.P):
            a.n_lo[p] = int(pan.n_lo[p])
        a.with_r2 = int(bool(r2))
        a.form = int(form)
        return a

    def run(self, kernel="fused", grid=0, r2=False, dynamic=True, form=0):
        """Launch one binning pass.  With ``r2`` (one population) the bins also
        receive the sum of r^2 and the number of pairs where r^2 is defined.
        ``dynamic`` selects the atomic-counter tile schedule of the fused
        kernel (False: static stride) and ``form`` its producer/consumer form
        (1) or per-wave form (0); the result is the same either way."""
        torch = _torch()
        pan = self.pan
        if r2 and pan.P != 1:
            raise NotImplementedError("r2 sums are built for one population")
        width = self.ns + (2 if r2 else 0)
        out = torch.zeros((self.nb, width + 1, 3), dtype=torch.int64, device=pan.device)
        a = self.args(out, grid, r2, form)
        sched = None
        if dynamic:
            sched = torch.zeros(_lib.SCHED_WORDS, dtype=torch.int32, device=pan.device)
            a.d_sched = _lib.ptr(sched).value
        fn = _lib.lib().pgld_binned_fused if kernel == "fused" else _lib.lib().pgld_binned_plain
        _lib.check(fn(ctypes.byref(pan.desc), ctypes.byref(a), _lib.stream_handle()))
        res = PendingResult(out, self.ns, lambda: self.n_pairs, len(self.tiles), r2)
        res._sched = sched  # kept alive until the launch has been waited for
        return res

    def capture(self, kernel="fused", grid=0, r2=False, dynamic=True, repack=False, form=0):
        """Record one binning pass (optionally preceded by re-packing the
        panels) as a HIP graph; ``GraphRun.replay()`` then relaunches it with
        one call.  Buffers are cleared by fill kernels inside the graph."""
        return GraphRun(self, kernel, grid, r2, dynamic, repack, form)


class GraphRun:
    """A captured binning pass: output and scheduler buffers are owned by the
    graph and rewritten by every replay, so a replay's result must be read
    (``finish()``) before the next replay."""

    def __init__(self, plan, kernel, grid, r2, dynamic, repack, form=0):
        torch = _torch()
        pan = plan.pan
        if r2 and pan.P != 1:
            raise NotImplementedError("r2 sums are built for one population")
        self.plan, self.r2, self.repack = plan, bool(r2), bool(repack)
        width = plan.ns + (2 if r2 else 0)
        self.out = torch.zeros((plan.nb, width + 1, 3), dtype=torch.int64, device=pan.device)
        self.sched = (torch.zeros(_lib.SCHED_WORDS, dtype=torch.int32, device=pan.device)
                      if dynamic else None)
        self.args = plan.args(self.out, grid, r2, form)
        if self.sched is not None:
            self.args.d_sched = _lib.ptr(self.sched).value
        self.fn = _lib.lib().pgld_binned_fused if kernel == "fused" else \
            _lib.lib().pgld_binned_plain
        side = torch.cuda.Stream(device=pan.device)   # default priority
        side.wait_stream(torch.cuda.current_stream(pan.device))
        with torch.cuda.stream(side):                  # warm-up outside the graph
            self._body()
        torch.cuda.current_stream(pan.device).wait_stream(side)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self._body()

    def _body(self):
        pan = self.plan.pan
        if self.repack:
            pan.repack()
        self.out.zero_()
        if self.sched is not None:
            self.sched.zero_()
        _lib.check(self.fn(ctypes.byref(pan.desc), ctypes.byref(self.args),
                           _lib.stream_handle()))

    def replay(self):
        self.graph.replay()
        return PendingResult(self.out, self.plan.ns, lambda: self.plan.n_pairs,
                             len(self.plan.tiles),
                             self.r2)


@dataclass
class BinnedResult:
    sums: np.ndarray        # [n_bins, n_stats] exact-summed FP64
    counts: np.ndarray      # [n_bins] pairs per bin
    _n_pairs: object        # pairs within the cutoff: an int, or a callable giving it
    n_tiles: int
    r2_sum: np.ndarray = None    # [n_bins] sum of r^2 where defined (with r2=True)
    r2_count: np.ndarray = None  # [n_bins] pairs with r^2 defined

    @property
    def n_pairs(self):
        """Enumerated pairs (within the cutoff), counted on first use."""
        if callable(self._n_pairs):
            self._n_pairs = int(self._n_pairs())
        return self._n_pairs

    @property
    def mean_r2(self):
        """Mean r^2 per bin over pairs where it is defined (NaN if none)."""
        if self.r2_sum is None:
            return None
        with np.errstate(invalid="ignore", divide="ignore"):
            return np.where(self.r2_count > 0, self.r2_sum / np.maximum(self.r2_count, 1), np.nan)


class LDPanels:
    """Packed per-population operand panels of one int8 matrix on one GPU.

    mat:      torch int8 tensor [n_rows, n_sites] on a CUDA device (sample-major)
    pop_rows: list (1..4) of row-index lists, one per population
    mode:     "hap" (values 0/1/-1) or "geno" (0/1/2/-1)
    pack:     False keeps only the per-site tallies (heterozygosity needs no
              panels)

    Every launch goes to torch's current stream at the time of the call, the
    stream the buffers are allocated, cleared and read on.
    """

    def __init__(self, mat, pop_rows, mode="hap", pack=True):
        torch = _torch()
        if mat.dtype != torch.int8 or mat.dim() != 2 or not mat.is_cuda:
            raise ValueError("mat must be a 2-D int8 CUDA tensor")
        if not 1 <= len(pop_rows) <= 4:
            raise ValueError("1-4 populations supported")
        self.mode = MODE_HAP if mode == "hap" else MODE_GENO
        self.mat = mat if mat.stride(1) == 1 else mat.contiguous()
        self.device = mat.device
        self.n_rows, self.V = self.mat.shape
        self.ld = self.mat.stride(0)
        self.P = len(pop_rows)
        sizes = [len(r) for r in pop_rows]
        for r in pop_rows:
            if len(r) and (min(r) < 0 or max(r) >= self.n_rows):
                raise ValueError("population row index out of range")
        self.pop_sizes = sizes
        self.row_offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
        flat = np.concatenate([np.asarray(r, np.int32) for r in pop_rows]) if sum(sizes) else \
            np.zeros(0, np.int32)
        self.d_rows = torch.as_tensor(flat, dtype=torch.int32, device=self.device)
        L = _lib.lib()
        desc = self._desc([0] * self.P)
        offs = _lib.int32_array(self.row_offsets)
        self.site_stats = torch.zeros((self.P, 3, self.V), dtype=torch.int32, device=self.device)
        flags = torch.zeros(1, dtype=torch.int32, device=self.device)
        self._flags = flags
        h = _lib.stream_handle()
        _lib.check(L.pgld_scan_sites(_lib.ptr(self.mat), self.n_rows, self.V, self.ld,
                                     _lib.ptr(self.d_rows), offs, ctypes.byref(desc),
                                     _lib.ptr(self.site_stats), _lib.ptr(flags), h))
        miss = self.site_stats[:, 2, :].sum(dim=1).cpu().numpy() if self.V else np.zeros(self.P)
        if int(flags.item()):
            alph = "0, 1 or -1" if self.mode == MODE_HAP else "0, 1, 2 or -1"
            raise ValueError(f"matrix values must be {alph}")
        self.has_missing = [int(m > 0) for m in miss]
        # lower bound of any pair's valid-sample count: size - 2 * (max missing at a site)
        mx = (self.site_stats[:, 2, :].max(dim=1).values.cpu().numpy() if self.V
              else np.zeros(self.P, np.int64))
        self.n_lo = [max(0, int(k) - 2 * int(m)) for k, m in zip(sizes, mx)]
        self.desc = self._desc(self.has_missing)
        kmax = max(sizes) if sizes else 0
        self.rtab = torch.as_tensor(reciprocal_table(kmax), device=self.device)
        self.panels = None
        if not pack:
            return
        nbytes = L.pgld_panel_bytes(ctypes.byref(self.desc))
        if nbytes < 0:
            _lib.check(_lib.PGLD_EINVAL)
        self.panels = torch.zeros(max(int(nbytes), 16), dtype=torch.int8, device=self.device)
        _lib.check(L.pgld_pack(_lib.ptr(self.mat), self.n_rows, self.V, self.ld,
                               _lib.ptr(self.d_rows), offs, ctypes.byref(self.desc),
                               _lib.ptr(self.panels), h))

    def _desc(self, missing):
        d = _lib.PanelDesc()
        d.mode = self.mode
        d.n_pops = self.P
        d.n_sites = self.V
        for p in range(self.P):
            d.pop_size[p] = self.pop_sizes[p]
            d.has_missing[p] = missing[p]
        return d

    @property
    def n_stats(self):
        return int(_lib.lib().pgld_n_stats(self.mode, self.P))

    def rescan(self):
        """Re-run the site-tally kernel into the existing tally buffer (no
        allocation, no host synchronisation; the value check's flag word is
        cleared by a fill kernel and not read back)."""
        if self.V == 0:
            return
        self._flags.zero_()
        offs = _lib.int32_array(self.row_offsets)
        _lib.check(_lib.lib().pgld_scan_sites(_lib.ptr(self.mat), self.n_rows, self.V, self.ld,
                                              _lib.ptr(self.d_rows), offs,
                                              ctypes.byref(self.desc), _lib.ptr(self.site_stats),
                                              _lib.ptr(self._flags), _lib.stream_handle()))

    def repack(self):
        """Re-run the pack kernel into the existing panel buffer (no allocation)."""
        if self.panels is None:
            raise ValueError("these tallies were built without panels (pack=False)")
        offs = _lib.int32_array(self.row_offsets)
        _lib.check(_lib.lib().pgld_pack(_lib.ptr(self.mat), self.n_rows, self.V, self.ld,
                                        _lib.ptr(self.d_rows), offs, ctypes.byref(self.desc),
                                        _lib.ptr(self.panels), _lib.stream_handle()))

    def plan(self, positions, bins, max_dist=None, coords=None, tiles=None, square=None):
        """Device-side plan of one binning job (positions, edges, band tiles)."""
        return BinnedPlan(self, positions, bins, max_dist, coords, tiles, square)

    def binned(self, positions, bins, max_dist=None, coords=None, kernel="fused", grid=0,
               tiles=None, sync=True, r2=False, form=0):
        """Bin sums of every LD statistic over pairs within ``max_dist``.

        positions: sorted site positions (the cutoff coordinate), numpy or tensor
        bins:      increasing bin edges; pair distance d goes to bin b when
                   bins[b] <= d < bins[b+1] (np.digitize semantics)
        coords:    optional binning coordinate (e.g. Morgans); default positions
        r2:        one population only: also sum r^2 per bin (see BinnedResult.mean_r2)
        With ``sync=False`` the launch is left in flight and a PendingResult is
        returned; call ``.finish()`` on it to wait and decode.
        """
        pend = self.plan(positions, bins, max_dist, coords, tiles).run(kernel, grid, r2,
                                                                       form=form)
        return pend.finish() if sync else pend

    def count_block(self, pop, i0, ni, j0, nj):
        """Dense int32 counts [q, ni, nj] of one population (MFMA path)."""
        torch = _torch()
        q = 4 if self.mode == MODE_HAP else 9
        out = torch.zeros((q, ni, nj), dtype=torch.int32, device=self.device)
        _lib.check(_lib.lib().pgld_count_block(ctypes.byref(self.desc), _lib.ptr(self.panels),
                                                _lib.ptr(self.site_stats), pop, i0, ni, j0, nj,
                                                _lib.ptr(out), _lib.stream_handle()))
        return out
