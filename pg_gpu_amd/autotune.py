"""Launch-shape autotuner of the fused binning kernel and the window kernel.

The fused kernel's tile is fixed by the MFMA (32 x 32 sites per wave); what
varies with the data is how the band's tiles are ordered and dealt out:

- ``square``: the side (in tiles) of the squares the band tile list is
  ordered in.  The waves of one XCD work inside one square at a time, so its
  operand panels (2 x square 32-site blocks) must stay in that XCD's L2; the
  best side depends on the band width in tiles and the panel bytes per block
  (the sample counts).
- ``dynamic``: atomic-counter tile claiming per XCD, or a static stride.
- ``grid``: resident workgroups (the kernel's default fills every CU at its
  occupancy).
- ``form``: every wave counts and evaluates its own tiles (0), or counting
  waves hand their tiles to statistics waves through LDS (1, for the traits
  that run two or more waves per SIMD; the library uses form 0 elsewhere).

``tune_binned`` times every candidate on the job itself (HIP events on the
launch stream, best of ``reps`` after one warm-up) and returns the fastest;
results are cached per job shape (mode, populations and their sizes, missing
data, bins, band width in tiles), so later jobs of the same shape skip the
search.  The bin sums do not depend on the choice (integer limb reduction).
"""
from __future__ import annotations

import numpy as np

from . import _device as dv
from .engine import TILE, BinnedPlan

SQUARES = (8, 16, 32, 64)

_cache = {}


def job_key(pan, positions, bins, max_dist, r2):
    """Shape of a binning job as the launch shape sees it."""
    pos = np.asarray(positions, np.float64)
    md = float(bins[-1]) if max_dist is None else float(max_dist)
    span = (pos[-1] - pos[0]) / max(len(pos) - 1, 1) if len(pos) > 1 else 1.0
    band_tiles = int(round(md / max(span, 1e-12) / TILE))
    return (int(pan.mode), int(pan.P), tuple(pan.pop_sizes), tuple(pan.has_missing),
            int(len(bins) - 1), int(np.log2(max(band_tiles, 1))), bool(r2))


def _time(plan, cfg, r2, reps):
    torch = dv.torch()
    stream = torch.cuda.current_stream()
    plan.run("fused", cfg["grid"], r2, cfg["dynamic"], cfg["form"]).finish()   # warm-up
    best = float("inf")
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        pend = plan.run("fused", cfg["grid"], r2, cfg["dynamic"], cfg["form"])
        b.record(stream)
        b.synchronize()
        pend.finish()
        best = min(best, a.elapsed_time(b))
    return best


def pc_form_available(pan):
    """Whether the library has the producer/consumer form for this panel set
    (one haplotype population with missing calls, two haplotype populations,
    one genotype population); elsewhere it runs form 0 whatever is asked."""
    if pan.mode == 0:
        return (pan.P == 1 and bool(pan.has_missing[0])) or pan.P == 2
    return pan.P == 1


def tune_binned(pan, positions, bins, max_dist=None, coords=None, r2=False, reps=3,
                squares=SQUARES, grids=(0,), forms=(0, 1), use_cache=True):
    """-> (best config dict(square, dynamic, grid, form), [(config, ms)...] measured).
    The best config is cached per ``job_key``; a cached job returns it with an
    empty measurement list."""
    key = job_key(pan, positions, bins, max_dist, r2)
    if use_cache and key in _cache:
        return dict(_cache[key]), []
    timings = []
    forms = [f for f in forms if f == 0 or pc_form_available(pan)] or [0]
    for sq in squares:
        plan = BinnedPlan(pan, positions, bins, max_dist, coords, square=sq)
        for dyn in (True, False):
            for g in grids:
                for form in forms:
                    cfg = dict(square=int(sq), dynamic=dyn, grid=int(g), form=int(form))
                    timings.append((cfg, _time(plan, cfg, r2, reps)))
    best = min(timings, key=lambda t: t[1])[0]
    _cache[key] = dict(best)
    return dict(best), timings


def clear_cache():
    _cache.clear()
    _win_cache.clear()


# --------------------------------------------------------------------------- windows

_win_cache = {}


def window_job_key(pan, windows):
    """Shape of a window job: mode, population size, missing data and the
    quartiles of the window lengths (log2)."""
    w = np.asarray(windows, np.int64).reshape(-1, 2)
    k = np.maximum(w[:, 1] - w[:, 0], 1)
    q = tuple(int(np.log2(x)) for x in np.percentile(k, [25, 50, 75])) if len(k) else ()
    return (int(pan.mode), tuple(pan.pop_sizes), tuple(pan.has_missing), q)


WINDOW_TIE = 0.02


def pick_window_launch(timings, tie=WINDOW_TIE):
    """The launch to keep from [(dict(tile, group), ms)...]: among those
    within ``tie`` (relative) of the fastest, the one with the smallest
    group.  An XCD's waves work on about (waves x group) neighbouring items
    at a time: short turns keep those items' panels in the XCD's L2, long
    turns flush fewer atomics but re-read panels from HBM (on the ragged-window
    workload, 64 items per turn run as fast as 16 and fetch 5x the bytes)."""
    fastest = min(t for _, t in timings)
    near = [(c, t) for c, t in timings if t <= fastest * (1.0 + tie)]
    return min(near, key=lambda ct: (ct[0]["group"], ct[1]))[0]


def tune_windows(pan, windows, reps=3, tiles=(32, 16), groups=(16, 32, 64), use_cache=True):
    """-> (best dict(tile, group), [(config, ms)...]): times the window kernel
    with 32 x 32 and (haplotypes) 16 x 16 MFMA tiles and several work-item
    group sizes (items per wave turn) on the job, keeps the launch that
    ``pick_window_launch`` names, cached per ``window_job_key``."""
    from .windows import WindowPlan

    key = window_job_key(pan, windows)
    if use_cache and key in _win_cache:
        return dict(_win_cache[key]), []
    torch = dv.torch()
    stream = torch.cuda.current_stream()
    timings = []
    for tile in tiles:
        if tile == 16 and pan.mode != 0:
            continue
        for group in groups:
            plan = WindowPlan(pan, windows, tile=tile, group=group)
            plan.run().finish()                                   # warm-up
            best = float("inf")
            for _ in range(reps):
                a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a.record(stream)
                pend = plan.run()
                b.record(stream)
                b.synchronize()
                pend.finish()
                best = min(best, a.elapsed_time(b))
            timings.append((dict(tile=int(tile), group=int(group)), best))
    best = pick_window_launch(timings)
    _win_cache[key] = dict(best)
    return dict(best), timings
