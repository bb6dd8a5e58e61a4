"""Benchmark of the binned two-locus LD path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c1]

Workload (default c1): 2 populations x 500 haplotypes, 200,000 biallelic SNPs
on 20 Mb (founder-copying model, Balding-Nichols F=0.05, 5% missing calls),
every site pair within 200 kb, the 15 two-population DD/Dz/pi2 statistics
summed into 20 log-spaced distance bins (100 bp .. 200 kb).

A step = one pass of the hot path over the whole input already resident in
HBM: the per-site tallies, packing the int8 matrix into MFMA panels, and the
fused count/estimator/bin kernel.  The band's tile list and the reciprocal
tables are a per-dataset plan built once, outside the timed region; the
"end_to_end" object times the whole call including them.  Timing uses HIP events on the
stream the kernels run on; the same run also times the plain per-pair FP64
kernel, the CPU oracle on a bounded anchor prefix (before the GPU is touched,
one process per core), and compares the fused result with the oracle on that
prefix.  One JSON line goes to stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import math
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_I8_TOPS = 5000.0      # MI355X dense int8 MFMA, TOP/s
PEAK_FP4_TOPS = 10100.0    # MI355X dense FP4 MFMA (the counting stage's operands), TOP/s
PEAK_FP64_TFLOPS = 78.6    # MI355X FP64 (vector and matrix)
PEAK_HBM_GBS = 8000.0

R2_WORKLOADS = ("c0", "c2")   # mean r^2 per bin is part of these workloads

WORKLOADS = {
    "c0": "c0: 1 pop x 200 haplotypes, 5k SNPs / 1 Mb, pairs <= 100 kb, 10 log bins, "
          "DD/Dz/pi2 + mean r^2",
    "c1": "c1: 2 pops x 500 haplotypes, 200k SNPs / 20 Mb, 5% missing, pairs <= 200 kb, "
          "20 log bins 100 bp-200 kb, 15 stats",
    "c2": "c2: 1 pop x 1000 diploid genotypes, 100k SNPs / 10 Mb, 2% missing, pairs <= 100 kb, "
          "15 log bins, DD/Dz/pi2 + mean dosage r^2",
    "c3": "c3: 300 haplotypes, 5,000 x 50 kb windows, SNPs per window clip(Gamma(0.8, 500), "
          "20, 4000), all pairs per window, DD/Dz/pi2 + mean r^2 per window",
    "c4": "c4: 4 pops x 40/120/300/800 haplotypes, 500k SNPs / 50 Mb, 3% missing, 30% rare "
          "sites, pairs <= 500 kb, 25 log bins, 105 stats",
    "g4": "g4: the c4 haplotypes paired into 4 pops x 20/60/150/400 diploid genotypes "
          "(not a headline configuration: the four-population genotype kernel)",
}


def traffic_from_profile(name):
    """HBM bytes per launch recorded by the rocprofv3 FETCH_SIZE / WRITE_SIZE
    passes in profiles/<name>_hbm_traffic.md (None if not profiled)."""
    path = os.path.join(ROOT, "profiles", f"{name}_hbm_traffic.md")
    try:
        for line in open(path):
            if "traffic_bytes_per_launch:" in line:
                return int(line.split(":")[1])
    except OSError:
        return None
    return None


def fp64_flops_from_profile(name):
    """FP64 vector FLOPs per launch of the kernel ((2 FMA + ADD + MUL) x 64
    lanes, from the rocprofv3 SQ_INSTS_VALU_*_F64 pass in
    profiles/<name>_fp64_valu.md; None if not profiled).  The count is fixed
    by the workload, so it pairs with the live kernel time."""
    path = os.path.join(ROOT, "profiles", f"{name}_fp64_valu.md")
    try:
        for line in open(path):
            if line.startswith("- fp64_flops_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        return None
    return None


def utilisation_from_profile(name):
    """MFMA / VALU utilisation of the fused kernel recorded by the rocprofv3
    counter pass in profiles/<name>_utilisation.md (None if not profiled)."""
    path = os.path.join(ROOT, "profiles", f"{name}_utilisation.md")
    out = {}
    try:
        for line in open(path):
            for key in ("mfma_util_percent", "valu_busy_percent", "mfma_ops_per_launch",
                        "valu_instructions_per_launch"):
                if line.startswith(f"- {key}:"):
                    out[key] = float(line.split(":")[1])
    except OSError:
        return None
    return out or None


def kernel_name(pan, geno, windows, r2, form):
    """The kernel instantiation the library dispatches for this job (the
    traits class: one haplotype population without missing calls has its own,
    one genotype population keeps its moment accumulators in the per-wave
    form)."""
    tail = ", r2" if r2 else ""
    nomiss = pan.P == 1 and not bool(pan.has_missing[0])
    if windows:
        t = "GenoTraits<1, r2>" if geno else ("HapNMRawTraits<r2>" if nomiss else "HapTraits<1, r2>")
        return f"window_kernel<{t}>"
    if geno:
        t = (f"GenoRawTraits<{'r2' if r2 else ''}>" if pan.P == 1 and form != 1
             else f"GenoTraits<{pan.P}{tail}>")
    else:
        t = f"HapNMRawTraits<{'r2' if r2 else ''}>" if nomiss else f"HapTraits<{pan.P}{tail}>"
    return f"fused_kernel<{t}>"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# --------------------------------------------------------------------------- CPU baseline

_W = None


def _cpu_worker(block):
    """The reference's form: chunked counts, FP64 estimators, np.add.at."""
    from oracle import ld_oracle as o

    w = _W
    t0 = time.perf_counter()
    sums, counts, n = o.binned_sums(w.mat, w.positions, w.pop_rows, w.bins, w.max_dist,
                                    chunk=100_000, exact=False, anchors=block, mode=w.mode,
                                    r2=w.name in R2_WORKLOADS)
    return sums, counts, n, None, time.perf_counter() - t0


def _cpu_worker_exact(block):
    """Correctly rounded bin sums and |term| sums (precision reference)."""
    from oracle import ld_oracle as o

    w = _W
    sums, counts, n, absr = o.binned_sums(w.mat, w.positions, w.pop_rows, w.bins, w.max_dist,
                                          chunk=100_000, with_abs=True, anchors=block,
                                          mode=w.mode, r2=w.name in R2_WORKLOADS)
    return sums, counts, n, absr, 0.0


def _cpu_worker_windows(block):
    from oracle import ld_oracle as o

    w = _W
    t0 = time.perf_counter()
    sums, counts = o.window_sums(w.mat, w.windows[block[0]:block[1]], mode=w.mode)
    return sums, counts, int(counts.sum()), None, time.perf_counter() - t0


# worker processes: every host core, up to the 16-core CPU share a GPU of an
# MI355X node gets (os.cpu_count() on a shared box reports the whole node)
CPU_SHARE = 16


def cpu_baseline(w, target_s=15.0, max_workers=CPU_SHARE):
    """Oracle (NumPy restatement of the reference estimators, np.add.at
    binning) timed on a prefix of the work (anchor sites, or windows), one
    process per core over contiguous blocks; then, untimed, correctly rounded
    sums over the first quarter of that prefix for the precision check."""
    global _W
    _W = w
    cores = max(1, min(os.cpu_count() or 1, max_workers))
    ctx = mp.get_context("fork")

    def pool_rate(worker, blocks):
        # aggregate pairs/s of all workers on small blocks (the rate of one
        # process alone overstates what `cores` processes reach together)
        t0 = time.perf_counter()
        with ctx.Pool(cores) as pool:
            parts = pool.map(worker, blocks)
        return sum(p[2] for p in parts) / max(time.perf_counter() - t0, 1e-6)

    if w.windows is not None:
        k = w.windows[:, 1] - w.windows[:, 0]
        cum = np.cumsum(k * (k - 1) // 2)
        rate = pool_rate(_cpu_worker_windows, [(2 * q, 2 * q + 2) for q in range(cores)])
        X = int(np.searchsorted(cum, rate * target_s)) + 1
        X = max(cores, min(len(k), X))
        edges = np.linspace(0, X, cores + 1).astype(int)
        blocks = [(int(edges[q]), int(edges[q + 1])) for q in range(cores)]
        log(f"cpu baseline: {cores} workers, windows [0, {X}) ~ {int(cum[X - 1]):,} pairs")
        worker = _cpu_worker_windows
    else:
        rate = pool_rate(_cpu_worker, [(64 * q, 64 * q + 64) for q in range(cores)])
        jmax = np.searchsorted(w.positions, w.positions + w.max_dist, side="right")
        per = np.maximum(0, jmax - np.arange(len(w.positions)) - 1)
        cum = np.cumsum(per)
        want = rate * target_s
        X = int(np.searchsorted(cum, want)) + 1
        X = max(64 * cores, min(len(w.positions), (X // (64 * cores)) * 64 * cores))
        edges = np.linspace(0, X, cores + 1).astype(int)
        blocks = [(int(edges[q]), int(edges[q + 1])) for q in range(cores)]
        log(f"cpu baseline: {cores} workers, anchors [0, {X}) ~ {int(cum[X - 1]):,} pairs")
        worker = _cpu_worker
    t0 = time.perf_counter()
    with ctx.Pool(cores) as pool:
        parts = pool.map(worker, blocks)
    wall = time.perf_counter() - t0
    n_pairs = sum(p[2] for p in parts)
    res = dict(value=n_pairs / wall, cores=cores, cpu_count=os.cpu_count(), X=X,
               n_pairs=n_pairs, wall=wall)
    if w.windows is not None:
        res.update(sums=np.concatenate([p[0] for p in parts]),
                   counts=np.concatenate([p[1] for p in parts]), abs=None, X_exact=X,
                   n_pairs_exact=n_pairs)
        return res
    # precision reference: exact sums over the first quarter of the prefix
    Xe = max(64 * cores, (X // 4) // (64 * cores) * 64 * cores)
    edges = np.linspace(0, Xe, cores + 1).astype(int)
    with ctx.Pool(cores) as pool:
        parts = pool.map(_cpu_worker_exact, [(int(edges[q]), int(edges[q + 1]))
                                             for q in range(cores)])
    nb, ns = w.bins.size - 1, parts[0][0].shape[1]
    res.update(sums=np.array([[math.fsum(p[0][b, s] for p in parts) for s in range(ns)]
                              for b in range(nb)]),
               counts=sum(p[1] for p in parts), abs=sum(p[3] for p in parts), X_exact=Xe,
               n_pairs_exact=sum(p[2] for p in parts))
    return res


# --------------------------------------------------------------------------- GPU


def counting_stage(pan, plan, launch, steps, n_pairs, ops_per_pair, nprod, w=None):
    """The counting stage alone (pgld_count_band: the fused kernel's traits,
    tiles and schedule with a checksum for an epilogue), timed with HIP
    events on the launch stream, its checksum compared with the oracle's
    closed form when the CPU leg runs."""
    import ctypes

    import torch

    from pg_gpu_amd import _lib

    stream = torch.cuda.current_stream()
    dummy = torch.zeros((1, 1, 3), dtype=torch.int64, device=pan.device)
    a = plan.args(dummy, launch["grid"])
    sched = torch.zeros(_lib.SCHED_WORDS, dtype=torch.int32, device=pan.device)
    sums = torch.zeros((pan.P, 4), dtype=torch.int64, device=pan.device)
    if launch.get("dynamic", True):
        a.d_sched = _lib.ptr(sched).value

    def once(ev=None):
        sched.zero_()
        sums.zero_()
        if ev is not None:
            ev[0].record(stream)
        _lib.check(_lib.lib().pgld_count_band(ctypes.byref(pan.desc), ctypes.byref(a),
                                              _lib.ptr(sums), _lib.stream_handle()))
        if ev is not None:
            ev[1].record(stream)

    for _ in range(2):
        once()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(steps)]
    for ev in evs:
        once(ev)
    torch.cuda.synchronize()
    ms = float(np.mean([x.elapsed_time(y) for x, y in evs]))
    got = sums.cpu().numpy()
    eq = None
    if w is not None:
        from oracle import ld_oracle as o

        eq = bool(np.array_equal(got, o.band_count_checksum(w.mat, w.pop_rows, plan.tiles,
                                                            marginals=pan.P > 1)))
    kpad = [64 * ((k + 63) // 64) for k in pan.pop_sizes]
    executed = 2 * sum(q * k for q, k in zip(nprod, kpad)) * 1024 * len(plan.tiles)
    ach = ops_per_pair * n_pairs / (ms * 1e-3) / 1e12
    return {"kernel": f"count_band_kernel<HapTraits<{pan.P}>>", "kernel_ms": ms,
            "achieved": ach, "unit": "TOP/s (in-band pairs x 2 g sum K)",
            "frac_of_fp4_peak": ach / PEAK_FP4_TOPS, "frac_of_i8_peak": ach / PEAK_I8_TOPS,
            "executed_ops": executed,
            "executed_frac_of_fp4_peak": executed / (ms * 1e-3) / 1e12 / PEAK_FP4_TOPS,
            "checksum": got.tolist(), "checksum_equal": eq}


def end_to_end(mat, w, r2, launch, n_pairs, reps=3):
    """Wall time of the whole call from the resident int8 matrix: site
    tallies (with the host read of the missing-data flags), packing, the
    host-side band tile list and its upload, the kernel, and the download
    and decoding of the limbs -- the median of ``reps``."""
    import torch

    from pg_gpu_amd.engine import LDPanels

    times = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pan = LDPanels(mat, w.pop_rows, w.mode)
        plan = pan.plan(w.positions, w.bins, w.max_dist, square=launch.get("square"))
        plan.run("fused", launch.get("grid", 0), r2, launch.get("dynamic", True),
                 launch.get("form", 0)).finish()
        times.append(time.perf_counter() - t0)
        del pan, plan
    ms = float(np.median(times)) * 1e3
    return {"ms": ms, "value": n_pairs / (ms * 1e-3), "unit": "site-pairs/s", "reps": reps,
            "includes": "tallies + flag read, pack, band tile list (host) + upload, kernel, "
                        "limb download and decode"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c1")
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-count-stage", action="store_true",
                    help="skip the counting-stage-only kernel leg")
    ap.add_argument("--no-tune", action="store_true",
                    help="use the default launch shape instead of the autotuned one")
    ap.add_argument("--form", type=int, default=0,
                    help="fused kernel form with --no-tune: 0 per-wave, 1 producer/consumer")
    ap.add_argument("--window-group", type=int, default=0,
                    help="window kernel work items per wave turn with --no-tune (0 = default)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a device-to-device copy of this many MiB three times "
                         "(the known-size kernel that scales the HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(the multi-GPU sharding is a design note, not built)"}))
        sys.exit(2)

    if args.copy_probe_mib:
        import torch

        n = args.copy_probe_mib << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n, "copies": 3}))
        return

    from pg_gpu_amd import synth

    t0 = time.perf_counter()
    w = synth.make(args.workload, scale=args.scale)
    V = len(w.positions)
    log(f"workload {w.name}: matrix {w.mat.shape}, {len(w.pop_rows)} pops, "
        f"generated in {time.perf_counter() - t0:.1f}s")

    cpu = None if args.no_cpu else cpu_baseline(w, args.cpu_seconds)
    if cpu:
        log(f"cpu baseline: {cpu['value']:.3e} pairs/s on {cpu['cores']} cores ({cpu['wall']:.1f}s)")

    import torch

    from pg_gpu_amd.engine import TILE, LDPanels, band_tiles
    from pg_gpu_amd.windows import WindowPlan

    dev = torch.device("cuda:0")
    mat = torch.as_tensor(w.mat, device=dev)
    pan = LDPanels(mat, w.pop_rows, w.mode)
    ns = pan.n_stats
    stream = torch.cuda.current_stream()
    windows = w.windows is not None
    r2 = w.name in R2_WORKLOADS
    if windows:
        launch = {"tile": 32, "group": args.window_group, "tuned": False}
        if not args.no_tune:
            from pg_gpu_amd import autotune

            best, timings = autotune.tune_windows(pan, w.windows)
            launch = dict(best, tuned=True, candidates=len(timings),
                          candidate_ms={f"tile{c['tile']}g{c['group']}": round(t, 4)
                                        for c, t in timings})
            log(f"autotune: {best} over {len(timings)} tile shapes and groups")
        plan = WindowPlan(pan, w.windows, tile=launch["tile"], group=launch["group"])
        tiles = None
        run = plan.run
        n_bins = None
    else:
        launch = {"square": None, "dynamic": True, "grid": 0, "form": args.form, "tuned": False}
        if not args.no_tune:
            from pg_gpu_amd import autotune

            best, timings = autotune.tune_binned(pan, w.positions, w.bins, w.max_dist, r2=r2)
            launch = dict(best, tuned=True, candidates=len(timings),
                          candidate_ms={f"sq{c['square']}{'d' if c['dynamic'] else 's'}"
                                        f"f{c['form']}": round(t, 4) for c, t in timings})
            log(f"autotune: {best} over {len(timings)} launch shapes")
        tiles = band_tiles(w.positions, w.max_dist, launch["square"])
        plan = pan.plan(w.positions, w.bins, w.max_dist, tiles=tiles)
        run = lambda kernel="fused": plan.run(kernel, launch["grid"], r2,  # noqa: E731
                                              launch["dynamic"], launch["form"])
        n_bins = int(len(w.bins) - 1)

    def step(kernel="fused", ev=None, pev=None, sev=None):
        # one pass of the hot path over the resident int8 matrix: site
        # tallies, MFMA panels, the fused count / estimator / bin kernel
        if sev is not None:
            sev[0].record(stream)
        pan.rescan()
        if sev is not None:
            sev[1].record(stream)
        if pev is not None:
            pev[0].record(stream)
        pan.repack()
        if pev is not None:
            pev[1].record(stream)
        if ev is not None:
            ev[0].record(stream)
        r = run(kernel)
        if ev is not None:
            ev[1].record(stream)
        return r

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log("warmup done")
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    pevs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
            for _ in range(args.steps)]
    sevs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
            for _ in range(args.steps)]
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record(stream)
    results = [step(ev=evs[k], pev=pevs[k], sev=sevs[k]) for k in range(args.steps)]
    stop.record(stream)
    torch.cuda.synchronize()
    total_ms = start.elapsed_time(stop)
    graph = None
    if windows:
        # the whole job as one captured HIP graph (tallies, panels, clear, the
        # persistent launch over every window's work items); its replays are
        # the timed steps
        g = plan.capture(repack=True)
        for _ in range(max(args.warmup, 1)):
            g.replay()
        torch.cuda.synchronize()
        start.record(stream)
        for _ in range(args.steps):
            g.replay()
        stop.record(stream)
        torch.cuda.synchronize()
        eager_ms = total_ms / args.steps
        total_ms = start.elapsed_time(stop)
        gres = g.replay().finish()
        assert np.array_equal(gres.dd, results[-1].finish().dd)
        graph = {"ms_per_replay": total_ms / args.steps, "eager_ms_per_step": eager_ms}
        log(f"graph replays: {graph['ms_per_replay']:.3f} ms (eager step {eager_ms:.3f} ms)")
    kern_ms = [a.elapsed_time(b) for a, b in evs]
    pack_ms = float(np.mean([a.elapsed_time(b) for a, b in pevs]))
    scan_ms = float(np.mean([a.elapsed_time(b) for a, b in sevs]))
    # pack: every (row, site) of each population read once, the panels written once
    pack_bytes = int(sum(pan.pop_sizes)) * V + int(pan.panels.numel())
    # tallies: every (row, site) of each population read once, 3 int32 per site written
    scan_bytes = int(sum(pan.pop_sizes)) * V + 12 * V * pan.P
    res = results[-1].finish()
    n_pairs = int(res.n_pairs.sum()) if windows else res.n_pairs
    ms_step = total_ms / args.steps
    log(f"fused: {ms_step:.3f} ms/step, kernel {np.mean(kern_ms):.3f} ms, {n_pairs:,} pairs")

    def main_sums(r):
        return np.stack([r.dd, r.dz, r.pi2], axis=1) if windows else r.sums

    # all steps must agree bit for bit (integer accumulation)
    for r in results[:-1]:
        assert np.array_equal(main_sums(r.finish()), main_sums(res))

    plain = None
    if not args.no_plain:
        ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
        rp = step(kernel="plain", ev=ev)
        torch.cuda.synchronize()
        p_ms = ev[0].elapsed_time(ev[1])
        rp = rp.finish()
        a, b = main_sums(rp), main_sums(res)
        rel = float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))
        cnt_eq = (np.array_equal(rp.n_pairs, res.n_pairs) if windows
                  else np.array_equal(rp.counts, res.counts))
        plain = dict(ms=p_ms, value=n_pairs / (p_ms * 1e-3), rel_vs_fused=rel,
                     counts_equal=bool(cnt_eq))
        log(f"plain: {p_ms:.1f} ms, {plain['value']:.3e} pairs/s, rel diff {rel:.2e}")

    prec = None
    if cpu:
        X = cpu["X_exact"]
        if windows:
            rg = WindowPlan(pan, w.windows[:X], tile=launch["tile"],
                            group=launch["group"]).run().finish()
            got = np.stack([rg.dd, rg.dz, rg.pi2], axis=1)
            ref = cpu["sums"][:, :3]
            scale = np.maximum(np.abs(ref), 1e-3 * np.abs(ref).max(axis=0))
            err = np.abs(got - ref) / np.maximum(scale, 1e-300)
            prec = dict(max_rel_err=float(err.max()),
                        max_rel_err_vs_sum=float(np.max(np.abs(got - ref) /
                                                        np.maximum(np.abs(ref), 1e-300))),
                        per_stat_max_rel_err={n: float(e) for n, e in
                                              zip(("DD", "Dz", "pi2"), err.max(axis=0))},
                        r2_sum_max_rel_err=float(np.max(np.abs(rg.r2_sum - cpu["sums"][:, 3]) /
                                                        np.maximum(cpu["sums"][:, 3], 1e-300))),
                        counts_equal=bool(np.array_equal(rg.n_pairs, cpu["counts"]) and
                                          np.array_equal(rg.r2_count,
                                                         np.rint(cpu["sums"][:, 4]).astype(int))),
                        sample_pairs=int(cpu["n_pairs_exact"]))
        else:
            sub = tiles[tiles[:, 0] < X // TILE]
            rg = pan.binned(w.positions, w.bins, w.max_dist, kernel="fused", tiles=sub, r2=r2)
            ref = cpu["sums"][:, :ns]
            scale = np.maximum(np.abs(ref), 1e-3 * cpu["abs"][:, :ns])
            err = np.abs(rg.sums - ref) / np.maximum(scale, 1e-300)
            from pg_gpu_amd.ld_pipeline import ld_names

            names = ld_names(pan.P)
            prec = dict(max_rel_err=float(err.max()),
                        max_rel_err_vs_sum=float(np.max(np.abs(rg.sums - ref) /
                                                        np.maximum(np.abs(ref), 1e-300))),
                        # per statistic, max over the bins (same scaled error)
                        per_stat_max_rel_err={n: float(e) for n, e in zip(names, err.max(axis=0))}
                        if len(names) == ns else None,
                        counts_equal=bool(np.array_equal(rg.counts, cpu["counts"])),
                        sample_pairs=int(cpu["n_pairs_exact"]))
            if r2:
                prec["r2_sum_max_rel_err"] = float(np.max(
                    np.abs(rg.r2_sum - cpu["sums"][:, ns]) /
                    np.maximum(np.abs(cpu["sums"][:, ns]), 1e-300)))
                prec["counts_equal"] = prec["counts_equal"] and bool(np.array_equal(
                    rg.r2_count, np.rint(cpu["sums"][:, ns + 1]).astype(np.int64)))
        log(f"precision vs oracle on {cpu['n_pairs_exact']:,} pairs: {prec}")

    k_ms = float(np.mean(kern_ms))
    geno = w.mode == "geno"
    # indicator products per pair and population: haplotypes H.H (+ H.V, V.H,
    # V.V with missing calls); genotypes the 2x2 (or, with missing, 3x3)
    # feature products.  2 ops (multiply + add) per sample and product.
    nprod = [((9 if m else 4) if geno else (4 if m else 1)) for m in pan.has_missing]
    ops_per_pair = 2 * sum(q * k for q, k in zip(nprod, pan.pop_sizes))

    counting = None
    if not windows and not geno and not args.no_count_stage:
        counting = counting_stage(pan, plan, launch, args.steps, n_pairs, ops_per_pair,
                                  nprod, w if cpu is not None else None)
        log(f"counting stage: {counting['kernel_ms']:.3f} ms, "
            f"{counting['frac_of_fp4_peak']:.3f} of the FP4 peak, "
            f"{counting['frac_of_i8_peak']:.3f} of the int8 peak, "
            f"checksum equal: {counting['checksum_equal']}")

    e2e = None
    if not windows:
        e2e = end_to_end(mat, w, r2, launch, n_pairs)
        log(f"end to end: {e2e['ms']:.2f} ms ({e2e['value']:.3e} pairs/s)")

    achieved = ops_per_pair * n_pairs / (k_ms * 1e-3) / 1e12
    kname = kernel_name(pan, geno, windows, r2, launch.get("form", 0))
    prof = f"{'windows' if windows else 'fused'}_{w.mode}{pan.P}_{w.name}"
    traffic = traffic_from_profile(prof) if args.scale == 1.0 else None
    util = utilisation_from_profile(prof) if args.scale == 1.0 else None
    f64 = fp64_flops_from_profile(prof) if args.scale == 1.0 else None
    fp64_valu = None if f64 is None else {
        "achieved": f64 / (k_ms * 1e-3) / 1e12, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
        "frac": f64 / (k_ms * 1e-3) / 1e12 / PEAK_FP64_TFLOPS, "flops_per_launch": f64,
        "profile": f"profiles/{prof}_fp64_valu.md"}
    line = {
        "metric": "site-pairs/sec binned DD/Dz/pi2, I8-MFMA fused vs plain FP64; max rel err",
        "value": n_pairs / (ms_step * 1e-3),
        "unit": "site-pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        # counting runs on FP4 (E2M1) 0/1 indicators with FP32 accumulation,
        # exact for counts below 2^24; the estimators and bin sums are FP64
        "dtype": ("fp4 indicators -> f32 accum (exact counts), f64 estimators" +
                  (", f32 r^2 ratios" if (r2 or windows) else "")),
        "data": "synthetic (founder-copying generator, seeded)",
        "config": {"workload": WORKLOADS.get(w.name, w.name) if args.scale == 1.0 else
                   f"{w.name} at scale {args.scale}",
                   "n_sites": V, "pairs": n_pairs, "n_bins": n_bins,
                   "n_windows": None if not windows else int(len(w.windows)),
                   "n_stats": ns, "r2": bool(r2 or windows), "scale": args.scale,
                   "params": w.params, "launch": launch},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP4_TOPS,
                     "unit": "TOP/s (FP4 MFMA on 0/1 indicators)",
                     "frac": achieved / PEAK_FP4_TOPS, "frac_of_i8_peak": achieved / PEAK_I8_TOPS,
                     # what limits the fused kernel: the vector-ALU issue of its FP64
                     # estimator epilogue, not the matrix cores (see fp64_valu, counters)
                     "limiter": "valu issue (fp64 estimator epilogue)",
                     "traffic": traffic,
                     "traffic_profile": f"profiles/{prof}_hbm_traffic.md" if traffic else None,
                     "kernel": kname, "ops_per_pair": ops_per_pair, "kernel_ms": k_ms,
                     "counters": util,
                     # the estimator epilogue's bound: FP64 vector work
                     "fp64_valu": fp64_valu,
                     "counters_profile": f"profiles/{prof}_utilisation.md" if util else None},
        "pack": {"ms": pack_ms, "bytes": pack_bytes,
                 "GB/s": pack_bytes / (pack_ms * 1e-3) / 1e9,
                 "frac_of_hbm_peak": pack_bytes / (pack_ms * 1e-3) / 1e9 / PEAK_HBM_GBS},
        "scan": {"ms": scan_ms, "bytes": scan_bytes,
                 "GB/s": scan_bytes / (scan_ms * 1e-3) / 1e9,
                 "frac_of_hbm_peak": scan_bytes / (scan_ms * 1e-3) / 1e9 / PEAK_HBM_GBS},
        "counting_stage": counting,
        "end_to_end": e2e,
        "graph": graph,
        "plain_fp64": plain,
        "speedup_vs_plain": (plain["ms"] / k_ms) if plain else None,
        "precision": prec,
        "cpu_baseline": None if cpu is None else {
            "value": cpu["value"], "unit": "site-pairs/s", "cores": cpu["cores"], "kind": "port",
            "cpu_count": cpu["cpu_count"],
            "core_cap": f"{CPU_SHARE} worker processes at most: one GPU's CPU share of the node",
            "sample": (f"oracle (NumPy restatement of the reference's chunked counts, FP64 "
                       f"estimators and np.add.at binning) on "
                       f"{'windows' if windows else 'anchors'} [0, {cpu['X']}): "
                       f"{cpu['n_pairs']:,} pairs, {cpu['wall']:.1f}s wall")},
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
