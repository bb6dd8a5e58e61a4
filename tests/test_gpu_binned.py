"""GPU parity of the binned LD path (fused MFMA kernel and plain comparator)
against the CPU oracle, through the C ABI.

Bars: pair counts per bin and the dense 4-way counts are bit-exact; bin sums
of DD/Dz/pi2 agree with the oracle's correctly rounded FP64 sums to within
1e-12 of max(|sum|, 1e-3 * sum|terms|) -- per-term rounding is a few ulp of
each term, so the error is judged against the terms' own scale where a bin
sum cancels.  Repeated runs and different grid sizes give identical limbs.
"""
import numpy as np
import pytest

from oracle import ld_oracle as o
from pg_gpu_amd import synth
from pg_gpu_amd.engine import LDPanels

pytestmark = pytest.mark.gpu
REL = 1e-12


def run_case(cuda, hap, pos, pops, bins, max_dist=None, coords=None, kernels=("fused", "fused_pc", "plain")):
    import torch

    print(f"case: hap {hap.shape} pops {[len(p) for p in pops]} bins {len(bins) - 1}", flush=True)
    mat = torch.as_tensor(hap, device=cuda)
    pan = LDPanels(mat, pops, "hap")
    ref, cnt, n_enum, absr = o.binned_sums_hap(hap, pos, pops, bins, max_dist=max_dist,
                                               coords=coords, with_abs=True)
    out = {}
    for k in kernels:
        # "fused_pc": the fused kernel's producer/consumer form
        r = pan.binned(pos, bins, max_dist=max_dist, coords=coords,
                       kernel="plain" if k == "plain" else "fused", form=int(k == "fused_pc"))
        torch.cuda.synchronize()
        np.testing.assert_array_equal(r.counts, cnt)
        assert r.n_pairs == n_enum
        scale = np.maximum(np.abs(ref), 1e-3 * absr)
        err = np.abs(r.sums - ref)
        bad = err > REL * scale + 1e-300
        assert not bad.any(), (k, np.max(err / np.maximum(scale, 1e-300)))
        out[k] = r
    return pan, out


def small_pops(rng, n_hap, V, span, missing, n_pops=2, sizes=None):
    pos = np.sort(rng.choice(np.arange(1, span + 1), size=V, replace=False))
    sizes = sizes or [n_hap // n_pops] * n_pops
    fr = rng.beta(0.3, 0.3, V)
    founders = (rng.random((10, V)) < fr).astype(np.int8)
    hap = synth._copy(rng, founders, pos, sum(sizes), 1e-4)
    if missing:
        hap[rng.random(hap.shape) < missing] = -1
    offs = np.concatenate([[0], np.cumsum(sizes)])
    return hap, pos, [list(range(offs[q], offs[q + 1])) for q in range(len(sizes))]


def test_single_pop_no_missing(cuda):
    w = synth.make("c0", scale=0.1)
    run_case(cuda, w.mat, w.positions, w.pop_rows, w.bins, w.max_dist)


def test_two_pops_missing(cuda):
    w = synth.make("c1", scale=0.005)
    run_case(cuda, w.mat, w.positions, w.pop_rows, w.bins, w.max_dist)


@pytest.mark.parametrize("seed,V,sizes,missing", [
    (1, 333, [37, 50], 0.0),        # ragged V and population sizes
    (2, 200, [64, 33], 0.08),       # missing in both
    (3, 129, [70], 0.03),           # one pop with missing
    (4, 65, [1, 2], 0.0),           # tiny pops: every estimator guard path
])
def test_ragged_shapes(cuda, seed, V, sizes, missing):
    rng = np.random.default_rng(seed)
    hap, pos, pops = small_pops(rng, sum(sizes), V, 40 * V, missing, sizes=sizes)
    bins = np.array([0.0, 5.0, 50.0, 300.0, 1000.0, 5000.0])
    run_case(cuda, hap, pos, pops, bins, max_dist=3000.0)


def test_missing_in_one_pop_only(cuda):
    rng = np.random.default_rng(9)
    hap, pos, pops = small_pops(rng, 120, 300, 30000, 0.0, sizes=[60, 60])
    hap[pops[1][0]:pops[1][-1] + 1][rng.random((60, 300)) < 0.05] = -1
    run_case(cuda, hap, pos, pops, synth.log_bins(10, 20000, 12))


def test_duplicate_positions_cutoff_and_genetic_coords(cuda):
    rng = np.random.default_rng(11)
    V = 260
    pos = np.sort(rng.integers(0, 2000, V))           # many duplicate positions
    hap = (rng.random((48, V)) < 0.4).astype(np.int8)
    pops = [list(range(48))]
    bins = np.array([0.0, 1.0, 10.0, 100.0, 400.0])
    run_case(cuda, hap, pos, pops, bins, max_dist=400.0)   # d == max_dist enumerated, dropped
    morgans = np.cumsum(rng.random(V)) * 1e-4
    run_case(cuda, hap, pos, pops, np.array([0.0, 1e-4, 5e-4, 2e-3]), max_dist=700.0,
             coords=morgans)


def test_empty_band(cuda):
    rng = np.random.default_rng(5)
    hap, pos, pops = small_pops(rng, 40, 100, 100000, 0.0, n_pops=1)
    pos = np.arange(100) * 1000
    _, out = run_case(cuda, hap, pos, pops, np.array([1.0, 10.0, 100.0]))
    assert out["fused"].counts.sum() == 0


def test_bitwise_reproducible_across_grids(cuda):
    import torch

    w = synth.make("c1", scale=0.004)
    pan = LDPanels(torch.as_tensor(w.mat, device=cuda), w.pop_rows, "hap")
    base = pan.binned(w.positions, w.bins, w.max_dist, kernel="fused")
    for grid in (8, 64, 1024):
        for form in (0, 1):
            r = pan.binned(w.positions, w.bins, w.max_dist, kernel="fused", grid=grid, form=form)
            np.testing.assert_array_equal(r.sums, base.sums)
            np.testing.assert_array_equal(r.counts, base.counts)


@pytest.mark.parametrize("missing", [0.0, 0.07])
def test_dense_counts_bit_exact(cuda, missing):
    import torch

    rng = np.random.default_rng(21)
    V, K = 150, 77
    hap = (rng.random((K, V)) < 0.5).astype(np.int8)
    if missing:
        hap[rng.random(hap.shape) < missing] = -1
    pan = LDPanels(torch.as_tensor(hap, device=cuda), [list(range(K))], "hap")
    got = pan.count_block(0, 0, V, 0, V).cpu().numpy()
    ii, jj = np.meshgrid(np.arange(V), np.arange(V), indexing="ij")
    c, _ = o.hap_counts(hap, ii.ravel(), jj.ravel())
    np.testing.assert_array_equal(got.reshape(4, -1).T, c)


@pytest.mark.parametrize("offset", [0, 3, 8])
def test_dense_counts_from_unaligned_views(cuda, offset):
    """Rows of a column-sliced view start off 16-byte boundaries (row stride
    1,000 + 37 bytes, base offset): the pack kernel's funnel-shifted loads and
    its byte path at the row ends must give the same counts as a copy."""
    import torch

    rng = np.random.default_rng(23 + offset)
    K, W = 90, 1037
    full = (rng.random((K, W)) < 0.4).astype(np.int8)
    full[rng.random(full.shape) < 0.05] = -1
    big = torch.as_tensor(full, device=cuda)
    view = big[:, offset:offset + 999]
    assert view.stride(1) == 1 and view.stride(0) == W
    hap = full[:, offset:offset + 999]
    pan = LDPanels(view, [list(range(0, K, 2)), list(range(1, K, 2))], "hap")
    assert pan.ld == W
    ii, jj = np.meshgrid(np.arange(999), np.arange(0, 999, 7), indexing="ij")
    for p, rows in enumerate([list(range(0, K, 2)), list(range(1, K, 2))]):
        got = pan.count_block(p, 0, 999, 0, 999).cpu().numpy()
        c, _ = o.hap_counts(hap, ii.ravel(), jj.ravel(), rows)
        np.testing.assert_array_equal(got[:, :, ::7].reshape(4, -1).T, c)


def test_three_pops_missing(cuda):
    rng = np.random.default_rng(31)
    hap, pos, pops = small_pops(rng, 0, 300, 30_000, 0.03, sizes=[20, 45, 70])
    run_case(cuda, hap, pos, pops, synth.log_bins(100, 12_000, 6), 12_000.0)


def test_four_pops_many_bins(cuda):
    """36 bins x 106 outputs of limbs leave no LDS room for the operand ring:
    the fused kernel falls back to register-streamed operands."""
    rng = np.random.default_rng(32)
    hap, pos, pops = small_pops(rng, 0, 260, 26_000, 0.04, sizes=[12, 30, 18, 44])
    run_case(cuda, hap, pos, pops, synth.log_bins(50, 10_000, 36), 10_000.0,
             kernels=("fused",))


def test_four_pops_64_bins_split_launches(cuda):
    """64 bins x 106 outputs of limbs exceed LDS: the fused entry point covers
    the bins in consecutive launches over bin ranges."""
    rng = np.random.default_rng(33)
    hap, pos, pops = small_pops(rng, 0, 240, 24_000, 0.03, sizes=[10, 26, 20, 40])
    run_case(cuda, hap, pos, pops, synth.log_bins(20, 9_000, 64), 9_000.0)


def test_two_pops_64_bins_one_launch(cuda):
    """The largest bin count in one launch (65 edges: one per lane plus the
    last), with the cutoff past the last edge so that pairs beyond the bins
    are skipped by the tile classification and the bin words."""
    rng = np.random.default_rng(34)
    hap, pos, pops = small_pops(rng, 0, 400, 30_000, 0.0, sizes=[30, 41])
    run_case(cuda, hap, pos, pops, synth.log_bins(20, 8_000, 64), 9_000.0)


def test_four_pops_workload_c4_scaled(cuda):
    """40/120/300/800 haplotypes, 3% missing, rare-allele sites: the 105-statistic set."""
    w = synth.make("c4", scale=0.0008)
    run_case(cuda, w.mat, w.positions, w.pop_rows, w.bins, w.max_dist)


def test_four_pops_no_missing_ragged(cuda):
    rng = np.random.default_rng(32)
    hap, pos, pops = small_pops(rng, 0, 150, 20_000, 0.0, sizes=[3, 33, 64, 97])
    run_case(cuda, hap, pos, pops, np.array([0.0, 300.0, 2_000.0, 8_000.0]), 8_000.0)


def check_r2(res, ref, n_ld, base):
    """r^2 columns: the defined-pair count is exact; the r^2 sum is formed from
    FP32 ratios (about 2^-22 relative per term).  The LD sums are bit-identical
    to the same launch without r^2 (checked against the oracle elsewhere)."""
    np.testing.assert_array_equal(res.r2_count, np.rint(ref[:, n_ld + 1]).astype(np.int64))
    np.testing.assert_allclose(res.r2_sum, ref[:, n_ld], rtol=2e-6, atol=1e-9)
    np.testing.assert_array_equal(res.sums, base.sums)
    np.testing.assert_array_equal(res.counts, base.counts)


@pytest.mark.parametrize("kernel,form", [("fused", 0), ("fused", 1), ("plain", 0)])
def test_single_pop_mean_r2(cuda, kernel, form):
    """Workload c0 shape: DD, Dz, pi2 and mean r^2 per bin."""
    import torch

    w = synth.make("c0", scale=0.1)
    hap = w.mat.copy()
    hap[:, 7] = 0                                     # a monomorphic site: r^2 undefined
    pan = LDPanels(torch.as_tensor(hap, device=cuda), w.pop_rows, "hap")
    ref, cnt, _ = o.binned_sums_hap(hap, w.positions, w.pop_rows, w.bins, max_dist=w.max_dist,
                                    r2=True)
    res = pan.binned(w.positions, w.bins, w.max_dist, kernel=kernel, r2=True, form=form)
    base = pan.binned(w.positions, w.bins, w.max_dist, kernel=kernel, form=form)
    np.testing.assert_array_equal(res.counts, cnt)
    check_r2(res, ref, 3, base)
    assert np.all(res.r2_count <= res.counts) and res.r2_count.sum() < res.counts.sum()


def test_single_pop_mean_r2_missing(cuda):
    import torch

    rng = np.random.default_rng(33)
    hap, pos, pops = small_pops(rng, 0, 300, 30_000, 0.06, sizes=[90])
    pan = LDPanels(torch.as_tensor(hap, device=cuda), pops, "hap")
    bins = synth.log_bins(100, 10_000, 5)
    ref, cnt, _ = o.binned_sums_hap(hap, pos, pops, bins, max_dist=10_000.0, r2=True)
    res = pan.binned(pos, bins, 10_000.0, r2=True)
    np.testing.assert_array_equal(res.counts, cnt)
    check_r2(res, ref, 3, pan.binned(pos, bins, 10_000.0))


def test_static_and_dynamic_tile_schedules_agree(cuda):
    """The atomic-counter schedule and the static stride give the same limbs."""
    import torch

    w = synth.make("c1", scale=0.004)
    pan = LDPanels(torch.as_tensor(w.mat, device=cuda), w.pop_rows, "hap")
    plan = pan.plan(w.positions, w.bins, w.max_dist)
    a = plan.run(dynamic=True).finish()
    b = plan.run(dynamic=False).finish()
    c = plan.run(dynamic=True, grid=16).finish()
    np.testing.assert_array_equal(a.sums, b.sums)
    np.testing.assert_array_equal(a.sums, c.sums)
    np.testing.assert_array_equal(a.counts, b.counts)


def test_graph_replays_match_eager(cuda):
    """A captured pass (re-pack, clears, fused kernel) replayed several times,
    with eager launches and host reads in between, gives the eager limbs bit
    for bit every time."""
    import torch

    rng = np.random.default_rng(41)
    hap, pos, pops = small_pops(rng, 96, 700, 300_000, 0.04)
    bins = np.array([0.0, 500, 2_000, 8_000, 30_000, 100_000])
    pan = LDPanels(torch.as_tensor(hap, device=cuda), pops, "hap")
    plan = pan.plan(pos, bins, 60_000)
    eager = plan.run().finish()
    g = plan.capture(repack=True)
    for rep in range(4):
        got = g.replay().finish()
        np.testing.assert_array_equal(got.counts, eager.counts)
        np.testing.assert_array_equal(got.sums, eager.sums)
        # eager work between replays: another pass and unrelated device work
        again = plan.run(dynamic=rep % 2 == 0).finish()
        np.testing.assert_array_equal(again.sums, eager.sums)
        torch.cuda.synchronize()
        _ = (torch.arange(1000, device=cuda) * rep).sum().item()


def test_graph_replays_single_pop_r2(cuda):
    import torch

    rng = np.random.default_rng(42)
    hap, pos, pops = small_pops(rng, 64, 500, 200_000, 0.0, n_pops=1)
    bins = np.array([0.0, 1_000, 10_000, 50_000])
    pan = LDPanels(torch.as_tensor(hap, device=cuda), pops, "hap")
    plan = pan.plan(pos, bins, 50_000)
    eager = plan.run(r2=True).finish()
    for form in (0, 1):
        g = plan.capture(r2=True, form=form)
        for _ in range(3):
            got = g.replay().finish()
            np.testing.assert_array_equal(got.sums, eager.sums)
            np.testing.assert_array_equal(got.r2_sum, eager.r2_sum)
            np.testing.assert_array_equal(got.r2_count, eager.r2_count)
            torch.cuda.synchronize()
            _ = (torch.arange(1000, device=cuda) * 3).sum().item()


def test_autotuned_launch_shape_gives_identical_sums(cuda):
    import torch

    from pg_gpu_amd import autotune

    rng = np.random.default_rng(43)
    hap, pos, pops = small_pops(rng, 80, 900, 400_000, 0.03)
    bins = np.geomspace(100, 80_000, 9)
    pan = LDPanels(torch.as_tensor(hap, device=cuda), pops, "hap")
    autotune.clear_cache()
    best, timings = autotune.tune_binned(pan, pos, bins, 80_000, reps=1, squares=(4, 16))
    assert len(timings) == 8 and best in [t[0] for t in timings]
    assert all(ms > 0 for _, ms in timings)
    again, t2 = autotune.tune_binned(pan, pos, bins, 80_000)
    assert again == best and t2 == []              # cached per job shape
    ref = pan.plan(pos, bins, 80_000).run().finish()
    got = pan.plan(pos, bins, 80_000, square=best["square"]).run(
        grid=best["grid"], dynamic=best["dynamic"], form=best["form"]).finish()
    np.testing.assert_array_equal(got.sums, ref.sums)
    np.testing.assert_array_equal(got.counts, ref.counts)
