"""GPU parity of the ragged-window path (per-window all-pairs DD/Dz/pi2 and
r^2 sums) against the oracle.  Bars: pair counts and defined-r^2 counts are
exact; DD/Dz/pi2 sums within 1e-12 of max(|sum|, 1e-3 * sum|terms|); r^2
sums (FP32 ratios) within 2e-6 relative."""
import numpy as np
import pytest

from oracle import ld_oracle as o
from pg_gpu_amd import synth
from pg_gpu_amd.engine import LDPanels
from pg_gpu_amd.windows import WindowPlan, windows_from_positions

pytestmark = pytest.mark.gpu


def run_windows(cuda, mat, windows, mode="hap", kernels=("fused", "plain"), tiles=(32,)):
    import torch

    print(f"case: windows {len(windows)} {mode} {mat.shape} tiles {tiles}", flush=True)
    pan = LDPanels(torch.as_tensor(mat, device=cuda), [list(range(mat.shape[0]))], mode)
    ref, cnt = o.window_sums(mat, windows, mode=mode)
    plan = None
    runs = [(k, 32) for k in kernels] + [("fused", t) for t in tiles if t != 32]
    for k, tile in runs:
        plan = WindowPlan(pan, windows, tile=tile)
        r = plan.run(k).finish()
        np.testing.assert_array_equal(r.n_pairs, cnt)
        np.testing.assert_array_equal(r.r2_count, np.rint(ref[:, 4]).astype(np.int64))
        np.testing.assert_allclose(r.r2_sum, ref[:, 3], rtol=2e-6, atol=1e-9)
        got = np.stack([r.dd, r.dz, r.pi2], axis=1)
        scale = np.maximum(np.abs(ref[:, :3]), 1e-3 * np.abs(ref[:, :3]).max(axis=0))
        assert np.all(np.abs(got - ref[:, :3]) <= 1e-12 * scale + 1e-300), (k, tile)
    return WindowPlan(pan, windows)


def test_windows_c3_scaled(cuda):
    """Workload c3 shape: Gamma(0.8)-distributed SNP counts per 50 kb window."""
    w = synth.make("c3", scale=0.004)
    sel = [2, 3, 5, 6, 9, 13, 17, 18, 19]          # 20 .. 683 sites (a subset keeps the oracle fast)
    run_windows(cuda, w.mat, w.windows[sel], tiles=(32, 16))


def test_windows_ragged_edges_and_empty(cuda):
    rng = np.random.default_rng(41)
    V = 300
    hap = (rng.random((70, V)) < rng.beta(0.4, 0.4, V)).astype(np.int8)
    hap[rng.random(hap.shape) < 0.04] = -1
    hap[:, 10] = 0
    windows = np.array([[0, 0], [0, 1], [1, 33], [33, 34], [34, 97], [97, 97], [97, 300]])
    run_windows(cuda, hap, windows, tiles=(32, 16))


def test_windows_genotypes(cuda):
    rng = np.random.default_rng(42)
    V = 260
    pos = np.sort(rng.choice(np.arange(1, 30_001), V, replace=False))
    hap = (rng.random((120, V)) < rng.beta(0.4, 0.4, V)).astype(np.int16)
    geno = (hap[0::2] + hap[1::2]).astype(np.int8)
    geno[rng.random(geno.shape) < 0.02] = -1
    run_windows(cuda, geno, windows_from_positions(pos, 7_000.0, start=0.0), mode="geno")


def test_windows_graph_replays_match_eager(cuda):
    """One captured job (re-pack, clear, one persistent launch over every
    window's work items) replayed with eager work in between equals the eager
    launch bit for bit."""
    import torch

    w = synth.make("c3", scale=0.004)
    sel = [2, 3, 5, 6, 9, 13]
    plan = run_windows(cuda, w.mat, w.windows[sel], kernels=("fused",))
    eager = plan.run().finish()
    g = plan.capture(repack=True)
    for rep in range(4):
        got = g.replay().finish()
        for f in ("dd", "dz", "pi2", "r2_sum", "r2_count", "n_pairs"):
            np.testing.assert_array_equal(getattr(got, f), getattr(eager, f))
        again = plan.run().finish()
        np.testing.assert_array_equal(again.dd, eager.dd)
        _ = (torch.arange(100, device=cuda) + rep).sum().item()


@pytest.mark.parametrize("missing", [0.0, 0.05])
def test_windows_16_site_tiles_odd_ksteps(cuda, missing):
    """16 x 16 tiles with a sample count whose last 128-sample MFMA step is
    half empty (300 haplotypes: five 64-sample k-steps) and windows starting
    and ending inside 16-site blocks."""
    rng = np.random.default_rng(43)
    V = 200
    hap = (rng.random((300, V)) < rng.beta(0.4, 0.4, V)).astype(np.int8)
    if missing:
        hap[rng.random(hap.shape) < missing] = -1
    windows = np.array([[3, 9], [9, 40], [40, 41], [41, 200]])
    run_windows(cuda, hap, windows, kernels=("fused",), tiles=(32, 16))


def test_window_autotune_picks_a_measured_tile(cuda):
    import torch

    from pg_gpu_amd import autotune

    w = synth.make("c3", scale=0.004)
    pan = LDPanels(torch.as_tensor(w.mat, device=cuda), [list(range(w.mat.shape[0]))], "hap")
    best, timings = autotune.tune_windows(pan, w.windows, use_cache=False)
    assert best["tile"] in (16, 32) and len(timings) == 6
    assert best == autotune.pick_window_launch(timings)
    fastest = min(t for _, t in timings)
    assert min(t for c, t in timings if c == best) <= fastest * (1 + autotune.WINDOW_TIE)


def test_window_groups_give_identical_sums(cuda):
    """The work-item group size changes only which wave takes which items:
    the integer limb sums are bit-identical for every group size and tile."""
    import torch

    w = synth.make("c3", scale=0.004)
    sel = [2, 3, 5, 6, 9, 13, 17]
    pan = LDPanels(torch.as_tensor(w.mat, device=cuda), [list(range(w.mat.shape[0]))], "hap")
    for tile in (32, 16):
        print(f"case: groups tile {tile}", flush=True)
        base = WindowPlan(pan, w.windows[sel], tile=tile).run().finish()
        for group in (1, 3, 64, 4096):
            got = WindowPlan(pan, w.windows[sel], tile=tile, group=group).run().finish()
            for f in ("dd", "dz", "pi2", "r2_sum", "r2_count", "n_pairs"):
                np.testing.assert_array_equal(getattr(got, f), getattr(base, f), err_msg=f)
    with pytest.raises(ValueError):
        WindowPlan(pan, w.windows[sel], group=5000)


def test_window_slot_overflow_matches_oracle(cuda):
    """Haplotype windows without missing calls combine the four waves'
    window totals in LDS slots, four windows per workgroup turn; later
    windows of a turn add theirs straight to the output.  Many one-item and
    empty windows (140 windows, 95 of them with work items) on a grid of 8
    workgroups put up to 18 windows in a turn (76 windows take the direct
    path at 16 items per wave turn, 7 at one item): the limbs are
    bit-identical to launches on the default grid and match the oracle."""
    import torch

    rng = np.random.default_rng(43)
    V = 2400
    hap = (rng.random((90, V)) < rng.beta(0.4, 0.4, V)).astype(np.int8)
    wins, s = [], 0
    while s < V - 400:
        k = int(rng.choice([0, 1, 2, 5, 20, 33, 40]))
        wins.append((s, s + k))
        s += k + int(rng.integers(0, 4))
    wins.append((s, min(V, s + 300)))
    windows = np.array(wins, np.int64)
    print(f"case: slot overflow, {len(windows)} windows", flush=True)
    pan = LDPanels(torch.as_tensor(hap, device=cuda), [list(range(hap.shape[0]))], "hap")
    assert not any(pan.has_missing)
    ref, cnt = o.window_sums(hap, windows, mode="hap")
    limbs = []
    for grid, group in ((8, 16), (8, 1), (0, 1), (0, 16)):
        pend = WindowPlan(pan, windows, tile=32, group=group).run(grid=grid)
        limbs.append(pend.out.cpu().numpy().copy())
        r = pend.finish()
        np.testing.assert_array_equal(r.n_pairs, cnt)
        np.testing.assert_array_equal(r.r2_count, np.rint(ref[:, 4]).astype(np.int64))
        np.testing.assert_allclose(r.r2_sum, ref[:, 3], rtol=2e-6, atol=1e-9)
        got = np.stack([r.dd, r.dz, r.pi2], axis=1)
        scale = np.maximum(np.abs(ref[:, :3]), 1e-3 * np.abs(ref[:, :3]).max(axis=0))
        assert np.all(np.abs(got - ref[:, :3]) <= 1e-12 * scale + 1e-300), (grid, group)
    for x in limbs[1:]:
        np.testing.assert_array_equal(x, limbs[0])
