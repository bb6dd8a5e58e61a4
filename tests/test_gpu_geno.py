"""GPU parity of the binned LD path for genotype data (0/1/2/-1 dosages):
fused MFMA kernel and plain comparator against the oracle's kernel-form
estimators (pg_gpu/genotype_kernels.py via moments_ld.py:248-343).

Bars as for haplotypes: per-bin pair counts and the dense 9-way counts are
bit-exact; bin sums agree with the oracle's correctly rounded FP64 sums to
within 1e-12 of max(|sum|, 1e-3 * sum|terms|).
"""
import numpy as np
import pytest

from oracle import geno_oracle as go
from oracle import ld_oracle as o
from pg_gpu_amd import synth
from pg_gpu_amd.engine import LDPanels

pytestmark = pytest.mark.gpu
REL = 1e-12


def run_geno(cuda, geno, pos, pops, bins, max_dist=None, coords=None,
             kernels=("fused", "fused_pc", "plain")):
    import torch

    print(f"case: geno {geno.shape} pops {[len(p) for p in pops]} bins {len(bins) - 1}",
          flush=True)
    mat = torch.as_tensor(geno, device=cuda)
    pan = LDPanels(mat, pops, "geno")
    ref, cnt, n_enum, absr = o.binned_sums_geno(geno, pos, pops, bins, max_dist=max_dist,
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


def make_geno(rng, sizes, V, span, missing, missing_pops=None):
    pos = np.sort(rng.choice(np.arange(1, span + 1), size=V, replace=False))
    fr = rng.beta(0.3, 0.3, V)
    founders = (rng.random((12, V)) < fr).astype(np.int8)
    hap = synth._copy(rng, founders, pos, 2 * sum(sizes), 1e-4)
    geno = (hap[0::2].astype(np.int16) + hap[1::2]).astype(np.int8)
    offs = np.concatenate([[0], np.cumsum(sizes)])
    pops = [list(range(offs[q], offs[q + 1])) for q in range(len(sizes))]
    if missing:
        for q in (range(len(sizes)) if missing_pops is None else missing_pops):
            blk = geno[offs[q]:offs[q + 1]]
            blk[rng.random(blk.shape) < missing] = -1
    return geno, pos, pops


def test_geno_single_pop_no_missing(cuda):
    rng = np.random.default_rng(11)
    geno, pos, pops = make_geno(rng, [60], 700, 70_000, 0.0)
    run_geno(cuda, geno, pos, pops, synth.log_bins(100, 20_000, 8), 20_000.0)


def test_geno_single_pop_missing(cuda):
    rng = np.random.default_rng(12)
    geno, pos, pops = make_geno(rng, [45], 600, 60_000, 0.04)
    run_geno(cuda, geno, pos, pops, synth.log_bins(100, 20_000, 8), 20_000.0)


def test_geno_two_pops_missing(cuda):
    rng = np.random.default_rng(13)
    geno, pos, pops = make_geno(rng, [40, 70], 500, 40_000, 0.03)
    run_geno(cuda, geno, pos, pops, synth.log_bins(100, 15_000, 6), 15_000.0)


def test_geno_two_pops_missing_in_one(cuda):
    rng = np.random.default_rng(14)
    geno, pos, pops = make_geno(rng, [33, 64], 420, 30_000, 0.05, missing_pops=[1])
    run_geno(cuda, geno, pos, pops, synth.log_bins(100, 12_000, 5), 12_000.0)


def test_geno_two_pops_no_missing_ragged(cuda):
    rng = np.random.default_rng(15)
    geno, pos, pops = make_geno(rng, [7, 97], 333, 25_000, 0.0)
    run_geno(cuda, geno, pos, pops, np.array([0.0, 50.0, 400.0, 3_000.0, 10_000.0]), 10_000.0)


def test_geno_genetic_coords(cuda):
    rng = np.random.default_rng(16)
    geno, pos, pops = make_geno(rng, [50], 400, 40_000, 0.02)
    coords = np.cumsum(rng.exponential(1e-6, len(pos)))
    bins = np.array([0.0, 1e-5, 5e-5, 2e-4])
    run_geno(cuda, geno, pos, pops, bins, 30_000.0, coords=coords)


def test_geno_workload_c2_scaled(cuda):
    w = synth.make("c2", scale=0.004)
    run_geno(cuda, w.mat, w.positions, w.pop_rows, w.bins, w.max_dist)


@pytest.mark.parametrize("sizes,missing_pops", [([20, 35, 12], None), ([9, 30, 17, 25], [0, 2]),
                                                ([6, 70, 11, 40], None)])
def test_geno_three_and_four_populations(cuda, sizes, missing_pops):
    """Three and four genotype populations of uneven sizes (the population-set
    passes of the fused epilogue), missing calls in all or some of them."""
    rng = np.random.default_rng(sum(sizes))
    geno, pos, pops = make_geno(rng, sizes, 300, 30_000, 0.04, missing_pops=missing_pops)
    run_geno(cuda, geno, pos, pops, synth.log_bins(100, 15_000, 5), 15_000.0)


def test_geno_fused_reproducible_across_grids(cuda):
    import torch

    rng = np.random.default_rng(17)
    geno, pos, pops = make_geno(rng, [40, 40], 640, 50_000, 0.03)
    pan = LDPanels(torch.as_tensor(geno, device=cuda), pops, "geno")
    bins = synth.log_bins(100, 20_000, 6)
    a = pan.binned(pos, bins, 20_000.0, grid=8)
    b = pan.binned(pos, bins, 20_000.0, grid=0)
    np.testing.assert_array_equal(a.sums, b.sums)
    np.testing.assert_array_equal(a.counts, b.counts)


@pytest.mark.parametrize("missing", [0.0, 0.05])
def test_geno_dense_counts_bit_exact(cuda, missing):
    import torch

    rng = np.random.default_rng(18)
    geno, pos, pops = make_geno(rng, [75], 96, 10_000, missing)
    pan = LDPanels(torch.as_tensor(geno, device=cuda), pops, "geno")
    V = geno.shape[1]
    got = pan.count_block(0, 0, V, 0, V).cpu().numpy()
    ii, jj = np.meshgrid(np.arange(V), np.arange(V), indexing="ij")
    c, _ = o.geno_counts(geno, ii.ravel(), jj.ravel(), pops[0])
    want = go.to_moments(c).T.reshape(9, V, V).astype(np.int32)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("sizes,missing", [([1100], 0.02), ([1030, 40], 0.0)])
def test_geno_large_population_16bit_counts(cuda, sizes, missing):
    """Populations above 1023 individuals take the 16-bit count packing."""
    rng = np.random.default_rng(19)
    geno, pos, pops = make_geno(rng, sizes, 200, 20_000, missing)
    run_geno(cuda, geno, pos, pops, synth.log_bins(100, 10_000, 5), 10_000.0)


def test_coords_must_be_non_decreasing(cuda):
    import torch

    rng = np.random.default_rng(20)
    geno, pos, pops = make_geno(rng, [20], 64, 5_000, 0.0)
    pan = LDPanels(torch.as_tensor(geno, device=cuda), pops, "geno")
    coords = np.linspace(0.0, 1e-3, len(pos))
    coords[10] = coords[11] + 1e-4
    with pytest.raises(ValueError):
        pan.binned(pos, np.array([0.0, 1e-4, 1e-3]), 5_000.0, coords=coords)


@pytest.mark.parametrize("sizes,missing", [([20, 30, 25], 0.03), ([9, 16, 21, 40], 0.0)])
def test_geno_three_and_four_pops(cuda, sizes, missing):
    rng = np.random.default_rng(21)
    geno, pos, pops = make_geno(rng, sizes, 200, 20_000, missing)
    run_geno(cuda, geno, pos, pops, synth.log_bins(100, 8_000, 4), 8_000.0)


@pytest.mark.parametrize("kernel,form", [("fused", 0), ("fused", 1), ("plain", 0)])
def test_geno_mean_r2_c2_scaled(cuda, kernel, form):
    """Workload c2 shape: genotype DD, Dz, pi2 and mean dosage r^2 per bin."""
    import torch

    from test_gpu_binned import check_r2

    w = synth.make("c2", scale=0.004)
    pan = LDPanels(torch.as_tensor(w.mat, device=cuda), w.pop_rows, "geno")
    ref, cnt, _ = o.binned_sums_geno(w.mat, w.positions, w.pop_rows, w.bins,
                                     max_dist=w.max_dist, r2=True)
    res = pan.binned(w.positions, w.bins, w.max_dist, kernel=kernel, r2=True, form=form)
    np.testing.assert_array_equal(res.counts, cnt)
    check_r2(res, ref, 3, pan.binned(w.positions, w.bins, w.max_dist, kernel=kernel, form=form))
