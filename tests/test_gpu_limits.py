"""Population sizes whose valid-sample counts span too wide a range for the
reciprocal table in LDS (large populations with a heavily missing site): the
binned and window kernels then read the table from global memory.  Results
against the oracle, as for every other launch."""
import numpy as np
import pytest

from oracle import ld_oracle as o

pytestmark = pytest.mark.gpu


def data(seed, n_rows, V, hi, heavy):
    rng = np.random.default_rng(seed)
    pos = np.sort(rng.choice(np.arange(1, 40 * V), V, replace=False)).astype(np.int64)
    m = rng.integers(0, hi, (n_rows, V)).astype(np.int8)
    m[rng.random(m.shape) < 0.02] = -1
    m[rng.random(n_rows) < 0.6, heavy] = -1        # one site 60% missing
    return m, pos


def check(got, ref, cnt, absr, got_cnt):
    np.testing.assert_array_equal(got_cnt, cnt)
    scale = np.maximum(np.abs(ref), 1e-3 * absr)
    assert np.max(np.abs(got - ref) / np.maximum(scale, 1e-300)) < 1e-12


@pytest.mark.parametrize("sizes", [[6000], [3000, 3000]])
def test_binned_haplotypes_wide_table(cuda, sizes):
    import torch

    from pg_gpu_amd.engine import LDPanels

    hap, pos = data(1, sum(sizes), 160, 2, 37)
    offs = np.cumsum([0] + sizes)
    pops = [list(range(offs[k], offs[k + 1])) for k in range(len(sizes))]
    bins = np.array([1.0, 50, 300, 2000, 8000])
    pan = LDPanels(torch.as_tensor(hap, device=cuda), pops, "hap")
    assert pan.n_lo[0] * 2 < sizes[0]                      # the table spans most of 0..K
    ref, cnt, _, absr = o.binned_sums_hap(hap, pos, pops, bins, max_dist=8000.0, with_abs=True)
    r = pan.binned(pos, bins, 8000.0)
    check(r.sums, ref, cnt, absr, r.counts)
    if len(sizes) == 1:
        r2 = pan.binned(pos, bins, 8000.0, r2=True)
        ref2, cnt2, _, absr2 = o.binned_sums_hap(hap, pos, pops, bins, max_dist=8000.0,
                                                 with_abs=True, r2=True)
        check(r2.sums, ref2[:, :3], cnt2, absr2[:, :3], r2.counts)
        np.testing.assert_allclose(r2.r2_sum, ref2[:, 3], rtol=1e-6)
        np.testing.assert_array_equal(r2.r2_count, np.rint(ref2[:, 4]).astype(np.int64))


def test_binned_genotypes_wide_table(cuda):
    import torch

    from pg_gpu_amd.engine import LDPanels

    g, pos = data(2, 5600, 120, 3, 11)
    pops = [list(range(5600))]
    bins = np.array([1.0, 100, 1000, 6000])
    pan = LDPanels(torch.as_tensor(g, device=cuda), pops, "geno")
    ref, cnt, _, absr = o.binned_sums_geno(g, pos, pops, bins, max_dist=6000.0, with_abs=True)
    r = pan.binned(pos, bins, 6000.0)
    check(r.sums, ref, cnt, absr, r.counts)


def test_windows_wide_table(cuda):
    import torch

    from pg_gpu_amd.engine import LDPanels
    from pg_gpu_amd.windows import WindowPlan

    hap, pos = data(3, 6000, 150, 2, 70)
    windows = np.array([[0, 40], [40, 41], [41, 150]])
    pan = LDPanels(torch.as_tensor(hap, device=cuda), [list(range(6000))], "hap")
    res = WindowPlan(pan, windows).run().finish()
    sums, counts = o.window_sums(hap, windows)
    np.testing.assert_array_equal(res.n_pairs, counts)
    got = np.stack([res.dd, res.dz, res.pi2], axis=1)
    scale = np.maximum(np.abs(sums[:, :3]), 1e-3 * np.abs(sums[:, :3]).max(axis=0))
    assert np.max(np.abs(got - sums[:, :3]) / np.maximum(scale, 1e-300)) < 1e-12
    np.testing.assert_array_equal(res.r2_count, np.rint(sums[:, 4]).astype(np.int64))
