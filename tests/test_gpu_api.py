"""GPU tests of the reference-compatible Python API (through libpgld) against
the oracle and the moments golden fixture."""
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import ld_oracle as o

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"


def test_iter_pairs_matches_oracle(cuda):
    from pg_gpu_amd.ld_pipeline import iter_pairs_within_distance

    rng = np.random.default_rng(12345)
    pos = np.unique(np.sort(rng.integers(0, 100_000, size=500)))
    ref_i, ref_j = o.pairs_within_distance(pos, 5000)
    for cs in (1, 100, 10_000, 1_000_000):
        parts = list(iter_pairs_within_distance(pos, 5000, cs))
        ii = np.concatenate([p[0].cpu().numpy() for p in parts])
        jj = np.concatenate([p[1].cpu().numpy() for p in parts])
        np.testing.assert_array_equal(ii, ref_i)
        np.testing.assert_array_equal(jj, ref_j)
    assert list(iter_pairs_within_distance(np.array([42]), 100, 10)) == []
    assert list(iter_pairs_within_distance(np.array([0, 10, 20]), 0, 10)) == []


@pytest.mark.parametrize("missing", [0.0, 0.05])
def test_counts_for_pairs(cuda, missing):
    import torch

    from pg_gpu_amd.ld_pipeline import compute_counts_for_pairs

    rng = np.random.default_rng(3)
    hap = (rng.random((90, 140)) < 0.4).astype(np.int8)
    if missing:
        hap[rng.random(hap.shape) < missing] = -1
    ii, jj = o.pairs_within_distance(np.arange(140), 60)
    rows = sorted(rng.choice(90, 61, replace=False).tolist())
    got_c, got_n = compute_counts_for_pairs(torch.as_tensor(hap, device=cuda), ii, jj, rows)
    ref_c, ref_n = o.hap_counts(hap, ii, jj, rows)
    np.testing.assert_array_equal(got_c.cpu().numpy(), ref_c)
    if ref_n is None:
        assert got_n is None
    else:
        np.testing.assert_array_equal(got_n.cpu().numpy(), ref_n)


@pytest.mark.parametrize("missing", [0.0, 0.05])
def test_genotype_counts_for_pairs(cuda, missing):
    import torch

    from pg_gpu_amd.ld_pipeline import compute_genotype_counts_for_pairs

    rng = np.random.default_rng(42)
    g = rng.integers(0, 3, (70, 100)).astype(np.int8)
    if missing:
        g[rng.random(g.shape) < missing] = -1
    ii, jj = o.pairs_within_distance(np.arange(100), 40)
    got_c, got_n = compute_genotype_counts_for_pairs(torch.as_tensor(g, device=cuda), ii, jj)
    ref_c, ref_n = o.geno_counts(g, ii, jj)
    np.testing.assert_array_equal(got_c.cpu().numpy(), ref_c)
    assert (got_n is None) == (ref_n is None)
    if ref_n is not None:
        np.testing.assert_array_equal(got_n.cpu().numpy(), ref_n)


def _close(a, b, rtol=1e-12):
    a = a.cpu().numpy() if hasattr(a, "cpu") else np.asarray(a)
    np.testing.assert_allclose(a, b, rtol=rtol, atol=1e-300)


def test_per_pair_estimators(cuda):
    from pg_gpu_amd import ld_statistics as ls

    rng = np.random.default_rng(0)
    P = 4
    cs = [rng.integers(0, 30, (3000, 4)) for _ in range(P)]
    cs[0][:20] = rng.integers(0, 2, (20, 4))       # tiny samples: guard paths
    counts = np.concatenate(cs, axis=1)
    pops = [o.HapPop(c) for c in cs]
    _close(ls.dd(cs[0]), o.hap_dd(pops, 0, 0))
    _close(ls.dz(cs[0]), o.hap_dz(pops, 0, 0, 0))
    _close(ls.pi2(cs[0]), o.hap_pi2(pops, 0, 0, 0, 0))
    for pp in [(0, 1), (2, 3), (1, 1)]:
        _close(ls.dd(counts, pp), o.hap_dd(pops, *pp))
    for pp in [(0, 0, 1), (0, 1, 0), (0, 1, 1), (2, 1, 3), (3, 3, 3)]:
        _close(ls.dz(counts, pp), o.hap_dz(pops, *pp))
    for pp in [(0, 0, 0, 1), (0, 1, 1, 1), (0, 1, 0, 1), (0, 0, 2, 3), (0, 1, 2, 2),
               (0, 1, 0, 2), (1, 2, 3, 1), (0, 1, 2, 3)]:
        _close(ls.pi2(counts, pp), o.hap_pi2(pops, *pp))
    # pi2(i,i,k,k) averages both orientations
    _close(ls.pi2(counts, (0, 0, 1, 1)),
           0.5 * (o.hap_pi2(pops, 0, 0, 1, 1) + o.hap_pi2(pops, 1, 1, 0, 0)))
    # explicit n_valid overrides the count sums
    nv = cs[1].sum(axis=1) + 3
    _close(ls.dd(cs[1], n_valid=nv), o.hap_dd([o.HapPop(cs[1], nv)], 0, 0))


def test_r_family(cuda):
    from pg_gpu_amd import ld_statistics as ls

    c = np.array([[5, 0, 0, 5], [25, 25, 25, 25], [0, 5, 5, 0], [5, 5, 0, 0]], np.int32)
    r2 = ls.r_squared(c).cpu().numpy()
    np.testing.assert_allclose(r2[[0, 2]], 1.0, rtol=1e-10)
    assert abs(r2[1]) < 1e-10 and np.isnan(r2[3])
    np.testing.assert_allclose(ls.r(c).cpu().numpy()[2], -1.0, rtol=1e-10)
    rng = np.random.default_rng(5)
    c = rng.integers(0, 40, (1000, 4))
    ref = o.r_family(c)
    for got, exp in zip((ls.r(c), ls.r_squared(c), ls.d_prime(c)), ref):
        g = got.cpu().numpy()
        assert np.array_equal(np.isnan(g), np.isnan(exp))
        ok = ~np.isnan(exp)
        np.testing.assert_allclose(g[ok], exp[ok], rtol=1e-12)


def test_two_pop_batch(cuda):
    from pg_gpu_amd.ld_pipeline import (compute_single_pop_statistics_batch,
                                        compute_two_pop_statistics_batch)

    rng = np.random.default_rng(8)
    a, b = rng.integers(0, 50, (2000, 4)), rng.integers(0, 50, (2000, 4))
    nva, nvb = a.sum(1), b.sum(1)
    got = compute_two_pop_statistics_batch(a, b, nva, nvb).cpu().numpy()
    ref = o.lds_two_pop_15(a, b, nva, nvb)
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-300)
    got1 = compute_single_pop_statistics_batch(a, None).cpu().numpy()
    np.testing.assert_allclose(got1, o.multi_pop_stats_hap([a], [None]), rtol=1e-12, atol=1e-300)


def test_moments_golden_end_to_end(cuda):
    from pg_gpu_amd.moments_ld import compute_ld_statistics

    fx = json.loads((GOLD / "moments_im_2pop.json").read_text())
    out = compute_ld_statistics(str(GOLD / "im-parsing-example.vcf.gz"),
                                pop_file=str(GOLD / "im_pop.txt"), pops=fx["pops"],
                                bp_bins=fx["bp_bins"], report=False)
    assert out["stats"] == (fx["ld_names"], fx["het_names"])
    assert out["bins"] == [tuple(b) for b in fx["bins"]]
    got = np.array(out["sums"][:-1])
    ref = np.array(fx["sums"])
    assert np.max(np.abs(got - ref) / np.abs(ref)) < 1e-12
    np.testing.assert_allclose(out["sums"][-1], fx["het"], rtol=1e-13)


def test_haplotype_matrix_two_pops_raw_and_mean(cuda):
    from pg_gpu_amd.haplotype_matrix import HaplotypeMatrix

    fx = json.loads((GOLD / "moments_im_2pop.json").read_text())
    hm = HaplotypeMatrix.from_vcf(str(GOLD / "im-parsing-example.vcf.gz"))
    hm.load_pop_file(str(GOLD / "im_pop.txt"), pops=fx["pops"])
    raw = hm.compute_ld_statistics_gpu_two_pops(fx["bp_bins"], "deme0", "deme1", raw=True)
    mean = hm.compute_ld_statistics_gpu_two_pops(fx["bp_bins"], "deme0", "deme1", raw=False)
    for k, (lo, hi) in enumerate(fx["bins"]):
        r = np.array(list(raw[(lo, hi)].values()))
        assert np.max(np.abs(r - fx["sums"][k]) / np.abs(fx["sums"][k])) < 1e-12
        m = np.array(list(mean[(lo, hi)].values()))
        ratio = r / m
        assert np.allclose(ratio, ratio[0], rtol=1e-12) and ratio[0] > 1
    single = hm.compute_ld_statistics_gpu_single_pop(fx["bp_bins"], raw=True)
    assert len(single) == 3 and all(len(v) == 3 for v in single.values())


def _rand_geno_counts(rng, N, n_ind, missing):
    g = rng.integers(0, 3, (n_ind, 2 * N)).astype(np.int8)
    if missing:
        g[rng.random(g.shape) < missing] = -1
    c, nv = o.geno_counts(g, np.arange(N), np.arange(N, 2 * N))
    return c, nv


@pytest.mark.parametrize("P", [1, 2, 3])
def test_genotype_batch_and_raw_calls(cuda, P):
    """compute_all_{dd,dz,pi2}_geno and compute_multi_pop_statistics_batch_geno
    against the oracle's kernel form (genotype_kernels.py:561-936)."""
    from oracle import geno_oracle as go
    from pg_gpu_amd import genotype_kernels as gk
    from pg_gpu_amd.ld_pipeline import generate_stat_specs

    rng = np.random.default_rng(50 + P)
    sizes = [3, 9, 25][:P] if P == 3 else [14, 30][:P]
    cl, nl = zip(*[_rand_geno_counts(rng, 400, k, 0.06) for k in sizes])
    ref_pops = [go.GenoPop(c, nv) for c, nv in zip(cl, nl)]
    specs = generate_stat_specs(P)
    def same(got, ref):
        # where the reference divides by a non-positive falling factorial
        # (too few valid individuals) it yields inf/NaN; the build gives 0
        fin = np.isfinite(ref)
        np.testing.assert_allclose(got[fin], ref[fin], rtol=1e-10, atol=1e-13)
        assert np.all(got[~fin] == 0.0)

    got = gk.compute_multi_pop_statistics_batch_geno(list(cl), list(nl), None, specs)
    with np.errstate(invalid="ignore"):
        ref = go.multi_pop_stats_geno(list(cl), list(nl))
    same(got.cpu().numpy(), ref)
    # a custom spec list goes through the raw-call path
    sub = specs[::3]
    got2 = gk.compute_multi_pop_statistics_batch_geno(list(cl), list(nl), None, sub).cpu().numpy()
    same(got2, ref[:, ::3])
    pops = [gk.PopDataGeno(c, nv) for c, nv in zip(cl, nl)]
    for kind, fn in (("dd", gk.compute_all_dd_geno), ("dz", gk.compute_all_dz_geno),
                     ("pi2", gk.compute_all_pi2_geno)):
        calls = sorted({pp for _, cs in specs for _, st, pp in cs if st == kind})
        for pp, v in zip(calls, fn(pops, calls)):
            same(v.cpu().numpy(), go.kernel_call(ref_pops, kind, pp))


def test_haplotype_batch_custom_specs(cuda):
    from pg_gpu_amd import haplotype_kernels as hk
    from pg_gpu_amd.ld_pipeline import generate_stat_specs

    rng = np.random.default_rng(60)
    h = (rng.random((60, 400)) < 0.4).astype(np.int8)
    h[rng.random(h.shape) < 0.05] = -1
    ii, jj = np.arange(200), np.arange(200, 400)
    rows = [range(0, 12), range(12, 35), range(35, 60)]
    cl, nl = zip(*[o.hap_counts(h, ii, jj, list(r)) for r in rows])
    specs = generate_stat_specs(3)
    ref = o.multi_pop_stats_hap(list(cl), list(nl))
    got = hk.compute_multi_pop_statistics_batch_hap(list(cl), list(nl), None, specs)
    np.testing.assert_allclose(got.cpu().numpy(), ref, rtol=1e-10, atol=1e-13)
    sub = specs[1::4]
    got2 = hk.compute_multi_pop_statistics_batch_hap(list(cl), list(nl), None, sub)
    np.testing.assert_allclose(got2.cpu().numpy(), ref[:, 1::4], rtol=1e-10, atol=1e-13)


def test_sigma_d2_geno(cuda):
    """DD/pi2 of one population's genotype counts (ld_statistics_genotype.py:1518-1598)."""
    from oracle import geno_oracle as go
    from pg_gpu_amd.genotype_matrix import GenotypeMatrix
    from pg_gpu_amd.ld_statistics_genotype import sigma_d2_geno

    rng = np.random.default_rng(61)
    g = rng.integers(0, 3, (30, 40)).astype(np.int8)
    g[:, 3] = 1                                   # monomorphic dosage: pi2 <= 0 -> NaN
    gm = GenotypeMatrix(g, np.arange(40) * 10)
    got = sigma_d2_geno(gm).cpu().numpy()
    ii, jj = np.triu_indices(40, 1)
    c, _ = o.geno_counts(g, ii, jj)
    pop = go.GenoPop(c)
    dd, pi2 = go.kernel_call([pop], "dd", (0, 0)), go.kernel_call([pop], "pi2", (0, 0, 0, 0))
    with np.errstate(invalid="ignore", divide="ignore"):
        ref = np.where(pi2 > 0, dd / pi2, np.nan)
    np.testing.assert_array_equal(np.isnan(got), np.isnan(ref))
    ok = ~np.isnan(ref)
    np.testing.assert_allclose(got[ok], ref[ok], rtol=1e-10)
    g2 = g.copy()
    g2[0, 0] = -1
    with pytest.raises(ValueError):
        sigma_d2_geno(GenotypeMatrix(g2, np.arange(40) * 10))
