"""GPU parity of the dense tile consumers (pgld_pair_block: sigma_D^2, frequency
r^2, dosage r^2; ZnS and omega) against the oracle (ld_statistics.py:249-1021)."""
import numpy as np
import pytest

from oracle import ld_oracle as o

pytestmark = pytest.mark.gpu


def _hm(h):
    from pg_gpu_amd.haplotype_matrix import HaplotypeMatrix

    return HaplotypeMatrix(h, np.arange(h.shape[1]) * 100 + 1)


def _gm(g):
    from pg_gpu_amd.genotype_matrix import GenotypeMatrix

    return GenotypeMatrix(g, np.arange(g.shape[1]) * 100 + 1)


def _haps(seed, n, m, missing):
    rng = np.random.default_rng(seed)
    fr = rng.beta(0.5, 0.5, m)
    h = (rng.random((n, m)) < fr).astype(np.int8)
    h[:, 3] = 0                                   # monomorphic: dropped as non-segregating
    if missing:
        h[rng.random(h.shape) < missing] = -1
    return h


@pytest.mark.parametrize("missing", [0.0, 0.05])
def test_pair_matrices_match_oracle(cuda, missing):
    import torch

    from pg_gpu_amd.ld_tiles import _pair_matrix
    from pg_gpu_amd import _lib

    h = _haps(80, 70, 150, missing)
    print(f"case: pair matrices {h.shape} missing {missing}", flush=True)
    t = torch.as_tensor(h, device=cuda)
    val, flag = _pair_matrix(t, _lib.PAIR_SIGMA_D2, "hap")
    rv, rf = o.sigma_d2_matrix(h)
    np.testing.assert_allclose(val.cpu().numpy(), rv, rtol=1e-12, atol=1e-15)
    np.testing.assert_array_equal(flag.cpu().numpy(), rf)
    r2, _ = _pair_matrix(t, _lib.PAIR_R2_NAIVE, "hap")
    np.testing.assert_allclose(r2.cpu().numpy(), o.r2_naive_matrix(h), rtol=1e-12, atol=1e-15)
    g = (h[0::2].astype(np.int16) + h[1::2]).astype(np.int8)
    g[(h[0::2] < 0) | (h[1::2] < 0)] = -1
    d, _ = _pair_matrix(torch.as_tensor(g, device=cuda), _lib.PAIR_R2_DOSAGE, "geno")
    np.testing.assert_allclose(d.cpu().numpy(), o.r2_dosage_matrix(g), rtol=1e-10, atol=1e-14)


@pytest.mark.parametrize("missing", [0.0, 0.05])
def test_zns_and_omega(cuda, missing):
    from pg_gpu_amd.ld_tiles import omega, zns

    h = _haps(81, 60, 90, missing)
    hs = o.segregating(h)
    sig, flag = o.sigma_d2_matrix(hs)
    assert zns(_hm(h)) == pytest.approx(o.zns_value(sig, flag), rel=1e-12)
    assert omega(_hm(h)) == pytest.approx(o.omega_value(sig), rel=1e-10)
    r2 = o.r2_naive_matrix(hs)
    assert zns(_hm(h), estimator="r2") == pytest.approx(o.zns_value(r2), rel=1e-12)
    assert omega(_hm(h), estimator="r2") == pytest.approx(o.omega_value(r2), rel=1e-10)
    he = o.segregating(h, "exclude")
    assert zns(_hm(h), "exclude", "r2") == pytest.approx(o.zns_value(o.r2_naive_matrix(he)),
                                                         rel=1e-12)
    g = (h[0::2].astype(np.int16) + np.maximum(h[1::2], 0)).astype(np.int8)
    g[(h[0::2] < 0) | (h[1::2] < 0)] = -1
    rd = o.r2_dosage_matrix(g)
    assert zns(_gm(g)) == pytest.approx(o.zns_value(rd), rel=1e-10)
    assert omega(_gm(g)) == pytest.approx(o.omega_value(rd), rel=1e-9)
    with pytest.raises(ValueError):
        zns(_gm(g), estimator="sigma_d2")


def test_pairwise_r2_method(cuda):
    h = _haps(82, 40, 70, 0.0)
    got = _hm(h).pairwise_r2().cpu().numpy()
    np.testing.assert_allclose(got, o.r2_naive_matrix(h), rtol=1e-12, atol=1e-15)
    rh = _hm(h).pairwise_r2("rogers_huff").cpu().numpy()
    g = (h[0::2].astype(np.int16) + h[1::2]).astype(np.int8)
    np.testing.assert_allclose(rh, o.r2_dosage_matrix(g), rtol=1e-10, atol=1e-14)


@pytest.mark.parametrize("missing", [0.0, 0.05])
def test_windowed_r_squared(cuda, missing):
    from pg_gpu_amd.haplotype_matrix import HaplotypeMatrix

    h = _haps(83, 50, 120, missing)
    pos = np.sort(np.random.default_rng(5).choice(np.arange(1, 20_000), 120, replace=False))
    hm = HaplotypeMatrix(h, pos)
    hm.sample_sets = {"a": list(range(0, 30)), "b": list(range(30, 50))}
    bins = [0, 500, 2_000, 8_000, 30_000]
    got, cnt = hm.windowed_r_squared(bins, percentile=[10, 50, 90])
    ref, rcnt = o.windowed_r_squared(h, pos, bins, [10, 50, 90])
    np.testing.assert_array_equal(cnt, rcnt)
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-15)
    got_a, cnt_a = hm.windowed_r_squared(bins, pop="a")
    ref_a, rcnt_a = o.windowed_r_squared(h, pos, bins, 50, rows=list(range(30)))
    np.testing.assert_array_equal(cnt_a, rcnt_a)
    np.testing.assert_allclose(got_a, ref_a, rtol=1e-12, atol=1e-15)


def test_windowed_r_squared_rogers_huff(cuda):
    from pg_gpu_amd.haplotype_matrix import HaplotypeMatrix

    h = _haps(84, 40, 100, 0.0)
    h[:, 7] = 0  # a monomorphic site: its pairs are left out
    pos = np.arange(100) * 37 + 5
    hm = HaplotypeMatrix(h, pos)
    bins = [0, 300, 1_000, 4_000]
    got, cnt = hm.windowed_r_squared(bins, percentile=[25, 75], estimator="rogers_huff")
    ref, rcnt = o.windowed_r_squared(h, pos, bins, [25, 75], estimator="rogers_huff")
    np.testing.assert_array_equal(cnt, rcnt)
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-15)
