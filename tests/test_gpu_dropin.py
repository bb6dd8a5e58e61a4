"""GPU parity of the reference's drop-in entry points that wrap the fused
kernel: moments_ld.compute_ld_statistics with genotypes (union biallelic
filter + genotype heterozygosity) and with recombination-distance bins,
HaplotypeMatrix.compute_ld_statistics_gpu_single_pop, the GenotypeMatrix
conversions and filters, the all-pairs tallies, from_vcf subsetting, and the
error mapping of the C ABI -- each against the oracle."""
import ctypes
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import ld_oracle as o

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"


def founder_haps(seed, n_hap, V, span, missing=0.0, n_founders=12):
    from pg_gpu_amd import synth

    rng = np.random.default_rng(seed)
    pos = synth._positions(rng, V, span)
    p = rng.beta(0.4, 0.4, V)
    hap = synth._copy(rng, synth._founders(rng, n_founders, p), pos, n_hap, 2e-5)
    if missing:
        synth._missing(rng, hap, missing)
    return hap, pos


def assert_sums_close(got, ref, absr, tol=1e-12):
    got, ref, absr = (np.asarray(x, np.float64) for x in (got, ref, absr))
    scale = np.maximum(np.abs(ref), 1e-3 * absr)
    err = np.abs(got - ref) / np.maximum(scale, 1e-300)
    assert float(err.max(initial=0.0)) < tol, float(err.max())


def check_moments_dict(got, ref, P):
    assert got["stats"] == ref["stats"]
    assert got["bins"] == ref["bins"]
    nb = len(ref["bins"])
    assert len(got["sums"]) == nb + 1
    assert_sums_close(np.array(got["sums"][:nb]), np.array(ref["sums"][:nb]), ref["abs"][:, :])
    np.testing.assert_allclose(got["sums"][nb], ref["sums"][nb], rtol=1e-12, equal_nan=True)
    assert len(got["sums"][nb]) == P * (P + 1) // 2


@pytest.mark.parametrize("P", [1, 2, 3, 4])
def test_compute_ld_statistics_genotypes(cuda, P):
    """use_genotypes=True: union biallelic filter, nine-count estimators, and
    heterozygosity over the unfiltered genotypes (moments_ld.py:107-203,
    :256-388), with missing calls and overlapping populations."""
    from pg_gpu_amd.genotype_matrix import GenotypeMatrix
    from pg_gpu_amd.moments_ld import compute_ld_statistics

    hap, pos = founder_haps(100 + P, 2 * 24 * P, 700, 400_000)
    g = o.hap_to_geno(hap)
    rng = np.random.default_rng(P)
    g[rng.random(g.shape) < 0.04] = -1
    g[:, 10:20] = 0                         # monomorphic: removed by the union filter
    g[:, 30] = -1
    g[0, 30] = 1                            # one valid individual: removed, het NaN for others
    sizes = [24] * P
    offs = np.cumsum([0] + sizes)
    pops = [list(range(offs[k], offs[k + 1])) for k in range(P)]
    if P > 1:
        pops[1] = pops[1] + [0]             # an individual in two populations
    names = [f"p{k}" for k in range(P)]
    gm = GenotypeMatrix(g, pos, sample_sets=dict(zip(names, pops)))
    bins = np.logspace(2, np.log10(50_000), 7)
    got = compute_ld_statistics(genotype_matrix=gm, pops=names, bp_bins=bins,
                                use_genotypes=True, report=False)
    ref = o.moments_ld_stats(g, pos, pops, "geno", bp_bins=bins)
    check_moments_dict(got, ref, P)


def test_compute_ld_statistics_genotypes_from_haplotypes(cuda):
    """The haplotype_matrix= route of use_genotypes (GenotypeMatrix
    .from_haplotype_matrix on the device)."""
    from pg_gpu_amd.haplotype_matrix import HaplotypeMatrix
    from pg_gpu_amd.moments_ld import compute_ld_statistics

    hap, pos = founder_haps(7, 80, 500, 300_000, missing=0.03)
    sets = {"a": list(range(0, 40)), "b": list(range(40, 80))}
    hm = HaplotypeMatrix(hap, pos, sample_sets=sets)
    bins = np.logspace(2, 4.5, 6)
    got = compute_ld_statistics(haplotype_matrix=hm, pops=["a", "b"], bp_bins=bins,
                                use_genotypes=True, report=False)
    g = o.hap_to_geno(hap)
    ref = o.moments_ld_stats(g, pos, [list(range(0, 20)), list(range(20, 40))], "geno",
                             bp_bins=bins)
    check_moments_dict(got, ref, 2)


@pytest.mark.parametrize("use_genotypes", [False, True])
def test_compute_ld_statistics_r_bins(cuda, use_genotypes, tmp_path):
    """r_bins with a recombination map: Morgan distances interpolated from the
    map, the bp cutoff from the smallest local rate (moments_ld.py:166-245)."""
    from pg_gpu_amd.moments_ld import compute_ld_statistics

    fx = json.loads((GOLD / "moments_im_2pop.json").read_text())
    r_bins = [0.0, 1e-6, 5e-6, 2e-5, 1e-4]
    kw = dict(vcf_file=str(GOLD / "im-parsing-example.vcf.gz"),
              rec_map_file=str(GOLD / "rec_map_im.txt"), pop_file=str(GOLD / "im_pop.txt"),
              pops=["deme0", "deme1"], r_bins=r_bins, report=False, use_genotypes=use_genotypes)
    got = compute_ld_statistics(**kw)
    from tests.test_oracle_golden import load_im

    hap, pos, pops = load_im()
    if use_genotypes:
        from pg_gpu_amd.vcf import read_vcf_gt

        d = read_vcf_gt(GOLD / "im-parsing-example.vcf.gz")
        gt = d["gt"]
        n0 = np.sum(gt == 0, axis=(1, 2))
        n1 = np.sum(gt == 1, axis=(1, 2))
        keep = (n0 > 0) & (n1 > 0) & (np.sum(gt >= 2, axis=(1, 2)) == 0)
        g = np.where((gt[keep] >= 0).all(axis=2), gt[keep].sum(axis=2), -1).astype(np.int8).T
        keep_g = o.biallelic_filter_geno(g)
        n = g.shape[0]
        gpops = [sorted(set(r % n for r in p)) for p in pops]
        ref = o.moments_ld_stats(g[:, keep_g], d["pos"][keep][keep_g], gpops, "geno",
                                 r_bins=r_bins, rec_map_file=GOLD / "rec_map_im.txt")
    else:
        ref = o.moments_ld_stats(hap, pos, pops, "hap", r_bins=r_bins,
                                 rec_map_file=GOLD / "rec_map_im.txt")
    assert ref["counts"].sum() > 1000 and np.all(ref["counts"] > 0)
    check_moments_dict(got, ref, 2)
    assert fx["ld_names"] == list(got["stats"][0])


def test_single_pop_values(cuda):
    """compute_ld_statistics_gpu_single_pop: sums with raw=True, per-pair
    means otherwise, zeros for empty bins (haplotype_matrix.py:1592-1716)."""
    from pg_gpu_amd.haplotype_matrix import HaplotypeMatrix

    hap, pos = founder_haps(11, 60, 400, 200_000, missing=0.02)
    hap[:, 5] = 1                        # removed by the biallelic filter
    hm = HaplotypeMatrix(hap, pos)
    bins = [0.5, 1, 100, 1_000, 10_000, 60_000]   # distances are >= 1: the first bin is empty
    keep = o.biallelic_filter_hap(hap)
    ref, cnt, _, absr = o.binned_sums_hap(hap[:, keep], pos[keep], [list(range(60))], bins,
                                          with_abs=True)
    assert cnt[0] == 0 and cnt[-1] > 0
    raw = hm.compute_ld_statistics_gpu_single_pop(bins, raw=True)
    mean = hm.compute_ld_statistics_gpu_single_pop(bins)
    keys = [(float(a), float(b)) for a, b in zip(bins[:-1], bins[1:])]
    assert list(raw) == keys and list(mean) == keys
    got_raw = np.array([raw[k] for k in keys])
    got_mean = np.array([mean[k] for k in keys])
    assert_sums_close(got_raw, ref, absr)
    for b, k in enumerate(keys):
        if cnt[b] == 0:
            assert mean[k] == (0.0, 0.0, 0.0)
        else:
            assert_sums_close(got_mean[b] * cnt[b], ref[b], absr[b])


def test_genotype_matrix_conversions_and_filter(cuda):
    import torch

    from pg_gpu_amd.genotype_matrix import GenotypeMatrix
    from pg_gpu_amd.haplotype_matrix import HaplotypeMatrix

    rng = np.random.default_rng(21)
    hap = rng.integers(-1, 2, (30, 257)).astype(np.int8)
    sets = {"x": [0, 1, 7], "y": [12, 13, 29]}
    for dev in ("cpu", "gpu"):
        hm = HaplotypeMatrix(hap, np.arange(257) * 10 + 5, sample_sets=sets)
        if dev == "gpu":
            hm.transfer_to_gpu()
        gm = GenotypeMatrix.from_haplotype_matrix(hm)
        assert gm.device == hm.device
        g = gm.genotypes.cpu().numpy() if dev == "gpu" else gm.genotypes
        np.testing.assert_array_equal(g, o.hap_to_geno(hap))
        assert gm.sample_sets == {"x": [0, 3], "y": [6, 14]}
        back = gm.to_haplotype_matrix()
        h2 = back.haplotypes.cpu().numpy() if dev == "gpu" else back.haplotypes
        np.testing.assert_array_equal(h2, o.geno_to_hap(o.hap_to_geno(hap)))
    g = rng.integers(0, 3, (25, 300)).astype(np.int8)
    g[:, :30] = 2
    g[rng.random(g.shape) < 0.3] = -1
    g[:24, 100] = -1
    gm = GenotypeMatrix(torch.as_tensor(g, device=cuda), np.arange(300) + 1)
    f = gm.apply_biallelic_filter()
    keep = o.biallelic_filter_geno(g)
    assert 0 < len(keep) < 300
    np.testing.assert_array_equal(f.genotypes.cpu().numpy(), g[:, keep])
    np.testing.assert_array_equal(f.positions.cpu().numpy(), keep + 1)


def test_haplotype_biallelic_filter_on_device(cuda):
    import torch

    from pg_gpu_amd.haplotype_matrix import HaplotypeMatrix

    rng = np.random.default_rng(22)
    hap = rng.integers(-1, 2, (40, 500)).astype(np.int8)
    hap[:, :20] = 0
    hap[3, 50:60] = 2                   # a third allele: removed
    hm = HaplotypeMatrix(torch.as_tensor(hap, device=cuda), np.arange(500) * 3 + 1)
    f = hm.apply_biallelic_filter()
    keep = o.biallelic_filter_hap(hap)
    np.testing.assert_array_equal(f.haplotypes.cpu().numpy(), hap[:, keep])
    assert f.chrom_start == int(keep[0] * 3 + 1) and f.chrom_end == int(keep[-1] * 3 + 1)


def test_genotype_from_vcf_and_haplotype_subsets(cuda):
    from pg_gpu_amd.genotype_matrix import GenotypeMatrix
    from pg_gpu_amd.haplotype_matrix import HaplotypeMatrix
    from pg_gpu_amd.vcf import read_vcf_gt

    path = GOLD / "im-parsing-example.vcf.gz"
    d = read_vcf_gt(path)
    gt = d["gt"]
    keep = (np.sum(gt == 0, axis=(1, 2)) > 0) & (np.sum(gt == 1, axis=(1, 2)) > 0) & \
        (np.sum(gt >= 2, axis=(1, 2)) == 0)     # is_biallelic_01
    g = np.where((gt[keep] >= 0).all(axis=2), gt[keep].sum(axis=2), -1).astype(np.int8).T
    gm = GenotypeMatrix.from_vcf(str(path))
    np.testing.assert_array_equal(gm.genotypes, g)
    # region (1-based inclusive) and a sample subset in file order
    names = [d["samples"][7], d["samples"][2]]
    hm = HaplotypeMatrix.from_vcf(str(path), region="1:100000-300000", samples=names)
    sel = (d["pos"] >= 100_000) & (d["pos"] <= 300_000)
    assert hm.samples == [d["samples"][2], d["samples"][7]]
    assert (hm.chrom_start, hm.chrom_end) == (100_000, 300_000)
    exp = np.concatenate([gt[sel][:, [2, 7], 0].T, gt[sel][:, [2, 7], 1].T])
    np.testing.assert_array_equal(hm.haplotypes, exp)
    with pytest.raises(ValueError, match="No variants"):
        HaplotypeMatrix.from_vcf(str(path), region="2:1-10")


@pytest.mark.parametrize("missing", [0.0, 0.05])
def test_tally_gpu_haplotypes(cuda, missing):
    from pg_gpu_amd.haplotype_matrix import HaplotypeMatrix

    hap, pos = founder_haps(31, 150, 101, 50_000, missing=missing)
    sets = {"a": list(range(0, 130)), "b": list(range(130, 150))}
    hm = HaplotypeMatrix(hap, pos, sample_sets=sets)
    c, nv = hm.tally_gpu_haplotypes()
    rc, rn = o.tally_upper(hap)
    np.testing.assert_array_equal(c.cpu().numpy(), rc)
    assert (nv is None) == (rn is None)
    if rn is not None:
        np.testing.assert_array_equal(nv.cpu().numpy(), rn)
    c, nv = hm.tally_gpu_haplotypes(pop="a")
    rc, rn = o.tally_upper(hap, sets["a"])
    np.testing.assert_array_equal(c.cpu().numpy(), rc)   # n11 up to 130 (> 127)
    c2, n1, n2 = hm.tally_gpu_haplotypes_two_pops("a", "b")
    rb, rbn = o.tally_upper(hap, sets["b"])
    np.testing.assert_array_equal(c2.cpu().numpy(), np.concatenate([rc, rb], axis=1))
    if missing:
        np.testing.assert_array_equal(n1.cpu().numpy(), rn)
        np.testing.assert_array_equal(n2.cpu().numpy(), rbn)
    else:
        assert n1 is None and n2 is None
    c3, m1, m2 = hm.tally_gpu_haplotypes_two_pops_with_missing("a", "b")
    np.testing.assert_array_equal(c3.cpu().numpy(), c2.cpu().numpy())
    np.testing.assert_array_equal(m1.cpu().numpy(), rc.sum(axis=1))
    with pytest.raises(KeyError):
        hm.tally_gpu_haplotypes(pop="zz")
    with pytest.raises(ValueError):
        HaplotypeMatrix(hap, pos).tally_gpu_haplotypes(pop="a")


def test_compute_ld_statistics_per_pair_wrapper(cuda):
    """ld_statistics.compute_ld_statistics (ld_statistics.py:1030-1076)."""
    import torch

    from pg_gpu_amd import ld_statistics as ls

    rng = np.random.default_rng(40)
    counts = rng.integers(0, 30, (500, 8)).astype(np.int32)
    counts[:5] = [3, 0, 0, 4, 1, 1, 1, 1]
    cnt = torch.as_tensor(counts, device=cuda)
    got = ls.compute_ld_statistics(cnt, ["dd", "dz", "pi2", "r", "r_squared"],
                                   populations={"dd": (0, 1), "dz": (1, 0, 1),
                                                "pi2": (0, 1, 0, 1)})
    cl = [counts[:, :4], counts[:, 4:]]
    nl = [None, None]
    want = {k: o.multi_pop_stats_hap(cl, nl, [(k, [(1.0, k, pp)])])[:, 0]
            for k, pp in (("dd", (0, 1)), ("dz", (1, 0, 1)), ("pi2", (0, 1, 0, 1)))}
    r, r2, _ = o.r_family(counts[:, :4])
    want["r"], want["r_squared"] = r, r2
    assert set(got) == set(want)
    for k, v in want.items():
        np.testing.assert_allclose(got[k].cpu().numpy(), v, rtol=1e-12, atol=1e-300,
                                   equal_nan=True)
    with pytest.raises(ValueError, match="Unknown statistic"):
        ls.compute_ld_statistics(cnt, ["nope"])


def test_invalid_argument_raises_value_error_through_the_abi(cuda):
    """A bad argument returns PGLD_EINVAL (1) from the library, which the
    binding raises as ValueError; an unbuilt case returns PGLD_EUNSUPPORTED."""
    import torch

    from pg_gpu_amd import _lib
    from pg_gpu_amd.engine import LDPanels

    pan = LDPanels(torch.zeros((8, 64), dtype=torch.int8, device=cuda), [list(range(8))])
    out = torch.zeros((1, 5, 3), dtype=torch.int64, device=cuda)
    plan = pan.plan(np.arange(64) * 10.0, [1.0, 100.0])
    a = plan.args(out)
    a.n_bins = 0
    rc = _lib.lib().pgld_binned_fused(ctypes.byref(pan.desc), ctypes.byref(a),
                                      _lib.stream_handle())
    assert rc == _lib.PGLD_EINVAL == 1
    assert "n_bins" in _lib.lib().pgld_last_error().decode()
    with pytest.raises(ValueError):
        _lib.check(rc)
    a.n_bins = 1
    a.with_r2 = 1
    d2 = _lib.PanelDesc()
    ctypes.memmove(ctypes.byref(d2), ctypes.byref(pan.desc), ctypes.sizeof(d2))
    d2.n_pops = 2
    d2.pop_size[1] = 8
    rc = _lib.lib().pgld_binned_fused(ctypes.byref(d2), ctypes.byref(a), _lib.stream_handle())
    assert rc == _lib.PGLD_EUNSUPPORTED
    with pytest.raises(NotImplementedError):
        _lib.check(rc)
    with pytest.raises(ValueError, match="values must be"):
        LDPanels(torch.full((4, 40), 3, dtype=torch.int8, device=cuda), [[0, 1, 2, 3]])


def test_launches_follow_the_current_stream(cuda):
    """Every launch, fill and read uses torch's current stream: a pass run
    under a side stream gives the default stream's result bit for bit."""
    import torch

    from pg_gpu_amd.engine import LDPanels

    hap, pos = founder_haps(50, 96, 900, 600_000, missing=0.05)
    bins = np.logspace(2, 5, 9)
    ref = LDPanels(torch.as_tensor(hap, device=cuda), [list(range(48)), list(range(48, 96))])
    r0 = ref.binned(pos, bins)
    side = torch.cuda.Stream(device=cuda)
    with torch.cuda.stream(side):
        pan = LDPanels(torch.as_tensor(hap, device=cuda), [list(range(48)), list(range(48, 96))])
        r1 = pan.binned(pos, bins)
    np.testing.assert_array_equal(r1.sums, r0.sums)
    np.testing.assert_array_equal(r1.counts, r0.counts)


@pytest.mark.parametrize("mode,offset", [("hap", 0), ("hap", 3), ("geno", 0), ("geno", 7)])
def test_scan_sites_tallies(cuda, mode, offset):
    """pgld_scan_sites against NumPy: per population and site the counts of
    1, 2 and -1, on a column view that starts off a 16-byte boundary, with
    populations of more than 255 rows (the per-byte counters' flush) and a
    site count that is not a multiple of 16."""
    import torch

    from pg_gpu_amd.engine import LDPanels

    rng = np.random.default_rng(60 + offset)
    hi = 2 if mode == "hap" else 3
    full = rng.integers(-1, hi, (700, 1029 + offset)).astype(np.int8)
    mat = full[:, offset:]
    pops = [list(range(0, 600, 2)), list(range(1, 700, 2))[:290], [5, 699, 3]]
    dev = torch.as_tensor(full, device=cuda)[:, offset:]
    pan = LDPanels(dev, pops, mode, pack=False)
    got = pan.site_stats.cpu().numpy()
    for p, rows in enumerate(pops):
        x = mat[rows]
        np.testing.assert_array_equal(got[p, 0], (x == 1).sum(axis=0))
        np.testing.assert_array_equal(got[p, 1], (x == 2).sum(axis=0))
        np.testing.assert_array_equal(got[p, 2], (x == -1).sum(axis=0))
    bad = full.copy()
    bad[699, offset + 1028] = 2 if mode == "hap" else 3     # last site, last population
    with pytest.raises(ValueError):
        LDPanels(torch.as_tensor(bad, device=cuda)[:, offset:], pops, mode, pack=False)
