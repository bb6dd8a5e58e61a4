"""CPU checks that pin the oracle's container passes and the moments driver's
host steps: hap <-> geno conversion, the genotype and union biallelic filters,
genotype heterozygosity, the recombination-map interpolation and bp cutoff,
the dosage r^2 missing-call convention and the all-pairs tallies.  Each is
checked against brute force or a closed form, not against its own formula."""
import itertools
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import ld_oracle as o

GOLD = Path(__file__).resolve().parent / "golden"


def test_hap_to_geno_brute_force_and_round_trip():
    rng = np.random.default_rng(1)
    h = rng.integers(-1, 2, (10, 37)).astype(np.int8)
    g = o.hap_to_geno(h)
    for i in range(5):
        for s in range(37):
            a, b = int(h[2 * i, s]), int(h[2 * i + 1, s])
            assert g[i, s] == (-1 if a < 0 or b < 0 else a + b)
    # genotype -> two haplotypes -> genotype is the identity
    gg = rng.integers(-1, 3, (7, 29)).astype(np.int8)
    np.testing.assert_array_equal(o.hap_to_geno(o.geno_to_hap(gg)), gg)
    hh = o.geno_to_hap(gg)
    assert np.all(hh[0::2][gg == 1] == 1) and np.all(hh[1::2][gg == 1] == 0)   # hets as (1, 0)


def test_geno_filters_brute_force():
    rng = np.random.default_rng(2)
    g = rng.integers(0, 3, (12, 300)).astype(np.int8)
    g[:, :40] = 0                                   # monomorphic reference
    g[:, 40:80] = 2                                 # monomorphic alternative
    g[rng.random(g.shape) < 0.3] = -1
    g[:11, 80:90] = -1                              # one valid individual
    pops = [[0, 1, 2, 5], [5, 6, 7], [9, 10]]       # individual 5 in two populations
    for rows in (None, [r for p in pops for r in p]):
        keep = o.biallelic_filter_geno(g, rows)
        sel = list(range(12)) if rows is None else sorted(set(rows))
        want = []
        for s in range(g.shape[1]):
            vals = [int(g[r, s]) for r in sel if g[r, s] >= 0]
            alleles = [a for v in vals for a in ([1] * v + [0] * (2 - v))]
            if len(vals) >= 2 and 0 in alleles and 1 in alleles:
                want.append(s)
        np.testing.assert_array_equal(keep, want)


def test_heterozygosity_geno_by_allele_enumeration():
    """H_i_i: probability that two distinct allele copies of population i
    differ; H_i_j: that one copy from i and one from j differ -- enumerated
    over explicit allele copies at every site."""
    rng = np.random.default_rng(3)
    g = rng.integers(0, 3, (9, 25)).astype(np.int8)
    g[rng.random(g.shape) < 0.2] = -1
    g[0:3, 0] = [1, -1, -1]                          # pop 0 has one valid individual at site 0
    pops = [[0, 1, 2], [3, 4, 5, 6], [7, 8]]
    got = o.heterozygosity_geno(g, pops)
    copies = [[[a for r in p if g[r, s] >= 0 for a in [1] * g[r, s] + [0] * (2 - g[r, s])]
               for s in range(g.shape[1])] for p in pops]
    for i, j in itertools.combinations_with_replacement(range(3), 2):
        tot = 0.0
        for s in range(g.shape[1]):
            ci, cj = copies[i][s], copies[j][s]
            if i == j:
                pairs = [(x, y) for a, x in enumerate(ci) for b, y in enumerate(ci) if a != b]
            else:
                pairs = [(x, y) for x in ci for y in cj]
            tot += (sum(x != y for x, y in pairs) / len(pairs)) if pairs else np.nan
        assert got[f"H_{i}_{j}"] == pytest.approx(tot, rel=1e-13, nan_ok=True)


def test_heterozygosity_geno_nan_for_empty_population_site():
    g = np.array([[0, 1], [-1, 1], [-1, 2]], np.int8)
    h = o.heterozygosity_geno(g, [[1, 2], [0]])
    assert np.isnan(h["H_0_0"]) and np.isnan(h["H_0_1"]) and np.isfinite(h["H_1_1"])


def test_rec_map_interpolation_closed_forms():
    path = GOLD / "rec_map_im.txt"
    mp, mv = o.read_rec_map(path)          # header line skipped
    np.testing.assert_array_equal(mp, [0, 250_000, 500_000, 1_000_000])
    np.testing.assert_allclose(mv, [0, 0.0025, 0.0075, 0.0125])
    pos = np.array([0, 125_000, 250_000, 375_000, 750_000, 1_000_000, 1_200_000])
    got = o.interpolate_genetic_distances(pos, path)
    # rate 1e-8 M/bp, then 2e-8, then 1e-8; clamped past the last knot
    np.testing.assert_allclose(got, [0, 0.00125, 0.0025, 0.005, 0.01, 0.0125, 0.0125],
                               rtol=1e-15, atol=1e-18)


def test_max_bp_for_r_dist_hand_cases():
    pos = np.array([0, 100, 200, 300])
    gen = np.array([0, 1e-6, 3e-6, 4e-6])        # local rates 1e-8, 2e-8, 1e-8
    assert o.max_bp_for_r_dist(pos, gen, 1e-6) == pytest.approx(110.0, rel=1e-12)
    assert o.max_bp_for_r_dist(pos, gen, 1e-5) == 300.0          # capped at the span
    assert o.max_bp_for_r_dist(pos, np.zeros(4), 1e-6) == 300.0  # no positive rate
    assert o.max_bp_for_r_dist(np.array([5, 5, 5]), np.zeros(3), 1e-6) == 0.0


def test_r2_dosage_matrix_is_corrcoef_of_mean_imputed_columns():
    rng = np.random.default_rng(4)
    g = rng.integers(0, 3, (40, 12)).astype(np.int8)
    g[rng.random(g.shape) < 0.15] = -1
    g[:, 5] = 1                                     # no variance: r^2 = 0
    got = o.r2_dosage_matrix(g)
    imp = g.astype(np.float64)
    for s in range(g.shape[1]):
        v = g[:, s] >= 0
        imp[~v, s] = g[v, s].mean()
    for i in range(12):
        for j in range(12):
            if i == j or 5 in (i, j):
                assert got[i, j] == 0.0
            else:
                assert got[i, j] == pytest.approx(np.corrcoef(imp[:, i], imp[:, j])[0, 1] ** 2,
                                                  rel=1e-12)


def test_tally_upper_brute_force():
    rng = np.random.default_rng(5)
    h = rng.integers(0, 2, (15, 9)).astype(np.int8)
    c, nv = o.tally_upper(h)
    assert nv is None and c.shape == (36, 4)
    h[rng.random(h.shape) < 0.2] = -1
    c, nv = o.tally_upper(h, rows=range(2, 15))
    for k, (i, j) in enumerate(zip(*np.triu_indices(9, 1))):
        a, b = h[2:, i], h[2:, j]
        ok = (a >= 0) & (b >= 0)
        want = [np.sum(ok & (a == x) & (b == y)) for x, y in ((1, 1), (1, 0), (0, 1), (0, 0))]
        np.testing.assert_array_equal(c[k], want)
        assert nv[k] == ok.sum()


def test_moments_driver_restatement_matches_golden_fixture():
    """moments_ld_stats (bp bins, haplotypes) reproduces moments' own output."""
    from tests.test_oracle_golden import load_im

    fx = json.loads((GOLD / "moments_im_2pop.json").read_text())
    hap, pos, pops = load_im()
    res = o.moments_ld_stats(hap, pos, pops, "hap", bp_bins=fx["bp_bins"])
    ref = np.array(fx["sums"])
    got = np.array(res["sums"][:-1])
    assert np.max(np.abs(got - ref) / np.abs(ref)) < 1e-12
    np.testing.assert_allclose(res["sums"][-1], fx["het"], rtol=1e-14)
    assert res["stats"] == (fx["ld_names"], fx["het_names"])


@pytest.mark.parametrize("missing", [False, True])
def test_band_count_checksum_brute_force(missing):
    """The closed form of the counting-stage checksum against a pair-by-pair
    loop over whole tiles (padded sites past the end included)."""
    rng = np.random.default_rng(6 + missing)
    K, V = 11, 70
    h = rng.integers(0, 2, (K + 5, V)).astype(np.int8)
    if missing:
        h[rng.random(h.shape) < 0.2] = -1
    rows = list(range(3, 3 + K))
    tiles = np.array([[0, 0], [0, 1], [1, 2], [2, 2], [0, 2]])
    x = h[rows]
    miss = bool(np.any(x < 0))
    tot = np.zeros((2, 4), np.int64)
    for rb, cb in tiles:
        for i in range(32 * rb, 32 * rb + 32):
            for j in range(32 * cb, 32 * cb + 32):
                a = x[:, i] if i < V else None
                b = x[:, j] if j < V else None
                if miss:
                    if a is None or b is None:
                        f = (0, 0, 0, 0)
                    else:
                        ok = (a >= 0) & (b >= 0)
                        f = (np.sum(ok & (a == 1) & (b == 1)), np.sum(ok & (a == 1)),
                             np.sum(ok & (b == 1)), np.sum(ok))
                else:
                    ki = int(np.sum(a == 1)) if a is not None else 0
                    kj = int(np.sum(b == 1)) if b is not None else 0
                    n11 = int(np.sum((a == 1) & (b == 1))) if a is not None and b is not None else 0
                    f = (n11, ki, kj, K)
                tot[0] += f
                tot[1] += (f[0], f[1] - f[0], f[2] - f[0], f[3] - f[1] - f[2] + f[0])
    np.testing.assert_array_equal(o.band_count_checksum(h, [rows], tiles, True)[0], tot[0])
    np.testing.assert_array_equal(o.band_count_checksum(h, [rows], tiles, False)[0], tot[1])
