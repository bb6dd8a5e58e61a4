"""Pins the CPU oracle's haplotype path (runs without a GPU).

Each check is chosen so that a plausible transcription mistake (dropped term,
wrong sign, swapped population index) fails it:
  * pair enumeration against brute force (reference tests/test_ld_pair_iteration.py);
  * counts against the reference's hand-VCF answers and brute force;
  * estimators: two independent transcriptions of the reference agree, and
    each raw call is an unbiased estimator of its population target by exact
    multinomial enumeration;
  * r / r^2 / D' closed forms (reference tests/test_ld_r_squared.py,
    tests/test_haplotype_kernels_rsq.py);
  * names / spec counts (reference tests/test_moments_ld.py:245-283).
"""
import itertools
from fractions import Fraction as F
from math import factorial

import numpy as np
import pytest

from oracle import ld_oracle as o


# ---------------------------------------------------------------- enumeration

def naive_pairs(pos, max_dist):
    return {(i, j) for i in range(len(pos)) for j in range(i + 1, len(pos))
            if pos[j] - pos[i] <= max_dist}


@pytest.mark.parametrize("n", [2, 10, 50])
@pytest.mark.parametrize("factor", [0.1, 0.5, 1.0, 10.0])
def test_pairs_match_brute_force(n, factor):
    rng = np.random.default_rng(n * 997 + int(factor * 1000))
    pos = np.unique(np.sort(rng.integers(0, 10 * n, size=n)))
    if len(pos) < 2:
        pytest.skip("duplicates collapsed input")
    md = max(1, int((pos[-1] - pos[0]) * factor))
    ii, jj = o.pairs_within_distance(pos, md)
    assert set(zip(ii.tolist(), jj.tolist())) == naive_pairs(pos, md)
    chunks = list(o.iter_pair_chunks(pos, md, 7))
    if chunks:
        ci = np.concatenate([c[0] for c in chunks])
        cj = np.concatenate([c[1] for c in chunks])
        np.testing.assert_array_equal(ci, ii)
        np.testing.assert_array_equal(cj, jj)


def test_pairs_edge_cases():
    assert len(o.pairs_within_distance(np.array([], np.int64), 10)[0]) == 0
    assert len(o.pairs_within_distance(np.array([42]), 10)[0]) == 0
    pos = np.array([0, 10, 20, 30])
    assert len(o.pairs_within_distance(pos, 0)[0]) == 0
    assert len(o.pairs_within_distance(pos, 10**9)[0]) == 6
    # duplicate positions are distance 0 and enumerated
    ii, jj = o.pairs_within_distance(np.array([5, 5, 5]), 0)
    assert list(zip(ii, jj)) == [(0, 1), (0, 2), (1, 2)]


# ---------------------------------------------------------------- counting

def hand_hap():
    # reference tests/test_haplotype_counting.py:59-100 (ind0..ind3, 2 sites),
    # rows are allele-0 haplotypes then allele-1 haplotypes as from_vcf builds them.
    gt = np.array([[[1, 1], [0, 1], [1, 0], [0, 0]],
                   [[1, 0], [1, 0], [1, 1], [0, 0]]], np.int8)  # [site, sample, ploidy]
    return np.concatenate([gt[:, :, 0].T, gt[:, :, 1].T], axis=0)


def test_counts_hand_vcf():
    hap = hand_hap()
    c, nv = o.hap_counts(hap, [0], [1], rows=[0, 1, 4, 5])
    np.testing.assert_array_equal(c[0], [1, 2, 1, 0])
    assert nv is None
    c, _ = o.hap_counts(hap, [0], [1], rows=[2, 3, 6, 7])
    np.testing.assert_array_equal(c[0], [1, 0, 1, 2])


def test_counts_brute_force_with_missing():
    rng = np.random.default_rng(7)
    hap = rng.integers(0, 2, (30, 12)).astype(np.int8)
    hap[rng.random(hap.shape) < 0.1] = -1
    ii, jj = o.pairs_within_distance(np.arange(12), 12)
    c, nv = o.hap_counts(hap, ii, jj)
    for t, (i, j) in enumerate(zip(ii, jj)):
        a, b = hap[:, i], hap[:, j]
        ok = (a >= 0) & (b >= 0)
        exp = [np.sum(ok & (a == x) & (b == y)) for x, y in ((1, 1), (1, 0), (0, 1), (0, 0))]
        np.testing.assert_array_equal(c[t], exp)
        assert nv[t] == ok.sum() == c[t].sum()


def test_geno_counts_brute_force():
    rng = np.random.default_rng(42)
    g = rng.integers(0, 3, (25, 8)).astype(np.int8)
    g[rng.random(g.shape) < 0.05] = -1
    ii, jj = o.pairs_within_distance(np.arange(8), 8)
    c, nv = o.geno_counts(g, ii, jj)
    for t, (i, j) in enumerate(zip(ii, jj)):
        a, b = g[:, i], g[:, j]
        ok = (a >= 0) & (b >= 0)
        exp = [np.sum(ok & (a == x) & (b == y)) for x in range(3) for y in range(3)]
        np.testing.assert_array_equal(c[t], exp)
        assert nv[t] == ok.sum()
    g2 = np.abs(g)
    assert o.geno_counts(g2, ii, jj)[1] is None


# ---------------------------------------------------------------- estimators

def rand_counts(rng, N, lo=0, hi=40):
    return rng.integers(lo, hi, (N, 4)).astype(np.int32)


def test_two_transcriptions_agree_two_pops():
    """ld_statistics.py formulas (2-pop fast path) == haplotype_kernels.py
    kernel formulas through the moments spec averaging."""
    rng = np.random.default_rng(0)
    a, b = rand_counts(rng, 5000), rand_counts(rng, 5000)
    a[:50] = rng.integers(0, 2, (50, 4))  # tiny n: exercises the validity guards
    fast = o.lds_two_pop_15(a, b)
    spec = o.multi_pop_stats_hap([a, b], [None, None])
    scale = np.maximum(np.abs(fast), 1e-300)
    assert np.max(np.abs(fast - spec) / scale) < 1e-12


def test_single_pop_transcriptions_agree():
    rng = np.random.default_rng(1)
    c = rand_counts(rng, 4000, 0, 30)
    P = [o.HapPop(c)]
    np.testing.assert_allclose(o.lds_dd_single(c), o.hap_dd(P, 0, 0), rtol=1e-13, atol=0)
    np.testing.assert_allclose(o.lds_dz_single(c), o.hap_dz(P, 0, 0, 0), rtol=1e-13, atol=0)
    np.testing.assert_allclose(o.lds_pi2_single(c), o.hap_pi2(P, 0, 0, 0, 0), rtol=1e-13, atol=0)


def _comps(n, k=4):
    if k == 1:
        yield (n,)
        return
    for a in range(n + 1):
        for rest in _comps(n - a, k - 1):
            yield (a,) + rest


def _mult_p(c, f):
    p = F(factorial(sum(c)))
    for ci, fi in zip(c, f):
        p = p / factorial(ci) * fi ** ci
    return p


def _expect(call, pops, fs, ns):
    tot = F(0)
    for combo in itertools.product(*[list(_comps(n)) for n in ns]):
        pr = F(1)
        for c, f in zip(combo, fs):
            pr *= _mult_p(c, f)
        if pr:
            tot += pr * o.exact_hap_call(list(combo), call, pops)
    return tot


def _target(call, pops, fs):
    """Population value each raw call estimates.  f = (f11, f10, f01, f00);
    D = f11 f00 - f10 f01; p = f11 + f10 (allele 1 at A), q = f11 + f01."""
    D = lambda f: f[0] * f[3] - f[1] * f[2]
    p = lambda f: f[0] + f[1]
    q = lambda f: f[0] + f[2]
    if call == "dd":
        i, j = pops
        return D(fs[i]) * D(fs[j])
    if call == "dz":
        i, j, k = pops
        return D(fs[i]) * (1 - 2 * p(fs[j])) * (1 - 2 * q(fs[k]))
    if o.pi2_case(*pops) == "triple":
        # every "three of one pop" call uses the (t, t, t, s) estimator
        cnt = {x: pops.count(x) for x in pops}
        t = [x for x, c in cnt.items() if c == 3][0]
        s = [x for x, c in cnt.items() if c == 1][0]
        pops = (t, t, t, s)
    i, j, k, l = pops
    H = lambda u, v, g: g(u) * (1 - g(v)) + g(v) * (1 - g(u))
    return F(1, 4) * H(fs[i], fs[j], p) * H(fs[k], fs[l], q)


def _freqs(seed):
    rng = np.random.default_rng(seed)
    w = [F(int(x)) for x in rng.integers(1, 10, 4)]
    s = sum(w)
    return [x / s for x in w]


UNBIASED_CASES = [
    ("dd", (0, 0), [5]), ("dz", (0, 0, 0), [5]), ("pi2", (0, 0, 0, 0), [5]),
    ("dd", (0, 1), [3, 3]), ("dz", (0, 0, 1), [4, 2]), ("dz", (0, 1, 0), [4, 2]),
    ("dz", (0, 1, 1), [3, 3]), ("pi2", (0, 0, 0, 1), [4, 2]), ("pi2", (0, 1, 1, 1), [2, 4]),
    ("pi2", (1, 0, 0, 0), [4, 2]), ("pi2", (0, 0, 1, 1), [3, 3]), ("pi2", (0, 1, 0, 1), [3, 3]),
    ("pi2", (0, 1, 1, 0), [3, 3]),
    ("dz", (0, 1, 2), [3, 2, 2]), ("pi2", (0, 0, 1, 2), [3, 2, 2]), ("pi2", (0, 1, 2, 2), [2, 2, 3]),
    ("pi2", (0, 1, 0, 2), [3, 2, 2]), ("pi2", (1, 0, 2, 0), [3, 2, 2]),
    ("pi2", (0, 1, 2, 3), [2, 2, 2, 2]),
]


@pytest.mark.parametrize("call,pops,ns", UNBIASED_CASES)
def test_estimators_unbiased(call, pops, ns):
    fs = [_freqs(s) for s in range(len(ns))]
    assert _expect(call, pops, fs, ns) == _target(call, pops, fs)


def test_unbiasedness_check_catches_a_planted_mistake():
    """The enumeration check is sharp: flipping one coefficient breaks it."""
    fs = [_freqs(0)]
    orig = o.hap_dd
    try:
        def bad(P, i, j):
            a = P[i]
            num = (a.c1 * (a.c1 - 1) * a.c4 * (a.c4 - 1) + a.c2 * (a.c2 - 1) * a.c3 * (a.c3 - 1)
                   - 1 * a.c1 * a.c2 * a.c3 * a.c4)
            return o._ratio(num, o._ff(a.n, 4))
        o.hap_dd = bad
        assert _expect("dd", (0, 0), fs, [5]) != _target("dd", (0, 0), fs)
    finally:
        o.hap_dd = orig


def test_validity_guards():
    # n < 4 -> single-pop statistics are 0 (ld_statistics.py:1124-1126)
    c = np.array([[1, 1, 1, 0], [0, 0, 0, 0]])
    for f in (o.lds_dd_single, o.lds_dz_single, o.lds_pi2_single):
        assert np.all(f(c) == 0)


# ---------------------------------------------------------------- r family

def test_r_closed_forms():
    r, r2, dp = o.r_family(np.array([[5, 0, 0, 5], [25, 25, 25, 25], [0, 5, 5, 0], [5, 5, 0, 0],
                                     [50, 0, 0, 50], [0, 50, 50, 0]]))
    np.testing.assert_allclose(r2[[0, 2, 4, 5]], 1.0, rtol=1e-12)
    assert abs(r2[1]) < 1e-12
    np.testing.assert_allclose(r[[2, 5]], -1.0, rtol=1e-12)
    assert np.isnan(r2[3]) and np.isnan(r[3]) and np.isnan(dp[3])
    np.testing.assert_allclose(dp[[4, 5]], [1.0, -1.0], rtol=1e-12)


# ---------------------------------------------------------------- names / specs

def test_names_and_spec_counts():
    assert [len(o.ld_names(p)) for p in range(1, 5)] == [3, 15, 45, 105]
    assert [len(o.het_names(p)) for p in range(1, 5)] == [1, 3, 6, 10]
    uniq = [len({(t, pp) for _, c in o.stat_specs(p) for _, t, pp in c}) for p in range(1, 5)]
    assert uniq == [3, 27, 114, 330]
    assert o.ld_names(2) == [
        "DD_0_0", "DD_0_1", "DD_1_1", "Dz_0_0_0", "Dz_0_0_1", "Dz_0_1_1", "Dz_1_0_0", "Dz_1_0_1",
        "Dz_1_1_1", "pi2_0_0_0_0", "pi2_0_0_0_1", "pi2_0_0_1_1", "pi2_0_1_0_1", "pi2_0_1_1_1",
        "pi2_1_1_1_1"]
    for p in range(1, 5):
        for _, calls in o.stat_specs(p):
            assert abs(sum(w for w, _, _ in calls) - 1.0) < 1e-15


# ---------------------------------------------------------------- tile consumers

def test_sigma_d2_matrix_is_ratio_of_unbiased_estimators():
    """sigma_D^2 of the tile consumers is DD / pi2 of the pinned single-pop
    estimators (their common 1/(n)_4 cancels)."""
    rng = np.random.default_rng(70)
    h = (rng.random((30, 12)) < 0.45).astype(np.int8)
    h[rng.random(h.shape) < 0.05] = -1
    val, flag = o.sigma_d2_matrix(h)
    ii, jj = np.triu_indices(12, 1)
    c, nv = o.hap_counts(h, ii, jj)
    P = [o.HapPop(c, nv)]
    dd, pi2 = o.hap_dd(P, 0, 0), o.hap_pi2(P, 0, 0, 0, 0)
    ok = (P[0].n >= 4) & (pi2 != 0)
    np.testing.assert_allclose(val[ii, jj][ok], (dd / np.where(ok, pi2, 1))[ok], rtol=1e-12)
    np.testing.assert_array_equal(flag[ii, jj], (P[0].n >= 4).astype(float))


def test_r2_naive_matrix_brute_force():
    rng = np.random.default_rng(71)
    h = (rng.random((25, 9)) < 0.5).astype(np.int8)
    h[rng.random(h.shape) < 0.1] = -1
    got = o.r2_naive_matrix(h)
    for i in range(9):
        for j in range(9):
            if i == j:
                assert got[i, j] == 0
                continue
            vi, vj = h[:, i] >= 0, h[:, j] >= 0
            pi, pj = h[vi, i].mean(), h[vj, j].mean()
            both = vi & vj
            pab = np.mean((h[both, i] == 1) & (h[both, j] == 1)) if both.any() else 0.0
            den = pi * (1 - pi) * pj * (1 - pj)
            ref = (pab - pi * pj) ** 2 / den if den > 0 else 0.0
            assert abs(got[i, j] - ref) < 1e-12


def test_omega_prefix_sums_match_direct_sums():
    """The prefix-sum omega of the build (ld_tiles.omega) restates the same
    formula as the oracle's direct per-split sums; check the oracle's
    split-point bookkeeping on a matrix with a known block structure."""
    m = 8
    r2 = np.full((m, m), 0.1)
    r2[:4, :4] = 0.9
    r2[4:, 4:] = 0.9
    np.fill_diagonal(r2, 0)
    # at l = 4 within-mean 0.9, cross-mean 0.1 -> 9; other splits mix blocks
    assert abs(o.omega_value(r2) - 9.0) < 1e-12
    assert o.zns_value(r2) == pytest.approx((0.9 * 24 + 0.1 * 32) / 56)


def test_windowed_r_squared_matches_pearson_brute_force():
    """Count r^2 over the samples valid at both sites is the squared Pearson
    correlation of the two 0/1 columns there (np.corrcoef, pair by pair)."""
    rng = np.random.default_rng(11)
    h = (rng.random((16, 14)) < 0.35).astype(np.int8)
    h[rng.random(h.shape) < 0.1] = -1
    h[:, 2] = 0
    pos = np.sort(rng.choice(np.arange(1, 500), 14, replace=False))
    bins = [0, 60, 200, 600]
    vals = {k: [] for k in range(3)}
    for i in range(14):
        for j in range(i + 1, 14):
            ok = (h[:, i] >= 0) & (h[:, j] >= 0)
            x, y = h[ok, i].astype(float), h[ok, j].astype(float)
            if x.std() == 0 or y.std() == 0:
                continue
            k = np.digitize(pos[j] - pos[i], bins) - 1
            if 0 <= k < 3:
                vals[k].append(np.corrcoef(x, y)[0, 1] ** 2)
    got, cnt = o.windowed_r_squared(h, pos, bins, [20, 50, 100])
    for k in range(3):
        assert cnt[k] == len(vals[k])
        if vals[k]:
            np.testing.assert_allclose(got[k], np.percentile(vals[k], [20, 50, 100]), rtol=1e-12)
    h2 = np.where(h < 0, 0, h)
    got, cnt = o.windowed_r_squared(h2, pos, bins, 50, estimator="rogers_huff")
    g = h2[0::2].astype(float) + h2[1::2]
    for k in range(3):
        v = [np.corrcoef(g[:, i], g[:, j])[0, 1] ** 2 for i in range(14) for j in range(i + 1, 14)
             if g[:, i].std() > 0 and g[:, j].std() > 0
             and np.digitize(pos[j] - pos[i], bins) - 1 == k]
        assert cnt[k] == len(v)
        if v:
            np.testing.assert_allclose(got[k], np.percentile(v, 50), rtol=1e-12)
