"""Pins the genotype half of the oracle (CPU only).

* the single-population polynomials read from the two reference files
  (readable module and production kernels) are identical coefficient tables;
* every readable-form estimator is exactly unbiased for its haplotype-level
  target under random mating (factorial-moment expectation, exact rationals);
* the production (kernel) form -- the parity target of the GPU code -- agrees
  with the readable form on random counts for every raw call at P = 1..4;
* a planted coefficient error is caught by the unbiasedness check.
"""
import copy
from fractions import Fraction as F

import numpy as np
import pytest
import sympy as sp

from oracle import geno_oracle as go
from oracle import ld_oracle as o


def test_single_pop_tables_agree_between_reference_files():
    T = go.tables()
    for name in ("dd_geno_single", "dz_geno_single", "pi2_geno_single"):
        a = {tuple(m): F(c) for m, c in T["kernel_single"][name]}
        b = {tuple(m): F(c) for m, c in T["readable"][name]["numer"]}
        assert a == b, name


def _freqs(seed):
    rng = np.random.default_rng(seed + 100)
    w = [F(int(x)) for x in rng.integers(1, 10, 4)]
    s = sum(w)
    return [x / s for x in w]


def _target(kind, pops, fs):
    D = lambda f: f[0] * f[3] - f[1] * f[2]
    p = lambda f: f[0] + f[1]
    q = lambda f: f[0] + f[2]
    if kind == "dd":
        return D(fs[pops[0]]) * D(fs[pops[1]])
    if kind == "dz":
        return D(fs[pops[0]]) * (1 - 2 * p(fs[pops[1]])) * (1 - 2 * q(fs[pops[2]]))
    i, j, k, l = pops
    H = lambda u, v, g: g(u) * (1 - g(v)) + g(v) * (1 - g(u))
    return F(1, 4) * H(fs[i], fs[j], p) * H(fs[k], fs[l], q)


# readable function -> (statistic, target populations in terms of its argument slots)
CASES = {
    "dd_geno_single": ("dd", (0, 0)), "dd_geno_between": ("dd", (0, 1)),
    "dz_geno_single": ("dz", (0, 0, 0)), "dz_geno_p1p2": ("dz", (0, 0, 1)),
    "dz_geno_p1p3": ("dz", (0, 1, 0)), "dz_geno_p2p3": ("dz", (0, 1, 1)),
    "dz_geno_alldiff": ("dz", (0, 1, 2)), "pi2_geno_single": ("pi2", (0, 0, 0, 0)),
    "pi2_geno_triple_123": ("pi2", (0, 0, 0, 1)), "pi2_geno_triple_134": ("pi2", (0, 1, 0, 0)),
    "pi2_geno_iikk": ("pi2", (0, 0, 1, 1)), "pi2_geno_ijij": ("pi2", (0, 1, 0, 1)),
    "pi2_geno_iikl": ("pi2", (0, 0, 1, 2)), "pi2_geno_ijkk": ("pi2", (1, 2, 0, 0)),
    "pi2_geno_shared": ("pi2", (0, 1, 0, 2)), "pi2_geno_alldiff": ("pi2", (0, 1, 2, 3)),
}


def _estimate_expectation(t, ns, Qs):
    e = go.expected_table(t["numer"], Qs, ns)
    den = sp.sympify(t["denom"]).subs({sp.Symbol(f"n{k}"): ns[k] for k in range(len(ns))})
    return e * F(t["scale"]) / F(int(den))


@pytest.mark.parametrize("name", sorted(CASES))
def test_genotype_estimators_unbiased(name):
    T = go.tables()["readable"]
    t = T[name]
    kind, pops = CASES[name]
    P = t["n_pops"]
    fs = [_freqs(s) for s in range(P)]
    Qs = [go.genotype_class_probs(f) for f in fs]
    ns = [6] * P
    assert _estimate_expectation(t, ns, Qs) == _target(kind, pops, fs)


def test_unbiasedness_catches_planted_coefficient_error():
    t = copy.deepcopy(go.tables()["readable"]["dd_geno_single"])
    t["numer"][5][1] = str(F(t["numer"][5][1]) + F(1, 16))
    fs = [_freqs(0)]
    Qs = [go.genotype_class_probs(f) for f in fs]
    assert _estimate_expectation(t, [6], Qs) != _target("dd", (0, 0), fs)


@pytest.mark.parametrize("P", [1, 2, 3, 4])
def test_kernel_form_matches_readable_form(P):
    rng = np.random.default_rng(P)
    N = 200
    cs = [rng.integers(0, 12, (N, 9)) for _ in range(P)]
    cs[0][:5] = rng.integers(0, 2, (5, 9))          # small samples: guard paths
    Pk = [go.GenoPop(c) for c in cs]
    G = [go.to_moments(c) for c in cs]
    ns = [g.sum(1) for g in G]
    calls = sorted({(t, pp) for _, c in o.stat_specs(P) for _, t, pp in c})
    for kind, pp in calls:
        a = go.readable_call(G, ns, kind, pp)
        b = go.kernel_call(Pk, kind, pp)
        ok = np.isfinite(a) & np.isfinite(b) & (np.minimum.reduce([n for n in ns]) >= 4)
        np.testing.assert_allclose(b[ok], a[ok], rtol=1e-12, atol=1e-300, err_msg=str((kind, pp)))


def test_moments_order_mapping():
    c = np.zeros((1, 9))
    c[0, 8] = 5      # n22 -> g1
    c[0, 0] = 2      # n00 -> g9
    g = go.to_moments(c)
    assert g[0, 0] == 5 and g[0, 8] == 2


def test_dosage_r2_matches_brute_force_pearson():
    """r2_dosage from the nine counts equals numpy's Pearson r^2 over the
    individuals valid at both loci (and NaN exactly where a locus is constant)."""
    from oracle import ld_oracle as o

    rng = np.random.default_rng(7)
    g = rng.integers(0, 3, (40, 30)).astype(np.int8)
    g[rng.random(g.shape) < 0.1] = -1
    g[:, 5] = np.where(g[:, 5] >= 0, 1, -1)           # monomorphic among the valid
    ii, jj = np.triu_indices(30, 1)
    c, _ = o.geno_counts(g, ii, jj)
    got = go.r2_dosage(c)
    for k, (i, j) in enumerate(zip(ii, jj)):
        ok = (g[:, i] >= 0) & (g[:, j] >= 0)
        x, y = g[ok, i].astype(float), g[ok, j].astype(float)
        if x.std() == 0 or y.std() == 0:
            assert np.isnan(got[k])
        else:
            assert abs(got[k] - np.corrcoef(x, y)[0, 1] ** 2) < 1e-12


def test_haplotype_r2_matches_brute_force_pearson():
    """ld_statistics.r_squared's count formula is the Pearson r^2 of the 0/1
    alleles over samples valid at both loci."""
    from oracle import ld_oracle as o

    rng = np.random.default_rng(8)
    h = (rng.random((50, 25)) < 0.4).astype(np.int8)
    h[rng.random(h.shape) < 0.08] = -1
    ii, jj = np.triu_indices(25, 1)
    c, nv = o.hap_counts(h, ii, jj)
    got = o.r_family(c, nv)[1]
    for k, (i, j) in enumerate(zip(ii, jj)):
        ok = (h[:, i] >= 0) & (h[:, j] >= 0)
        x, y = h[ok, i].astype(float), h[ok, j].astype(float)
        if x.std() == 0 or y.std() == 0:
            assert np.isnan(got[k])
        else:
            assert abs(got[k] - np.corrcoef(x, y)[0, 1] ** 2) < 1e-12


def test_centered_moment_forms_equal_reference_polynomials():
    """The device code evaluates the reference's three single-population
    sub-polynomials (genotype_kernels.py:27-71, 235-272, 364-405; the
    coefficient tables in oracle/geno_kernel_polys.json) in centered
    dosage moments.  Substituting the moments' definitions as linear forms of
    the nine class counts, each form expands to the reference polynomial
    exactly; so do the features D, pA, qA, pB, qB, A, B the other estimators
    use."""
    import json
    from pathlib import Path

    path = Path(__file__).resolve().parents[1] / "oracle" / "geno_kernel_polys.json"
    data = json.loads(path.read_text())
    g = sp.symbols("g0:9")
    X = [2, 2, 2, 1, 1, 1, 0, 0, 0]   # left dosage of class m (moments order)
    Y = [2, 1, 0, 2, 1, 0, 2, 1, 0]   # right dosage

    def mom(k, l):
        return sum(sp.Integer((X[m] - 1) ** k * (Y[m] - 1) ** l) * g[m] for m in range(9))

    n, a1, a2, b1, b2 = mom(0, 0), mom(1, 0), mom(2, 0), mom(0, 1), mom(0, 2)
    ab11, ab12, ab21, ab22 = mom(1, 1), mom(1, 2), mom(2, 1), mom(2, 2)
    u = a2 * b2 - a1 ** 2 * b2 - a2 * b1 ** 2
    t4, t56, s11 = a1 * b1 * ab11, a1 * ab12 + ab21 * b1, ab11 ** 2
    forms = {
        "dd_poly": 2 * (t4 + (n - 1) * t56) + u - s11 * (3 * n - 2) - ab22 * n * (n - 1),
        "dz_quartic": t4 - u - (n + 1) * t56 - s11 * (n - 1) + 2 * ab22 * n,
        "pi2_corr": (u + 4 * (t56 - t4) + (a1 ** 2 + b1 ** 2) * (5 * n - 6)
                     + (a2 + b2) * (n - 2) * (n - 3) + 2 * s11 - 6 * ab22
                     - n * (6 * n ** 2 - 11 * n + 6)),
    }
    for name, form in forms.items():
        ref = sum(sp.Integer(int(c)) * sp.Mul(*[g[m] ** e for m, e in enumerate(exps)])
                  for exps, c in data[name])
        assert sp.expand(form - ref) == 0, name
    half, quarter = sp.Rational(1, 2), sp.Rational(1, 4)
    gxAB = g[0] + half * g[1] + half * g[3] + quarter * g[4]
    gxAb = half * g[1] + g[2] + quarter * g[4] + half * g[5]
    gxaB = half * g[3] + quarter * g[4] + g[6] + half * g[7]
    gxab = quarter * g[4] + half * g[5] + half * g[7] + g[8]
    feats = {  # reference feature (genotype_kernels.py:412-492) vs centered form
        "D": (gxAB * gxab - gxAb * gxaB, (n * ab11 - a1 * b1) / 4),
        "pA": (g[0] + g[1] + g[2] + half * (g[3] + g[4] + g[5]), (n + a1) / 2),
        "qA": (half * (g[3] + g[4] + g[5]) + g[6] + g[7] + g[8], (n - a1) / 2),
        "pB": (g[0] + half * g[1] + g[3] + half * g[4] + g[6] + half * g[7], (n + b1) / 2),
        "qB": (half * g[1] + g[2] + half * g[4] + g[5] + half * g[7] + g[8], (n - b1) / 2),
        "A": (-g[0] + g[2] - g[3] + g[5] - g[6] + g[8], -b1),
        "B": (-g[0] - g[1] - g[2] + g[6] + g[7] + g[8], -a1),
    }
    for name, (ref, form) in feats.items():
        assert sp.expand(ref - form) == 0, name
