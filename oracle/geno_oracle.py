"""CPU oracle for pg_gpu's genotype (unphased diploid) LD estimators
(TEST INFRASTRUCTURE ONLY -- imported by tests/, never by the product).

Two transcriptions of the reference are kept side by side:

* ``readable`` -- the polynomial numerators of
  pg_gpu/ld_statistics_genotype.py:37-1505, evaluated from the exact
  coefficient tables in geno_polys.json.gz (see extract_geno_polys.py), with
  that module's denominators, validity masks and scale factors;
* ``kernel``   -- the production formulas of pg_gpu/genotype_kernels.py:27-405
  on the per-population features of _PopDataGeno (:412-492), with the case
  dispatch of compute_all_{dd,dz,pi2}_geno (:561-878).  This is what the
  reference's binned path runs, so it is the parity target.

Genotype counts per pair are ordered (n00, n01, ..., n22) (n_ab: genotype a at
the left locus, b at the right).  The moments order used by the formulas is
g1 = n22, g2 = n21, g3 = n20, g4 = n12, g5 = n11, g6 = n10, g7 = n02,
g8 = n01, g9 = n00 (genotype_kernels.py:424-435).

Pins: the two transcriptions agree; every readable estimator is an exactly
unbiased estimator of its haplotype-frequency target under random mating
(tests/test_oracle_geno.py, by factorial-moment expectation); the single-pop
polynomials of both reference files agree symbolically.
"""
from __future__ import annotations

import itertools
from fractions import Fraction
from functools import lru_cache

import numpy as np

from .extract_geno_polys import load as _load_tables
from .ld_oracle import pi2_case

# ---------------------------------------------------------------------------
# moments order
# ---------------------------------------------------------------------------


def to_moments(counts):
    """(N, 9) counts (n00..n22) -> (N, 9) float64 in moments order g1..g9."""
    c = np.asarray(counts, np.float64)
    return c[:, ::-1].copy()


@lru_cache(maxsize=1)
def tables():
    return _load_tables()


# ---------------------------------------------------------------------------
# readable form (ld_statistics_genotype.py)
# ---------------------------------------------------------------------------


def _eval_table(table, gs):
    """Sum of coef * prod g^e over a coefficient table; gs = list of (N, 9)."""
    N = gs[0].shape[0]
    pw = [np.stack([g ** k for k in range(5)], axis=-1) for g in gs]  # (N, 9, 5)
    out = np.zeros(N)
    for mon, coef in table:
        term = np.full(N, float(Fraction(coef)))
        for p, g in enumerate(pw):
            e = mon[9 * p:9 * p + 9]
            for m in range(9):
                if e[m]:
                    term = term * g[:, m, e[m]]
        out += term
    return out


def _den(expr, ns):
    import sympy as sp

    f = sp.lambdify([sp.Symbol(f"n{k}") for k in range(len(ns))], sp.sympify(expr), "numpy")
    return np.asarray(f(*ns), np.float64)


def readable(name, gs, ns):
    """One readable-form estimator; gs moments-order arrays, ns sample sizes."""
    t = tables()["readable"][name]
    num = _eval_table(t["numer"], gs)
    den = _den(t["denom"], ns)
    valid = _VALID[name](ns)
    out = np.zeros_like(num)
    out[valid] = num[valid] / np.maximum(den[valid], 1)
    return t["scale"] * out


_VALID = {
    "dd_geno_single": lambda n: n[0] >= 4, "dz_geno_single": lambda n: n[0] >= 4,
    "pi2_geno_single": lambda n: n[0] >= 4,
    "dd_geno_between": lambda n: (n[0] >= 2) & (n[1] >= 2),
    "dz_geno_p1p2": lambda n: (n[0] >= 3) & (n[1] >= 1),
    "dz_geno_p1p3": lambda n: (n[0] >= 3) & (n[1] >= 1),
    "dz_geno_p2p3": lambda n: (n[0] >= 2) & (n[1] >= 2),
    "dz_geno_alldiff": lambda n: (n[0] >= 2) & (n[1] >= 1) & (n[2] >= 1),
    "pi2_geno_triple_123": lambda n: (n[0] >= 3) & (n[1] >= 1),
    "pi2_geno_triple_134": lambda n: (n[0] >= 3) & (n[1] >= 1),
    "pi2_geno_iikk": lambda n: (n[0] >= 2) & (n[1] >= 2),
    "pi2_geno_ijij": lambda n: (n[0] >= 2) & (n[1] >= 2),
    "pi2_geno_iikl": lambda n: (n[0] >= 2) & (n[1] >= 1) & (n[2] >= 1),
    "pi2_geno_ijkk": lambda n: (n[0] >= 2) & (n[1] >= 1) & (n[2] >= 1),
    "pi2_geno_shared": lambda n: (n[0] >= 2) & (n[1] >= 1) & (n[2] >= 1),
    "pi2_geno_alldiff": lambda n: (n[0] >= 1) & (n[1] >= 1) & (n[2] >= 1) & (n[3] >= 1),
}


def readable_call(G, ns, kind, pops):
    """Raw call on per-pop moments arrays G[p] and sizes ns[p], readable form,
    with the population slot conventions of ld_statistics_genotype.py."""
    if kind == "dd":
        i, j = pops
        if i == j:
            return readable("dd_geno_single", [G[i]], [ns[i]])
        return readable("dd_geno_between", [G[i], G[j]], [ns[i], ns[j]])
    if kind == "dz":
        i, j, k = pops
        if i == j == k:
            return readable("dz_geno_single", [G[i]], [ns[i]])
        if i == j:
            return readable("dz_geno_p1p2", [G[i], G[k]], [ns[i], ns[k]])
        if i == k:
            return readable("dz_geno_p1p3", [G[i], G[j]], [ns[i], ns[j]])
        if j == k:
            return readable("dz_geno_p2p3", [G[i], G[j]], [ns[i], ns[j]])
        return readable("dz_geno_alldiff", [G[i], G[j], G[k]], [ns[i], ns[j], ns[k]])
    i, j, k, l = pops
    case = pi2_case(*pops)
    if case == "same":
        return readable("pi2_geno_single", [G[i]], [ns[i]])
    if case == "triple":
        cnt = {x: pops.count(x) for x in pops}
        t = [x for x, c in cnt.items() if c == 3][0]
        s = [x for x, c in cnt.items() if c == 1][0]
        name = "pi2_geno_triple_123" if (i == t and j == t) else "pi2_geno_triple_134"
        return readable(name, [G[t], G[s]], [ns[t], ns[s]])
    if case == "iikk":
        return readable("pi2_geno_iikk", [G[i], G[k]], [ns[i], ns[k]])
    if case == "iikl":
        return readable("pi2_geno_iikl", [G[i], G[k], G[l]], [ns[i], ns[k], ns[l]])
    if case == "ijkk":
        return readable("pi2_geno_ijkk", [G[k], G[i], G[j]], [ns[k], ns[i], ns[j]])
    if case == "ijij":
        return readable("pi2_geno_ijij", [G[i], G[j]], [ns[i], ns[j]])
    if case == "shared":
        if i == k:
            s, a, b = i, j, l
        elif i == l:
            s, a, b = i, j, k
        elif j == k:
            s, a, b = j, i, l
        else:
            s, a, b = j, i, k
        return readable("pi2_geno_shared", [G[s], G[a], G[b]], [ns[s], ns[a], ns[b]])
    return readable("pi2_geno_alldiff", [G[i], G[j], G[k], G[l]], [ns[i], ns[j], ns[k], ns[l]])


# ---------------------------------------------------------------------------
# kernel form (genotype_kernels.py)
# ---------------------------------------------------------------------------


class GenoPop:
    """Per-pop features of one pair batch (genotype_kernels.py:412-492)."""

    def __init__(self, counts, n_valid=None):
        g = to_moments(counts)
        self.g = g
        g1, g2, g3, g4, g5, g6, g7, g8, g9 = (g[:, m] for m in range(9))
        self.n = (np.asarray(n_valid, np.float64) if n_valid is not None else g.sum(axis=1))
        n = self.n
        self.D = (-(g2 / 2 + g3 + g5 / 4 + g6 / 2) * (g4 / 2 + g5 / 4 + g7 + g8 / 2)
                  + (g1 + g2 / 2 + g4 / 2 + g5 / 4) * (g5 / 4 + g6 / 2 + g8 / 2 + g9))
        self.pA = g1 + g2 + g3 + g4 / 2 + g5 / 2 + g6 / 2
        self.qA = g4 / 2 + g5 / 2 + g6 / 2 + g7 + g8 + g9
        self.pB = g1 + g2 / 2 + g4 + g5 / 2 + g7 + g8 / 2
        self.qB = g2 / 2 + g3 + g5 / 2 + g6 + g8 / 2 + g9
        HL, HR = g4 + g5 + g6, g2 + g5 + g8
        nn1 = np.maximum(4 * n * (n - 1), 1)
        self.BL = (4 * self.pA * self.qA - HL) / nn1
        self.BR = (4 * self.pB * self.qB - HR) / nn1
        xAB = g1 + g2 / 2 + g4 / 2 + g5 / 4
        xAb = g2 / 2 + g3 + g5 / 4 + g6 / 2
        xaB = g4 / 2 + g5 / 4 + g7 + g8 / 2
        xab = g5 / 4 + g6 / 2 + g8 / 2 + g9
        self.C00 = (self.pA * self.pB - xAB) / nn1
        self.C01 = (self.pA * self.qB - xAb) / nn1
        self.C10 = (self.qA * self.pB - xaB) / nn1
        self.C11 = (self.qA * self.qB - xab) / nn1
        self.X01, self.X10, self.X11 = xAb, xaB, xab
        self.Bsgn = xAB - xAb - xaB + xab
        P, R, Q, S = self.pA, self.qA, self.pB, self.qB
        L0 = g5 + 2 * g6 + 2 * g8 + 4 * g9
        L1 = g5 + 2 * g6
        M1 = g5 + 2 * g8
        negR3 = 3 * g4 + 3 * g5 + 3 * g6 + 4 * g7 + 4 * g8 + 4 * g9
        negR5 = 5 * g4 + 5 * g5 + 5 * g6 + 8 * g7 + 8 * g8 + 8 * g9
        b5A = -2 * g2 - 4 * g3 + 4 * g4 + g5 - 2 * g6 + 8 * g7 + 4 * g8 + 4
        b6A = -g5 - 2 * g6 - 4 * g7 - 4 * g8 - 4 * g9
        self.AQ3 = -(S * negR3) / 8 + P * S * R / 2 + (R * L0) / 8 + L1 / 8 - (P * L0) / 8
        self.AS3 = -(Q * negR5) / 8 + P * Q * R / 2 + (R * b5A) / 8 + b6A / 8 + (P * L0) / 8
        negL3 = 3 * g2 + 4 * g3 + 3 * g5 + 4 * g6 + 3 * g8 + 4 * g9
        negL5 = 5 * g2 + 8 * g3 + 5 * g5 + 8 * g6 + 5 * g8 + 8 * g9
        b5B = 4 * g2 + 8 * g3 - 2 * g4 + g5 + 4 * g6 - 4 * g7 - 2 * g8 + 4
        b6B = -4 * g3 - g5 - 4 * g6 - 2 * g8 - 4 * g9
        self.BP3 = -(R * negL3) / 8 + Q * S * R / 2 + (S * L0) / 8 + M1 / 8 - (Q * L0) / 8
        self.BR3 = -(P * negL5) / 8 + P * Q * S / 2 + (S * b5B) / 8 + b6B / 8 + (Q * L0) / 8
        # Dz precompute (genotype_kernels.py:176-197); its n is the count sum
        ns = g.sum(axis=1)
        self.ns = ns
        self.A = -g1 + g3 - g4 + g6 - g7 + g9
        self.B = -g1 - g2 - g3 + g7 + g8 + g9
        c = -g1 + g3 + g7 - g9
        MA = (2 * g2 * g7 + 4 * g3 * g7 + 2 * g5 * g7 + 4 * g6 * g7 - 2 * g1 * g8 + 2 * g3 * g8
              - 2 * g4 * g8 + 2 * g6 * g8 - 4 * g1 * g9 - 2 * g2 * g9 - 4 * g4 * g9 - 2 * g5 * g9)
        MB = (2 * g3 * g4 + 2 * g3 * g5 - 2 * g1 * g6 - 2 * g2 * g6 + 4 * g3 * g7 + 2 * g6 * g7
              + 4 * g3 * g8 + 2 * g6 * g8 - 4 * g1 * g9 - 4 * g2 * g9 - 2 * g4 * g9 - 2 * g5 * g9)
        d = xAB * xab - xAb * xaB
        self.Dz_D = d
        self.W = c + self.A * self.B
        self.U = d * (1 + self.B) + MA / 4
        self.Vc = d * (1 + self.A) + MB / 4
        with np.errstate(divide="ignore", invalid="ignore"):
            self.p, self.r = P / n, R / n
            self.q, self.s = Q / n, S / n


def _single_poly(name, G):
    """The single-pop polynomial numerators (equal in both reference files)."""
    return _eval_table(tables()["kernel_single"][name], [G.g])


def _ff(n, k):
    out = np.ones_like(n)
    for t in range(k):
        out = out * (n - t)
    return out


def kernel_call(P, kind, pops):
    """Raw call in the production kernel form (genotype_kernels.py)."""
    with np.errstate(divide="ignore", invalid="ignore"):
        return _kernel_call(P, kind, pops)


def _kernel_call(P, kind, pops):
    if kind == "dd":
        i, j = pops
        if i == j:  # _DD_GENO_SINGLE_KERN (:27-71), 0 if n < 4
            a = P[i]
            v = 4.0 * _single_poly("dd_geno_single", a) / _ff(a.ns, 4)
            return np.where(a.ns < 4, 0.0, v)
        a, b = P[i], P[j]  # _DD_GENO_BETWEEN_KERN (:73-81)
        den = _ff(a.n, 2) * _ff(b.n, 2)
        return np.where(den > 0, 4.0 * a.D * b.D / np.where(den > 0, den, 1), 0.0)
    if kind == "dz":
        i, j, k = pops
        if i == j == k:  # _DZ_III_KERN (:235-272)
            a = P[i]
            v = 2.0 * _single_poly("dz_geno_single", a) / _ff(a.ns, 4)
            return np.where(a.ns < 4, 0.0, v)
        if i == j:   # _DZ_IIJ_KERN (:208-215)
            a, b = P[i], P[k]
            return 2.0 * a.U * b.A / (b.ns * a.ns * (a.ns - 1) * (a.ns - 2))
        if i == k:   # _DZ_IJI_KERN (:217-224)
            a, b = P[i], P[j]
            return 2.0 * a.Vc * b.B / (b.ns * a.ns * (a.ns - 1) * (a.ns - 2))
        if j == k:   # _DZ_IJJ_KERN (:226-233)
            a, b = P[i], P[j]
            return 2.0 * a.Dz_D * b.W / (b.ns * (b.ns - 1) * a.ns * (a.ns - 1))
        a, b, c = P[i], P[j], P[k]  # _DZ_DISTINCT_KERN (:199-206)
        return 2.0 * a.Dz_D * b.B * c.A / (a.ns * (a.ns - 1) * b.ns * c.ns)
    i, j, k, l = pops
    case = pi2_case(*pops)
    if case == "same":  # _PI2_IIII_KERN (:364-405)
        a = P[i]
        v = _single_poly("pi2_geno_single", a) / _ff(a.ns, 4)
        return np.where(a.ns < 4, 0.0, v)
    if case == "triple":
        cnt = {x: pops.count(x) for x in pops}
        t = [x for x, c in cnt.items() if c == 3][0]
        s = [x for x, c in cnt.items() if c == 1][0]
        T, S = P[t], P[s]
        d = T.n * (T.n - 1) * (T.n - 2) * S.n
        if i == t and j == t:  # _PI2_IIIJ_KERN (:319-328)
            num = T.AQ3 * S.pB + T.AS3 * S.qB
        else:                  # _PI2_IJII_KERN (:330-339)
            num = T.BP3 * S.pA + T.BR3 * S.qA
        return np.where(d > 0, num / np.where(d > 0, d, 1), 0.0)
    if case == "iikk":  # :286-291
        return P[i].BL * P[k].BR
    if case == "iikl":  # :293-299
        return P[i].BL * 0.5 * (P[k].q * P[l].s + P[k].s * P[l].q)
    if case == "ijkk":  # :301-307
        return 0.5 * (P[i].p * P[j].r + P[i].r * P[j].p) * P[k].BR
    if case == "ijij":  # :341-362
        return _ijij(P[i], P[j])
    if case == "shared":  # :309-317
        if i == k:
            s, a, b = i, j, l
        elif i == l:
            s, a, b = i, j, k
        elif j == k:
            s, a, b = j, i, l
        else:
            s, a, b = j, i, k
        S, A, B = P[s], P[a], P[b]
        return A.r * (S.C00 * B.s + S.C01 * B.q) + A.p * (S.C10 * B.s + S.C11 * B.q)
    a, b, c, d = P[i], P[j], P[k], P[l]  # :279-284
    return 0.25 * (a.p * b.r + a.r * b.p) * (c.q * d.s + c.s * d.q)


def _ijij(I, J):
    Pi, Qi, Ri, Si = I.pA, I.pB, I.qA, I.qB
    Pj, Qj, Rj, Sj = J.pA, J.pB, J.qA, J.qB
    PQi, PRi, PSi, QRi, RSi = Pi * Qi, Pi * Ri, Pi * Si, Qi * Ri, Ri * Si
    PQj, PSj, QRj, RSj = Pj * Qj, Pj * Sj, Qj * Rj, Rj * Sj
    Tci = Ri * (1.0 + I.X01 - I.X10)
    Uci = Si * (1.0 + Ri)
    num = (-0.25 * RSi * Qj - 0.25 * I.X11 * PQj + 0.25 * RSi * PQj - 0.25 * Ri * PSj
           - 0.25 * PQi * J.X11 - 0.25 * RSi * J.X11 + 0.25 * PSi * J.X11 + 0.25 * QRi * J.X11
           - 0.25 * PRi * Sj + 0.25 * Tci * Sj + 0.25 * I.X11 * PSj + 0.25 * QRi * PSj
           + 0.25 * Uci * Rj - 0.25 * I.X01 * QRj + 0.25 * PSi * QRj
           - 0.25 * PSi * Rj - 0.25 * Pi * RSj + 0.25 * I.X01 * RSj + 0.25 * PQi * RSj
           + 0.25 * I.X11 * J.Bsgn + 0.25 * I.Bsgn * J.X11 - 1.0 * I.X11 * J.X11)
    d = I.n * (I.n - 1) * J.n * (J.n - 1)
    return np.where(d > 0, num / np.where(d > 0, d, 1), 0.0)


def multi_pop_stats_geno(counts_list, nvalid_list, specs=None):
    """(N, n_stats) through the kernel form (genotype_kernels.py:881-936)."""
    from .ld_oracle import stat_specs

    P = [GenoPop(c, nv) for c, nv in zip(counts_list, nvalid_list)]
    specs = stat_specs(len(P)) if specs is None else specs
    cache = {}
    out = np.zeros((P[0].n.shape[0], len(specs)))
    for s, (_, calls) in enumerate(specs):
        acc = None
        for w, t, pp in calls:
            if (t, pp) not in cache:
                cache[(t, pp)] = kernel_call(P, t, pp)
            term = cache[(t, pp)] if (w == 1.0 and len(calls) == 1) else w * cache[(t, pp)]
            acc = term if acc is None else acc + term
        out[:, s] = acc
    return out


# ---------------------------------------------------------------------------
# exact expectation of a table under multinomial genotype sampling
# ---------------------------------------------------------------------------


def genotype_class_probs(f):
    """Moments-order genotype-class probabilities of a random-mating diploid
    whose two haplotypes are drawn from f = (f11, f10, f01, f00)."""
    haps = [(1, 1), (1, 0), (0, 1), (0, 0)]
    q = [Fraction(0)] * 9
    for a, fa in zip(haps, f):
        for b, fb in zip(haps, f):
            gA, gB = a[0] + b[0], a[1] + b[1]
            q[3 * (2 - gA) + (2 - gB)] += fa * fb
    return q


@lru_cache(maxsize=None)
def _stirling2(e, j):
    if e == j:
        return 1
    if j == 0 or j > e:
        return 0
    return j * _stirling2(e - 1, j) + _stirling2(e - 1, j - 1)


def _ffi(n, k):
    r = 1
    for t in range(k):
        r *= (n - t)
    return r


def expected_table(table, Qs, ns):
    """E[sum coef prod G^e] exactly, G_p ~ Multinomial(ns[p], Qs[p]) independent,
    via factorial moments E[prod (G_m)_(j_m)] = (n)_(sum j) prod Q_m^j_m."""
    tot = Fraction(0)
    for mon, coef in table:
        val = Fraction(coef)
        for p in range(len(Qs)):
            e = mon[9 * p:9 * p + 9]
            acc = Fraction(0)
            for js in itertools.product(*[range(x + 1) for x in e]):
                s = 1
                for x, j in zip(e, js):
                    s *= _stirling2(x, j)
                    if s == 0:
                        break
                if s == 0:
                    continue
                term = Fraction(s * _ffi(ns[p], sum(js)))
                for m, j in enumerate(js):
                    if j:
                        term *= Qs[p][m] ** j
                acc += term
            val *= acc
            if val == 0:
                break
        tot += val
    return tot


# ---------------------------------------------------------------------------
# r^2 of dosages (pairwise-complete Pearson; Rogers-Huff without missing calls,
# ld_statistics.py:492-521)
# ---------------------------------------------------------------------------


def r2_dosage(counts):
    """(N, 9) counts (n00..n22, first digit locus i) -> r^2 [N], NaN where a
    locus has no dosage variance among the individuals valid at both."""
    c = np.asarray(counts, np.float64).reshape(-1, 3, 3)   # [N, x, y]
    dose = np.array([0.0, 1.0, 2.0])
    n = c.sum(axis=(1, 2))
    cx = c.sum(axis=2)                                     # [N, x]
    cy = c.sum(axis=1)
    sx, sxx = cx @ dose, cx @ (dose * dose)
    sy, syy = cy @ dose, cy @ (dose * dose)
    sxy = np.einsum("nxy,x,y->n", c, dose, dose)
    num = n * sxy - sx * sy
    vx, vy = n * sxx - sx * sx, n * syy - sy * sy
    ok = (vx > 0) & (vy > 0)
    with np.errstate(invalid="ignore", divide="ignore"):
        return np.where(ok, num * num / np.where(ok, vx * vy, 1.0), np.nan)

