"""Generate the per-population-count statistic sets (stats_hap.inc).

For P populations the moments-compatible statistic list (ld_names) and the
weighted raw-call specs (moments' _call_sgc averaging) follow
pg_gpu/ld_pipeline.py:357-457.  Every raw call is mapped to the estimator
kernel the reference dispatches it to (case grouping of
pg_gpu/haplotype_kernels.py:335-581); calls that evaluate the same formula on
the same populations are merged, so each distinct value is computed once per
pair and every statistic is a weighted sum of those values.

Run:  python pg_gpu_amd/csrc/gen_stats.py   (writes stats_hap.inc beside it)
"""
from __future__ import annotations

import math

from pathlib import Path

HERE = Path(__file__).resolve().parent


def ld_names(P):
    out = [f"DD_{a}_{b}" for a in range(P) for b in range(a, P)]
    out += [f"Dz_{a}_{b}_{c}" for a in range(P) for b in range(P) for c in range(b, P)]
    for a in range(P):
        for b in range(a, P):
            for c in range(a, P):
                for d in range(c, P):
                    if c == a == d and b != a:
                        continue
                    if a == c and d < b:
                        continue
                    out.append(f"pi2_{a}_{b}_{c}_{d}")
    return out


def specs(P):
    res = []
    for name in ld_names(P):
        kind, *pp = name.split("_")
        pp = tuple(int(x) for x in pp)
        if kind == "DD":
            res.append((name, [(1.0, "dd", pp)]))
        elif kind == "Dz":
            a, b, c = pp
            res.append((name, [(1.0, "dz", pp)] if b == c else
                        [(0.5, "dz", (a, b, c)), (0.5, "dz", (a, c, b))]))
        else:
            a, b, c, d = pp
            if a == b and c == d:
                calls = [(1.0, pp)] if a == c else [(0.5, (a, b, c, d)), (0.5, (c, d, a, b))]
            elif a == b:
                calls = [(0.25, x) for x in ((a, b, c, d), (a, b, d, c), (c, d, a, b), (d, c, a, b))]
            elif c == d:
                calls = [(0.25, x) for x in ((a, b, c, d), (b, a, c, d), (c, d, a, b), (c, d, b, a))]
            else:
                calls = [(0.125, x) for x in ((a, b, c, d), (a, b, d, c), (b, a, c, d), (b, a, d, c),
                                              (c, d, a, b), (d, c, a, b), (c, d, b, a), (d, c, b, a))]
            res.append((name, [(w, "pi2", x) for w, x in calls]))
    return res


def hap_key(kind, pp):
    """Canonical (device function, argument pops) for one raw haplotype call."""
    if kind == "dd":
        i, j = pp
        return ("hap_dd_same", (i,)) if i == j else ("hap_dd_pair", tuple(sorted(pp)))
    if kind == "dz":
        i, j, k = pp
        if i == j == k:
            return ("hap_dz_same", (i,))
        if i == j:
            return ("hap_dz_iik", (i, k))
        if i == k:
            return ("hap_dz_iji", (i, j))
        if j == k:
            return ("hap_dz_ijj", (i, j))
        return ("hap_dz_ijk", (i, j, k))
    i, j, k, l = pp
    cnt = {}
    for p in pp:
        cnt[p] = cnt.get(p, 0) + 1
    if len(cnt) == 1:
        return ("hap_pi2_same", (i,))
    if max(cnt.values()) == 3:
        t = [p for p, c in cnt.items() if c == 3][0]
        s = [p for p, c in cnt.items() if c == 1][0]
        return ("hap_pi2_triple", (t, s))
    if i == j and k == l:
        return ("hap_pi2_iikk", (i, k))
    if i == j:
        return ("hap_pi2_iikl", (i, min(k, l), max(k, l)))
    if k == l:
        return ("hap_pi2_ijkk", (min(i, j), max(i, j), k))
    if (i == k and j == l) or (i == l and j == k):
        return ("hap_pi2_ijij", (min(i, j), max(i, j)))
    if len(cnt) == 3:
        if i == k:
            s, a, b = i, j, l
        elif i == l:
            s, a, b = i, j, k
        elif j == k:
            s, a, b = j, i, l
        else:
            s, a, b = j, i, k
        return ("hap_pi2_shared", (s, a, b))
    return ("hap_pi2_alldiff", (min(i, j), max(i, j), min(k, l), max(k, l)))


def stat_set_plan(P, keyfn):
    """-> (values[(func, args)], stats[[(weight, value_index)]])."""
    values, index, stats = [], {}, []
    for _, calls in specs(P):
        terms = {}
        for w, kind, pp in calls:
            key = keyfn(kind, pp)
            if key not in index:
                index[key] = len(values)
                values.append(key)
            terms[index[key]] = terms.get(index[key], 0.0) + w
        stats.append(sorted(terms.items()))
    return values, stats


GROUP = 15  # statistics evaluated per pass of the fused epilogue (consecutive groups)

# Passes by population set: each pass evaluates the statistics whose
# populations form one of the listed sets, so a pass derives the bases of at
# most three populations (a basis is ~23 FP64 values).  Used where evaluating
# consecutive statistics keeps the bases of every population live at once and
# spills: four haplotype populations, three and four genotype populations.
_P4_PASSES = [
    [{0}, {1}, {0, 1}],
    [{2}, {3}, {2, 3}],
    [{0, 2}, {0, 3}, {0, 2, 3}],
    [{1, 2}, {1, 3}, {1, 2, 3}],
    [{0, 1, 2}, {0, 1, 3}, {0, 1, 2, 3}],
]
_P3_PASSES = [
    [{0}, {1}, {0, 1}],
    [{2}, {0, 2}],
    [{1, 2}, {0, 1, 2}],
]
# Genotype sets use the same passes.  Smaller passes (at most three
# populations, eight passes for four) remove the spills of the three-population
# kernel and most of the four-population one, but re-derive more bases: they
# measured slower (3 populations of 60/150/400 over 500k sites: 205 -> 216 ms;
# 4 of 20/60/150/400: 447 -> 500 ms).
POPSET_PASSES = {("hap", 4): _P4_PASSES, ("geno", 3): _P3_PASSES, ("geno", 4): _P4_PASSES}


def stat_pops(name):
    return frozenset(int(x) for x in name.split("_")[1:])


def plan_groups(prefix, P):
    """Statistic indices (ld_names order) of each pass of the fused epilogue."""
    names = ld_names(P)
    passes = POPSET_PASSES.get((prefix, P))
    if passes is None:
        return [list(range(k, min(k + GROUP, len(names)))) for k in range(0, len(names), GROUP)]
    groups = []
    for sets in passes:
        want = set(frozenset(x) for x in sets)
        groups.append([i for i, n in enumerate(names) if stat_pops(n) in want])
    assert sorted(i for g in groups for i in g) == list(range(len(names)))
    return groups


def emit_group_info(prefix, P, groups, n_stats):
    """Per-group size and statistic -> output column map (7-bit fields, 9 per
    64-bit word, selected with immediates: no memory access)."""
    lines = []
    for k, grp in enumerate(groups):
        lines.append(f"template <> __device__ constexpr int {prefix}_group_size<{P}, {k}>() "
                     f"{{ return {len(grp)}; }}")
        words = [sum(c << (7 * i) for i, c in enumerate(grp[w:w + 9]))
                 for w in range(0, len(grp), 9)]
        sel = "".join(f"q < {9 * (i + 1)} ? 0x{w:x}ull : " for i, w in enumerate(words)) + "0ull"
        lines.append(f"template <> __device__ __forceinline__ int {prefix}_stat_col<{P}, {k}>(int q) {{")
        lines.append(f"  if (q >= {len(grp)}) return {n_stats} + q - {len(grp)};")
        lines.append(f"  const unsigned long long w = {sel};")
        lines.append("  return (int)((w >> (7 * (q % 9))) & 127ull);")
        lines.append("}")
    return lines


def emit_set(prefix, popty, P, keyfn):
    """One device function per group of GROUP consecutive statistics (ld_names
    order); consecutive groups share no raw-call value, so splitting costs no
    re-evaluation.  Group k writes out[0 .. size-1] for statistics
    k*GROUP .. k*GROUP+size-1."""
    values, stats = stat_set_plan(P, keyfn)
    groups = plan_groups(prefix, P)
    ng = len(groups)
    gs = max(len(g) for g in groups)
    lines = [f"// P = {P}: {len(stats)} statistics from {len(values)} distinct raw-call values, "
             f"{ng} group(s) of up to {gs}",
             f"#define {prefix.upper()}_NSTATS_P{P} {len(stats)}",
             f"#define {prefix.upper()}_NGROUPS_P{P} {ng}",
             f"#define {prefix.upper()}_GS_P{P} {gs}"]
    lines += emit_group_info(prefix, P, groups, len(stats))
    for k in range(ng):
        grp = [stats[i] for i in groups[k]]
        used = sorted(set(v for terms in grp for v, _ in terms))
        lines.append(f"__device__ __forceinline__ void {prefix}_stats_P{P}_g{k}(const {popty}* p, "
                     "double* out) {")
        for v in used:
            fn, args = values[v]
            a = ", ".join(f"p[{x}]" for x in args)
            lines.append(f"  const double v{v} = {fn}({a});")
        for s, terms in enumerate(grp):
            if len(terms) == 1 and terms[0][1] == 1.0:
                expr = f"v{terms[0][0]}"
            else:
                expr = " + ".join(f"{w!r} * v{v}" for v, w in terms)
            lines.append(f"  out[{s}] = {expr};")
        lines.append("}")
        lines.append(f"template <> __device__ __forceinline__ void {prefix}_stats_group<{P}, {k}>("
                     f"const {popty}* p, double* out) {{ {prefix}_stats_P{P}_g{k}(p, out); }}")
    # whole set in ld_names order (used by the per-pair and plain kernels)
    lines.append(f"__device__ __forceinline__ void {prefix}_stats_P{P}(const {popty}* p, double* out) {{")
    for k in range(ng):
        if groups[k] == list(range(groups[k][0], groups[k][0] + len(groups[k]))):
            lines.append(f"  {prefix}_stats_P{P}_g{k}(p, out + {groups[k][0]});")
        else:
            lines.append(f"  {{ double t[{len(groups[k])}]; {prefix}_stats_P{P}_g{k}(p, t);")
            lines.append("    " + " ".join(f"out[{c}] = t[{q}];" for q, c in enumerate(groups[k])) + " }")
    lines.append("}")
    return "\n".join(lines)


# Reciprocal falling factorials of each haplotype raw call: (argument
# position, k) pairs, the call's value being num * prod 1/(n_arg)_k.
HAP_RECIP = {
    "hap_dd_same": ((0, 4),), "hap_dd_pair": ((0, 2), (1, 2)),
    "hap_dz_same": ((0, 4),), "hap_dz_iik": ((0, 3), (1, 1)), "hap_dz_iji": ((0, 3), (1, 1)),
    "hap_dz_ijj": ((0, 2), (1, 2)), "hap_dz_ijk": ((0, 2), (1, 1), (2, 1)),
    "hap_pi2_same": ((0, 4),), "hap_pi2_triple": ((0, 3), (1, 1)),
    "hap_pi2_iikk": ((0, 2), (1, 2)), "hap_pi2_iikl": ((0, 2), (1, 1), (2, 1)),
    "hap_pi2_ijkk": ((2, 2), (0, 1), (1, 1)), "hap_pi2_ijij": ((0, 2), (1, 2)),
    "hap_pi2_shared": ((0, 2), (1, 1), (2, 1)),
    "hap_pi2_alldiff": ((0, 1), (1, 1), (2, 1), (3, 1)),
}


# Constant power of two of each raw call whose device function has a _raw
# form (hap_X_num = hap_X_raw * INNER[X]); the merged Dz(i,i,k)/Dz(i,k,i)
# average carries 1/2 the same way.
HAP_INNER = {
    "hap_pi2_triple": 0.5, "hap_pi2_iikl": 0.5, "hap_pi2_ijkk": 0.5,
    "hap_pi2_ijij": 0.0625, "hap_pi2_shared": 0.0625, "hap_pi2_alldiff": 0.25,
}
HAP_AVG_INNER = 0.5
# population counts whose groups accumulate raw numerators: every count since
# the population-set passes (4 populations of 40-800 haplotypes over 500k
# sites: 159.1 ms scaled, 137.3 ms raw, both with the register operand ring)
HAP_FOLD_P = (1, 2, 3, 4)


def _log2_exact(x):
    e = math.frexp(x)[1] - 1
    if x <= 0 or 2.0 ** e != x:
        raise ValueError(f"{x} is not a power of two")
    return e


def emit_acc_set(P):
    """Accumulating form of the haplotype groups: acc[s] += wt * statistic s
    / 2^-e_s, with each statistic's reciprocal product folded into one FMA per
    distinct product (the averaged orientations of one statistic share it, so
    their exact numerators are summed first).  e_s is the statistic's largest
    constant power of two (the orientation weights times the raw calls' own
    constants); the sums are scaled by 2^-e_s where they become limbs
    (hap_scale_group)."""
    values, stats = stat_set_plan(P, hap_key)
    groups = plan_groups("hap", P)
    ng = len(groups)
    fold = P in HAP_FOLD_P
    inner = HAP_INNER if fold else {}
    lines = []
    for k in range(ng):
        grp = [stats[i] for i in groups[k]]
        used = sorted(set(v for terms in grp for v, _ in terms))
        lines.append(f"__device__ __forceinline__ void hap_acc_P{P}_g{k}(const HapPop* p, "
                     "double* acc, double wt) {")
        rkeys, rdecl, ndecl, body, exps = {}, [], [], [], []
        for v in used:
            fn, args = values[v]
            a = ", ".join(f"p[{x}]" for x in args)
            form = "raw" if fn in inner else "num"
            ndecl.append(f"  const double n{v} = {fn}_{form}({a});")
        for s, terms in enumerate(grp):
            # R key -> [(numerator name, constant multiplier)]
            by_r = {}
            for v, w in terms:
                fn, args = values[v]
                key = tuple(sorted((args[pos], kk) for pos, kk in HAP_RECIP[fn]))
                by_r.setdefault(key, []).append((f"n{v}", w * inner.get(fn, 1.0), (fn, args, w)))
            for key, tv in list(by_r.items()):
                # (Dz(i,i,k) + Dz(i,k,i)) / 2: one merged numerator
                fns = sorted(x[2] for x in tv)
                if (len(tv) == 2 and [f for f, _, _ in fns] == ["hap_dz_iik", "hap_dz_iji"] and
                        fns[0][1] == fns[1][1] and fns[0][2] == fns[1][2] == 0.5):
                    args = ", ".join(f"p[{x}]" for x in fns[0][1])
                    name = f"m{s}"
                    form = "raw" if fold else "num"
                    ndecl.append(f"  const double {name} = hap_dz_iik_iji_avg_{form}({args});")
                    by_r[key] = [(name, HAP_AVG_INNER if fold else 1.0, None)]
            scale = max(m for tv in by_r.values() for _, m, _ in tv) if fold else 1.0
            exps.append(-_log2_exact(scale))
            expr = f"acc[{s}]"
            for key, tv in by_r.items():
                if key not in rkeys:
                    rkeys[key] = f"R{len(rkeys)}"
                    prod = " * ".join(f"p[{x}].r{kk}" for x, kk in key)
                    rdecl.append(f"  const double {rkeys[key]} = {prod} * wt;")
                parts = []
                for name, m, _ in tv:
                    rel = m / scale
                    _log2_exact(rel)
                    parts.append(name if rel == 1.0 else f"{rel!r} * {name}")
                num = " + ".join(parts)
                expr = f"fma({num}, {rkeys[key]}, {expr})"
            body.append(f"  acc[{s}] = {expr};")
        lines += rdecl + ndecl + body
        lines.append("}")
        lines.append(f"template <> __device__ __forceinline__ void hap_acc_group<{P}, {k}>("
                     f"const HapPop* p, double* acc, double wt) {{ hap_acc_P{P}_g{k}(p, acc, wt); }}")
        # e_s in 4-bit fields, 16 statistics per word; 0 past the group's end
        assert all(0 <= e < 16 for e in exps)
        words = [sum(e << (4 * i) for i, e in enumerate(exps[w:w + 16]))
                 for w in range(0, len(exps), 16)]
        sel = "".join(f"q < {16 * (i + 1)} ? 0x{w:x}ull : " for i, w in enumerate(words)) + "0ull"
        lines.append(f"template <> __device__ __forceinline__ int hap_scale_group<{P}, {k}>(int q) {{")
        lines.append(f"  const unsigned long long w = {sel};")
        lines.append("  return (int)((w >> (4 * (q & 15))) & 15ull);")
        lines.append("}")
    return "\n".join(lines)


def geno_key(kind, pp):
    """Canonical device function for one raw genotype call (case dispatch of
    pg_gpu/genotype_kernels.py:561-878)."""
    if kind == "dd":
        i, j = pp
        return ("geno_dd_same", (i,)) if i == j else ("geno_dd_pair", tuple(sorted(pp)))
    if kind == "dz":
        i, j, k = pp
        if i == j == k:
            return ("geno_dz_same", (i,))
        if i == j:
            return ("geno_dz_iik", (i, k))
        if i == k:
            return ("geno_dz_iji", (i, j))
        if j == k:
            return ("geno_dz_ijj", (i, j))
        return ("geno_dz_ijk", (i, j, k))
    i, j, k, l = pp
    cnt = {}
    for p in pp:
        cnt[p] = cnt.get(p, 0) + 1
    if len(cnt) == 1:
        return ("geno_pi2_same", (i,))
    if max(cnt.values()) == 3:
        t = [p for p, c in cnt.items() if c == 3][0]
        s = [p for p, c in cnt.items() if c == 1][0]
        return ("geno_pi2_iiij" if (i == t and j == t) else "geno_pi2_ijii", (t, s))
    if i == j and k == l:
        return ("geno_pi2_iikk", (i, k))
    if i == j:
        return ("geno_pi2_iikl", (i, min(k, l), max(k, l)))
    if k == l:
        return ("geno_pi2_ijkk", (min(i, j), max(i, j), k))
    if (i == k and j == l) or (i == l and j == k):
        return ("geno_pi2_ijij", (i, j))
    if len(cnt) == 3:
        if i == k:
            s, a, b = i, j, l
        elif i == l:
            s, a, b = i, j, k
        elif j == k:
            s, a, b = j, i, l
        else:
            s, a, b = j, i, k
        return ("geno_pi2_shared", (s, a, b))
    return ("geno_pi2_alldiff", (min(i, j), max(i, j), min(k, l), max(k, l)))


def main():
    parts = ["// Generated by gen_stats.py -- do not edit by hand.",
             "#pragma once",
             '#include "hap_estimators.h"', "",
             f"#define STATS_GROUP {GROUP}",
             "template <int P, int G> __device__ void hap_stats_group(const HapPop* p, double* out);",
             ""]
    parts.append("template <int P, int G> __device__ void hap_acc_group(const HapPop* p, double* acc, "
                 "double wt);")
    parts.append("// exponent e of the 2^-e the fused kernel applies to statistic q of group G")
    parts.append("template <int P, int G> __device__ int hap_scale_group(int q);")
    parts.append("// statistics in pass G and the output column of its q-th statistic")
    parts.append("template <int P, int G> __device__ constexpr int hap_group_size();")
    parts.append("template <int P, int G> __device__ int hap_stat_col(int q);")
    for P in range(1, 5):
        parts.append(emit_set("hap", "HapPop", P, hap_key))
        parts.append(emit_acc_set(P))
        parts.append("")
    (HERE / "stats_hap.inc").write_text("\n".join(parts))
    parts = ["// Generated by gen_stats.py -- do not edit by hand.",
             "#pragma once",
             '#include "geno_estimators.h"', "",
             "template <int P, int G> __device__ void geno_stats_group(const GenoPop* p, double* out);",
             "template <int P, int G> __device__ constexpr int geno_group_size();",
             "template <int P, int G> __device__ int geno_stat_col(int q);",
             ""]
    for P in range(1, 5):
        parts.append(emit_set("geno", "GenoPop", P, geno_key))
        parts.append("")
    (HERE / "stats_geno.inc").write_text("\n".join(parts))


if __name__ == "__main__":
    main()
