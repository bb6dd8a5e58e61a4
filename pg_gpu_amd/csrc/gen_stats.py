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


GROUP = 15  # statistics evaluated per pass of the fused epilogue
# larger groups where the registers allow (one wave per SIMD): each pass
# re-derives every population's basis from the counts
GROUP_OF = {("hap", 4): 53}


def group_size(prefix, P):
    return GROUP_OF.get((prefix, P), GROUP)


def emit_set(prefix, popty, P, keyfn):
    """One device function per group of GROUP consecutive statistics (ld_names
    order); consecutive groups share no raw-call value, so splitting costs no
    re-evaluation.  Group k writes out[0 .. size-1] for statistics
    k*GROUP .. k*GROUP+size-1."""
    values, stats = stat_set_plan(P, keyfn)
    gs = group_size(prefix, P)
    ng = (len(stats) + gs - 1) // gs
    lines = [f"// P = {P}: {len(stats)} statistics from {len(values)} distinct raw-call values, "
             f"{ng} group(s) of up to {gs}",
             f"#define {prefix.upper()}_NSTATS_P{P} {len(stats)}",
             f"#define {prefix.upper()}_NGROUPS_P{P} {ng}",
             f"#define {prefix.upper()}_GS_P{P} {min(gs, len(stats))}"]
    for k in range(ng):
        grp = stats[k * gs:(k + 1) * gs]
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
    # whole set (used by the per-pair and plain kernels)
    lines.append(f"__device__ __forceinline__ void {prefix}_stats_P{P}(const {popty}* p, double* out) {{")
    for k in range(ng):
        lines.append(f"  {prefix}_stats_P{P}_g{k}(p, out + {k * gs});")
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


def emit_acc_set(P):
    """Accumulating form of the haplotype groups: acc[s] += wt * statistic s,
    with each statistic's reciprocal product folded into one FMA per distinct
    product (the averaged orientations of one statistic share it, so their
    exact numerators are summed first)."""
    values, stats = stat_set_plan(P, hap_key)
    gs = group_size("hap", P)
    ng = (len(stats) + gs - 1) // gs
    lines = []
    for k in range(ng):
        grp = stats[k * gs:(k + 1) * gs]
        used = sorted(set(v for terms in grp for v, _ in terms))
        lines.append(f"__device__ __forceinline__ void hap_acc_P{P}_g{k}(const HapPop* p, "
                     "double* acc, double wt) {")
        rkeys, rdecl, ndecl, body = {}, [], [], []
        for v in used:
            fn, args = values[v]
            a = ", ".join(f"p[{x}]" for x in args)
            ndecl.append(f"  const double n{v} = {fn}_num({a});")
        for s, terms in enumerate(grp):
            by_r = {}
            for v, w in terms:
                fn, args = values[v]
                key = tuple(sorted((args[pos], kk) for pos, kk in HAP_RECIP[fn]))
                by_r.setdefault(key, []).append((v, w))
            expr = f"acc[{s}]"
            for key, tv in list(by_r.items()):
                # (Dz(i,i,k) + Dz(i,k,i)) / 2: one merged numerator
                fns = sorted((values[v][0], values[v][1], w) for v, w in tv)
                if (len(tv) == 2 and [f for f, _, _ in fns] == ["hap_dz_iik", "hap_dz_iji"] and
                        fns[0][1] == fns[1][1] and fns[0][2] == fns[1][2] == 0.5):
                    args = ", ".join(f"p[{x}]" for x in fns[0][1])
                    name = f"m{s}"
                    ndecl.append(f"  const double {name} = hap_dz_iik_iji_avg_num({args});")
                    by_r[key] = [(name, None)]
            for key, tv in by_r.items():
                if key not in rkeys:
                    rkeys[key] = f"R{len(rkeys)}"
                    prod = " * ".join(f"p[{x}].r{kk}" for x, kk in key)
                    rdecl.append(f"  const double {rkeys[key]} = {prod} * wt;")
                if len(tv) == 1 and tv[0][1] is None:
                    num = tv[0][0]
                elif len(tv) == 1 and tv[0][1] == 1.0:
                    num = f"n{tv[0][0]}"
                else:
                    num = " + ".join(f"{w!r} * n{v}" for v, w in tv)
                expr = f"fma({num}, {rkeys[key]}, {expr})"
            body.append(f"  acc[{s}] = {expr};")
        lines += rdecl + ndecl + body
        lines.append("}")
        lines.append(f"template <> __device__ __forceinline__ void hap_acc_group<{P}, {k}>("
                     f"const HapPop* p, double* acc, double wt) {{ hap_acc_P{P}_g{k}(p, acc, wt); }}")
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
    for P in range(1, 5):
        parts.append(emit_set("hap", "HapPop", P, hap_key))
        parts.append(emit_acc_set(P))
        parts.append("")
    (HERE / "stats_hap.inc").write_text("\n".join(parts))
    parts = ["// Generated by gen_stats.py -- do not edit by hand.",
             "#pragma once",
             '#include "geno_estimators.h"', "",
             "template <int P, int G> __device__ void geno_stats_group(const GenoPop* p, double* out);",
             ""]
    for P in range(1, 5):
        parts.append(emit_set("geno", "GenoPop", P, geno_key))
        parts.append("")
    (HERE / "stats_geno.inc").write_text("\n".join(parts))


if __name__ == "__main__":
    main()
