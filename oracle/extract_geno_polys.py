"""Extract the genotype LD estimator polynomials from the reference source text.

The unbiased genotype-count estimators of DD, Dz and pi2 (Mathematica-derived
in moments, transcribed by pg_gpu) are polynomials with up to ~130
monomials.  Their coefficients are mathematical constants that any
implementation must reproduce, so instead of retyping them this script reads
them, as data, from two independent places in the reference:

  A. pg_gpu/ld_statistics_genotype.py -- the readable CuPy polynomials
     (``numer = ...``, ``denom = ...``, ``valid = ...``, ``return s * _safe_div``)
  B. pg_gpu/genotype_kernels.py -- the production kernel strings for the three
     single-population polynomials (``poly`` of _DD_GENO_SINGLE_KERN,
     ``quartic`` of _DZ_III_KERN, ``corr`` of _PI2_IIII_KERN)

and writes them as exact rational coefficient tables to
oracle/geno_polys.json.gz.  Nothing is executed: the Python source is parsed
with ``ast`` and the C expressions with sympy's parser.  The tests check
that table A and table B agree symbolically and that every estimator in table
A is unbiased (tests/test_oracle_geno.py).

Run (needs the reference checkout):  python oracle/extract_geno_polys.py /root/reference
"""
from __future__ import annotations

import ast
import gzip
import json
import re
import sys
from pathlib import Path

import sympy as sp

HERE = Path(__file__).resolve().parent

# symbol names: single-pop functions use g1..g9, multi-pop g{pop}{1..9}
G = {f"g{m}": sp.Symbol(f"g0_{m}") for m in range(1, 10)}
for k in range(1, 5):
    for m in range(1, 10):
        G[f"g{k}{m}"] = sp.Symbol(f"g{k - 1}_{m}")
NS = {"n": sp.Symbol("n0"), **{f"n{k}": sp.Symbol(f"n{k - 1}") for k in range(1, 5)}}


def _sympify(src):
    src = re.sub(r"(\d+)\.0\b", r"\1", src)            # integer-valued float literals
    return sp.sympify(src, locals={**G, **NS}, rational=True)


def _poly_table(expr, npops):
    gens = [sp.Symbol(f"g{p}_{m}") for p in range(npops) for m in range(1, 10)]
    P = sp.Poly(sp.expand(expr), *gens)
    return [[list(mon), str(coef)] for mon, coef in sorted(P.terms())]


def _scale_of(ret):
    """``s * _safe_div(numer, denom, valid)`` -> s (1 if bare)."""
    if isinstance(ret, ast.BinOp) and isinstance(ret.op, ast.Mult):
        return float(ret.left.value)
    return 1.0


def readable_tables(src_text):
    tree = ast.parse(src_text)
    out = {}
    for fn in tree.body:
        if not isinstance(fn, ast.FunctionDef) or fn.name.startswith("_") or \
                fn.name == "sigma_d2_geno":
            continue
        npops = len(fn.args.args)
        parts = {}
        for st in fn.body:
            if isinstance(st, ast.Assign) and isinstance(st.targets[0], ast.Name) and \
                    st.targets[0].id in ("numer", "denom", "valid"):
                parts[st.targets[0].id] = ast.get_source_segment(src_text, st.value)
            if isinstance(st, ast.Return):
                parts["scale"] = _scale_of(st.value)
        numer = _sympify(parts["numer"])
        out[fn.name] = {
            "n_pops": npops,
            "numer": _poly_table(numer, npops),
            "denom": str(_sympify(parts["denom"])),
            "valid": parts["valid"],
            "scale": parts["scale"],
        }
    return out


def kernel_tables(src_text):
    """Single-pop polynomials of the production kernels, in moments order."""
    nmap = {f"n{m}": G[f"g{m}"] for m in range(1, 10)}

    def grab(kernel, var):
        k = src_text.index(kernel)
        s = src_text.index(f"double {var}=", k) + len(f"double {var}=")
        e = src_text.index(";", s)
        body = re.sub(r"(\d+)\.0\b", r"\1", src_text[s:e].replace("\n", " "))
        return sp.sympify(body, locals={**nmap, "Dg": sp.Symbol("Dg")}, rational=True)

    g = [G[f"g{m}"] for m in range(1, 10)]
    Dg = -(g[1] / 2 + g[2] + g[4] / 4 + g[5] / 2) * (g[3] / 2 + g[4] / 4 + g[6] + g[7] / 2) + \
        (g[0] + g[1] / 2 + g[3] / 2 + g[4] / 4) * (g[4] / 4 + g[5] / 2 + g[7] / 2 + g[8])
    x00 = g[0] + g[1] / 2 + g[3] / 2 + g[4] / 4
    x01 = g[1] / 2 + g[2] + g[4] / 4 + g[5] / 2
    x10 = g[3] / 2 + g[4] / 4 + g[6] + g[7] / 2
    x11 = g[4] / 4 + g[5] / 2 + g[7] / 2 + g[8]
    A = -g[0] + g[2] - g[3] + g[5] - g[6] + g[8]
    B = -g[0] - g[1] - g[2] + g[6] + g[7] + g[8]
    a1 = g[0] + g[1] + g[2] + (g[3] + g[4] + g[5]) / 2
    a2 = g[0] + g[1] / 2 + g[3] + g[4] / 2 + g[6] + g[7] / 2
    a3 = g[1] / 2 + g[2] + g[4] / 2 + g[5] + g[7] / 2 + g[8]
    a4 = (g[3] + g[4] + g[5]) / 2 + g[6] + g[7] + g[8]
    poly = grab("_DD_GENO_SINGLE_KERN", "poly")
    quartic = grab("_DZ_III_KERN", "quartic")
    corr = grab("_PI2_IIII_KERN", "corr") * 16   # the kernel's corr carries a 1/16 factor
    dd = poly / 16 + Dg ** 2
    dz = quartic / 4 + A * B * (x00 * x11 - x01 * x10)
    pi2 = a1 * a2 * a3 * a4 + corr / 16
    full = {"dd_geno_single": _poly_table(dd, 1), "dz_geno_single": _poly_table(dz, 1),
            "pi2_geno_single": _poly_table(pi2, 1)}
    # integer sub-polynomials the device code evaluates (genotype_kernels.py
    # structure: DD = poly/16 + Dg^2, Dz = quartic/4 + A B D, pi2 = a1a2a3a4 + corr/16)
    parts = {"dd_poly": _poly_table(poly, 1), "dz_quartic": _poly_table(quartic, 1),
             "pi2_corr": _poly_table(corr, 1)}
    return full, parts


def main(ref_root):
    ref = Path(ref_root)
    readable = readable_tables((ref / "pg_gpu/ld_statistics_genotype.py").read_text())
    kern, parts = kernel_tables((ref / "pg_gpu/genotype_kernels.py").read_text())
    # kernel sub-polynomial tables, checked against the device's centred forms
    # by tests/test_oracle_geno.py
    csrc = HERE / "geno_kernel_polys.json"
    csrc.write_text(json.dumps({
        "source": "integer sub-polynomials of the single-population genotype estimators "
                  "(pg_gpu/genotype_kernels.py:27-71, :235-272, :364-405); variables g0_1..g0_9 "
                  "= moments genotype classes g1 (n22) .. g9 (n00)",
        **parts}, separators=(",", ":")))
    doc = {
        "source": "pg_gpu/ld_statistics_genotype.py (readable) and pg_gpu/genotype_kernels.py "
                  "(kernel single-pop polynomials); variables g{p}_{m}: population slot p, "
                  "moments genotype class m (g1 = n22, ..., g9 = n00)",
        "readable": readable,
        "kernel_single": kern,
    }
    with gzip.GzipFile(HERE / "geno_polys.json.gz", "wb", mtime=0) as fh:
        fh.write(json.dumps(doc, separators=(",", ":")).encode())


def load(path=HERE / "geno_polys.json.gz"):
    with gzip.open(path, "rt") as fh:
        return json.load(fh)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
