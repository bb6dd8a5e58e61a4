"""Minimal VCF reader for the LD containers.

Replaces the ``allel.read_vcf`` call of ``HaplotypeMatrix.from_vcf``
(pg_gpu/haplotype_matrix.py:292-359) and ``GenotypeMatrix.from_vcf``
(pg_gpu/genotype_matrix.py:342-390) for plain or gzip-compressed VCF text.
Only the fields the LD path needs are read: CHROM, POS, the sample names and
the GT sub-field of each call.  Allele indices are kept as written (0, 1, 2,
...); missing alleles (``.``) become -1.
"""
from __future__ import annotations

import gzip
import warnings

import numpy as np


def _open(path):
    with open(path, "rb") as fh:
        magic = fh.read(2)
    if magic == b"\x1f\x8b":
        return gzip.open(path, "rt")
    return open(path, "rt")


def _allele(tok):
    return -1 if tok in (".", "") else int(tok)


def parse_region(region):
    """'chrom', 'chrom:begin-end', 'chrom:begin-' or 'chrom:-end' ->
    (chrom, begin, end) with 1-based inclusive bounds (None where open);
    thousands separators are accepted."""
    chrom, _, rng = region.partition(":")
    begin = end = None
    if rng:
        b, _, e = rng.partition("-")
        begin = int(b.replace(",", "")) if b else None
        end = int(e.replace(",", "")) if e else None
    return chrom or None, begin, end


def _sample_selection(header_samples, samples):
    """Indices (header order) of the requested samples: names or integer
    indices; names not in the header are skipped with a warning."""
    if samples is None:
        return list(range(len(header_samples)))
    pick = set()
    for s in samples:
        if isinstance(s, (int, np.integer)):
            if not 0 <= int(s) < len(header_samples):
                raise ValueError(f"sample index {s} out of range")
            pick.add(int(s))
        elif s in header_samples:
            pick.add(header_samples.index(s))
        else:
            warnings.warn(f"sample not found: {s}", UserWarning)
    return sorted(pick)


def read_vcf_gt(path, region=None, samples=None):
    """Return dict(samples, chrom, pos int64[V], gt int8[V, n_samples, 2]).

    ``region`` keeps records on one chromosome, optionally within 1-based
    inclusive bounds; ``samples`` keeps the named (or indexed) samples in
    the file's order.  Haploid calls are stored as (a, -1).  Allele indices
    above 127 are not representable in int8 and raise ValueError.
    """
    r_chrom, r_begin, r_end = parse_region(region) if region else (None, None, None)
    header = None
    cols = None
    chrom, pos, rows = [], [], []
    with _open(path) as fh:
        for line in fh:
            if line.startswith("##"):
                continue
            if line.startswith("#CHROM"):
                header = line.rstrip("\n").split("\t")[9:]
                cols = _sample_selection(header, samples)
                continue
            if not line.strip():
                continue
            f = line.rstrip("\n").split("\t")
            if header is None:
                raise ValueError(f"{path}: data line before #CHROM header")
            if r_chrom is not None and f[0] != r_chrom:
                continue
            p = int(f[1])
            if (r_begin is not None and p < r_begin) or (r_end is not None and p > r_end):
                continue
            fmt = f[8].split(":")
            try:
                gi = fmt.index("GT")
            except ValueError:
                raise ValueError(f"{path}: record without a GT field at POS {f[1]}")
            calls = np.empty((len(cols), 2), dtype=np.int16)
            fields = f[9:9 + len(header)]
            for k, c in enumerate(cols):
                gt = fields[c].split(":")[gi]
                sep = "|" if "|" in gt else "/"
                parts = gt.split(sep)
                calls[k, 0] = _allele(parts[0])
                calls[k, 1] = _allele(parts[1]) if len(parts) > 1 else -1
            if calls.max(initial=-1) > 127:
                raise ValueError(f"{path}: allele index > 127 at POS {f[1]}")
            chrom.append(f[0])
            pos.append(p)
            rows.append(calls.astype(np.int8))
    if header is None:
        raise ValueError(f"{path}: no #CHROM header line")
    if not rows:
        raise ValueError(f"No variants found in {path}" + (f" for region {region}" if region else ""))
    return {
        "samples": [header[c] for c in cols],
        "chrom": chrom,
        "pos": np.asarray(pos, dtype=np.int64),
        "gt": np.stack(rows, axis=0),
    }


def read_pop_file(path):
    """sample -> population map from a whitespace-delimited two-column file.

    A header line whose first field is ``sample`` is skipped, as in
    pg_gpu/haplotype_matrix.py:475-512.
    """
    out = {}
    with open(path) as fh:
        for line in fh:
            parts = line.strip().split()
            if len(parts) >= 2 and parts[0] != "sample":
                out[parts[0]] = parts[1]
    return out
