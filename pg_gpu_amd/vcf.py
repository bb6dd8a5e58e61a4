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

import numpy as np


def _open(path):
    with open(path, "rb") as fh:
        magic = fh.read(2)
    if magic == b"\x1f\x8b":
        return gzip.open(path, "rt")
    return open(path, "rt")


def _allele(tok):
    return -1 if tok in (".", "") else int(tok)


def read_vcf_gt(path):
    """Return dict(samples, chrom, pos int64[V], gt int8[V, n_samples, 2]).

    Haploid calls are stored as (a, -1).  Allele indices above 127 are not
    representable in int8 and raise ValueError.
    """
    samples = None
    chrom, pos, rows = [], [], []
    with _open(path) as fh:
        for line in fh:
            if line.startswith("##"):
                continue
            if line.startswith("#CHROM"):
                samples = line.rstrip("\n").split("\t")[9:]
                continue
            if not line.strip():
                continue
            f = line.rstrip("\n").split("\t")
            if samples is None:
                raise ValueError(f"{path}: data line before #CHROM header")
            fmt = f[8].split(":")
            try:
                gi = fmt.index("GT")
            except ValueError:
                raise ValueError(f"{path}: record without a GT field at POS {f[1]}")
            calls = np.empty((len(samples), 2), dtype=np.int16)
            for s, field in enumerate(f[9:9 + len(samples)]):
                gt = field.split(":")[gi]
                sep = "|" if "|" in gt else "/"
                parts = gt.split(sep)
                calls[s, 0] = _allele(parts[0])
                calls[s, 1] = _allele(parts[1]) if len(parts) > 1 else -1
            if calls.max(initial=-1) > 127:
                raise ValueError(f"{path}: allele index > 127 at POS {f[1]}")
            chrom.append(f[0])
            pos.append(int(f[1]))
            rows.append(calls.astype(np.int8))
    if samples is None:
        raise ValueError(f"{path}: no #CHROM header line")
    if not rows:
        raise ValueError(f"No variants found in {path}")
    return {
        "samples": samples,
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
