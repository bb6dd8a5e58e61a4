"""Drop-in for moments.LD.Parsing.compute_ld_statistics on one MI355X
(pg_gpu/moments_ld.py).

Same arguments and the same output dict ({'bins', 'sums', 'stats', 'pops'}),
so moments' bootstrap and inference code can consume it unchanged.  The
binned sums come from the fused MFMA kernel (bit-reproducible fixed-point
accumulation), the heterozygosity statistics from pgld_heterozygosity.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _device as dv
from . import _lib, filters
from .engine import LDPanels, decode_limbs
from .genotype_matrix import GenotypeMatrix
from .haplotype_matrix import HaplotypeMatrix
from .ld_pipeline import het_names, ld_names


def compute_ld_statistics(vcf_file=None, rec_map_file=None, pop_file=None, pops=None,
                          r_bins=None, bp_bins=None, use_genotypes=False, report=True,
                          ac_filter=True, haplotype_matrix=None, genotype_matrix=None,
                          accessible_bed=None):
    """Binned LD statistics for 1-4 populations (moments_ld.py:40-203)."""
    if pops is None:
        pops = ["pop0", "pop1"]
    num_pops = len(pops)
    if num_pops < 1 or num_pops > 4:
        raise ValueError("1-4 populations supported")
    if r_bins is None and bp_bins is None:
        raise ValueError("Either r_bins or bp_bins must be provided")

    if use_genotypes:
        if genotype_matrix is not None:
            mat = genotype_matrix
        elif haplotype_matrix is not None:
            mat = GenotypeMatrix.from_haplotype_matrix(haplotype_matrix)
        else:
            if vcf_file is None:
                raise ValueError("vcf_file or genotype_matrix required")
            if pop_file is None:
                raise ValueError("pop_file is required when loading from VCF")
            if report:
                print(f"Loading {vcf_file} (genotypes) ...")
            mat = GenotypeMatrix.from_vcf(vcf_file)
            mat.load_pop_file(pop_file, pops=pops)
            if ac_filter:
                mat = mat.apply_biallelic_filter()
            if accessible_bed is not None and not mat.has_accessible_mask:
                mat.set_accessible_mask(accessible_bed)
        if report:
            print(f"  {mat.num_individuals} individuals, {mat.num_variants:,} variants")
    else:
        if haplotype_matrix is not None:
            mat = haplotype_matrix
        else:
            if vcf_file is None:
                raise ValueError("vcf_file or haplotype_matrix is required")
            if pop_file is None:
                raise ValueError("pop_file is required when loading from VCF")
            if report:
                print(f"Loading {vcf_file} ...")
            mat = HaplotypeMatrix.from_vcf(vcf_file)
            mat.load_pop_file(pop_file, pops=pops)
            if ac_filter:
                mat = mat.apply_biallelic_filter()
            if accessible_bed is not None and not mat.has_accessible_mask:
                mat.set_accessible_mask(accessible_bed)
        if report:
            print(f"  {mat.num_haplotypes} hap, {mat.num_variants:,} variants")
    mat.transfer_to_gpu()

    pos_all = dv.to_numpy(mat.positions).astype(np.int64)
    if r_bins is not None:
        if rec_map_file is None:
            raise ValueError("rec_map_file required with r_bins")
        bins = np.asarray(r_bins, dtype=np.float64)
        gen_all = _interpolate_genetic_distances(pos_all, rec_map_file)
        max_bp = _max_bp_for_r_dist(pos_all, gen_all, float(bins[-1]))
    else:
        bins = np.asarray(bp_bins, dtype=np.float64)
        gen_all = None
        max_bp = float(bins[-1])
    n_bins = len(bins) - 1
    if report:
        print(f"  Computing LD ({n_bins} bins, {num_pops} pops) ...")

    data = mat.genotypes if use_genotypes else mat.haplotypes
    pop_rows = [[int(i) for i in mat.sample_sets[p]] for p in pops]
    mode = "geno" if use_genotypes else "hap"
    if use_genotypes:
        # heterozygosity over every site of the matrix (moments_ld.py:346-388
        # is called on the unfiltered genotypes): tallies only, no panels
        het = heterozygosity(LDPanels(data, pop_rows, mode, pack=False))
        # sites biallelic across the union of the populations' individuals
        # (moments_ld.py:256-276), tested and compacted on the device
        idx = _union_biallelic(data, pop_rows)
        data = filters.take_columns(data, idx)
        idx_np = idx.cpu().numpy()
        pos = pos_all[idx_np]
        coords = None if gen_all is None else gen_all[idx_np]
        pan = LDPanels(data, pop_rows, mode)
    else:
        pos, coords = pos_all, gen_all
        pan = LDPanels(data, pop_rows, mode)
        het = heterozygosity(pan)   # the same site tallies
    res = pan.binned(pos, bins, max_bp, coords=coords)
    if report:
        print("  Done.")
    hn = het_names(num_pops)
    return {
        "bins": [(float(bins[i]), float(bins[i + 1])) for i in range(n_bins)],
        "sums": [res.sums[i] for i in range(n_bins)] + [np.array([het[h] for h in hn])],
        "stats": (ld_names(num_pops), hn),
        "pops": pops,
    }


def heterozygosity(pan):
    """{H_i_j: float} summed over sites (moments_ld.py:346-388)."""
    t = dv.torch()
    P = pan.P
    n_het = P * (P + 1) // 2
    out = t.zeros((n_het, 3), dtype=t.int64, device=pan.device)
    bad = t.zeros(n_het, dtype=t.int32, device=pan.device)
    _lib.check(_lib.lib().pgld_heterozygosity(ctypes.byref(pan.desc), _lib.ptr(pan.site_stats),
                                              _lib.ptr(out), _lib.ptr(bad),
                                              _lib.stream_handle()))
    raw = out.cpu().numpy().reshape(n_het, 1, 3)
    vals = np.array([decode_limbs(np.concatenate([raw[e:e + 1], np.zeros((1, 1, 3), np.int64)],
                                                 axis=1), 1)[0][0, 0] for e in range(n_het)])
    nb = bad.cpu().numpy()
    vals[nb > 0] = np.nan
    return dict(zip(het_names(P), (float(v) for v in vals)))


def _union_biallelic(g, pop_rows):
    """Indices (CUDA int64) of the sites biallelic across the union of the
    selected populations' individuals, each individual counted once
    (moments_ld.py:256-279)."""
    rows = sorted(set(r for pr in pop_rows for r in pr))
    return filters.keep_index(filters.biallelic_keep(g, _lib.FILTER_GENO_BIALLELIC, rows))


def _interpolate_genetic_distances(positions, rec_map_file):
    """Per-variant map positions in Morgans by linear interpolation of a
    (pos, cM) map (moments_ld.py:211-223)."""
    mp, mv = [], []
    with open(rec_map_file) as fh:
        for line in fh:
            parts = line.strip().split()
            if len(parts) >= 2:
                try:
                    mp.append(float(parts[0]))
                    mv.append(float(parts[1]) / 100.0)
                except ValueError:
                    continue
    return np.interp(positions, np.array(mp), np.array(mv))


def _max_bp_for_r_dist(positions, gen_dists, max_r):
    """Conservative bp cutoff for a recombination-distance cutoff
    (moments_ld.py:226-245)."""
    chrom_len = float(positions[-1] - positions[0])
    if len(positions) < 2:
        return chrom_len
    bp = np.diff(positions).astype(np.float64)
    rd = np.diff(gen_dists)
    ok = bp > 0
    if not np.any(ok):
        return chrom_len
    rates = rd[ok] / bp[ok]
    min_rate = np.min(rates[rates > 0]) if np.any(rates > 0) else max_r / chrom_len
    return min(max_r / min_rate * 1.1, chrom_len)
