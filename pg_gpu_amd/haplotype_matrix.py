"""HaplotypeMatrix container and its binned LD methods (drop-in for the LD
part of pg_gpu/haplotype_matrix.py).

Data: ``haplotypes`` (n_haplotypes, n_variants) int8 with 0/1 and -1 for
missing, ``positions`` (n_variants,), ``sample_sets`` {population: [row
indices]}.  On the CPU the arrays are NumPy; after ``transfer_to_gpu`` they
are torch CUDA tensors.  The binned LD methods run the fused MFMA kernel
(pgld_binned_fused) via engine.LDPanels.
"""
from __future__ import annotations

from collections import OrderedDict

import numpy as np

from . import _device as dv
from . import _lib
from .accessible import read_bed_mask


class HaplotypeMatrix:
    """Phased haplotypes with positions and population labels
    (pg_gpu/haplotype_matrix.py:10-290)."""

    def __init__(self, genotypes, positions, chrom_start=None, chrom_end=None,
                 sample_sets=None, n_total_sites=None, samples=None, accessible_mask=None):
        if genotypes is None or getattr(genotypes, "size", None) == 0 or \
                (hasattr(genotypes, "numel") and genotypes.numel() == 0):
            raise ValueError("genotypes cannot be empty")
        if positions is None or len(positions) == 0:
            raise ValueError("positions cannot be empty")
        t = dv.torch()
        if isinstance(genotypes, t.Tensor) and genotypes.is_cuda:
            self._device = "GPU"
            positions = dv.as_tensor(positions, t.int64, genotypes.device) \
                if not (isinstance(positions, t.Tensor) and positions.is_cuda) else positions
        elif isinstance(genotypes, (np.ndarray, t.Tensor)):
            self._device = "CPU"
            genotypes = dv.to_numpy(genotypes)
            positions = dv.to_numpy(positions)
        else:
            raise ValueError("genotypes must be a numpy array or torch tensor")
        self._haplotypes = genotypes
        self._positions = positions
        self.chrom_start = chrom_start
        self.chrom_end = chrom_end
        self._sample_sets = sample_sets
        self.n_total_sites = n_total_sites
        self.samples = samples
        self._accessible_mask = None
        self._accessible_idx = None
        if accessible_mask is not None:
            self.accessible_mask = accessible_mask

    # ------------------------------------------------------------ data views
    @property
    def haplotypes(self):
        if self._accessible_idx is None:
            return self._haplotypes
        return self._haplotypes[:, self._take_idx()]

    @haplotypes.setter
    def haplotypes(self, value):
        self._haplotypes = value

    @property
    def positions(self):
        if self._accessible_idx is None:
            return self._positions
        return self._positions[self._take_idx()]

    @positions.setter
    def positions(self, value):
        self._positions = value

    def _take_idx(self):
        if self._device == "GPU":
            return dv.as_tensor(self._accessible_idx, dv.torch().int64, self._haplotypes.device)
        return self._accessible_idx

    @property
    def accessible_mask(self):
        return self._accessible_mask

    @accessible_mask.setter
    def accessible_mask(self, mask):
        """A boolean function of positions: keep variants where mask(pos) is True."""
        self._accessible_mask = mask
        if mask is None:
            self._accessible_idx = None
            return
        keep = np.asarray(mask(np.asarray(dv.to_numpy(self._positions)).astype(np.int64)), bool)
        self._accessible_idx = None if keep.all() else np.nonzero(keep)[0]

    @property
    def has_accessible_mask(self):
        return self._accessible_mask is not None

    def set_accessible_mask(self, mask_or_path, chrom=None):
        """Keep variants inside BED intervals (0-based half-open) of ``chrom``."""
        self.accessible_mask = read_bed_mask(mask_or_path, chrom) \
            if isinstance(mask_or_path, (str, bytes)) or hasattr(mask_or_path, "__fspath__") \
            else mask_or_path
        return self

    @property
    def device(self):
        return self._device

    @property
    def num_haplotypes(self):
        return int(self.haplotypes.shape[0])

    @property
    def num_variants(self):
        return int(self.haplotypes.shape[1])

    @property
    def sample_sets(self):
        if self._sample_sets is None:
            return {"all": list(range(self.haplotypes.shape[0]))}
        return self._sample_sets

    @sample_sets.setter
    def sample_sets(self, sample_sets):
        if not isinstance(sample_sets, dict):
            raise ValueError("sample_sets must be a dictionary")
        for v in sample_sets.values():
            if not isinstance(v, list):
                raise ValueError("values in sample_sets must be lists")
        self._sample_sets = sample_sets

    def transfer_to_gpu(self):
        if self._device == "CPU":
            t = dv.torch()
            dev = dv.device()
            self._haplotypes = t.as_tensor(np.ascontiguousarray(self._haplotypes, np.int8),
                                           device=dev)
            self._positions = t.as_tensor(np.asarray(self._positions, np.int64), device=dev)
            self._device = "GPU"

    def transfer_to_cpu(self):
        if self._device == "GPU":
            self._haplotypes = dv.to_numpy(self._haplotypes)
            self._positions = dv.to_numpy(self._positions)
            self._device = "CPU"

    # ------------------------------------------------------------ loading
    @classmethod
    def from_vcf(cls, path, region=None, samples=None, include_invariant=False,
                 accessible_bed=None):
        """Phased VCF -> haplotypes; allele-0 rows of every sample first, then
        allele-1 rows (haplotype_matrix.py:292-359).  ``region``
        ('chrom[:begin-end]', 1-based inclusive) keeps the records inside it
        and sets chrom_start / chrom_end to its bounds; ``samples`` keeps the
        named samples, in the file's order."""
        from .vcf import parse_region, read_vcf_gt

        d = read_vcf_gt(path, region=region, samples=samples)
        gt = d["gt"]
        n = gt.shape[1]
        hap = np.empty((2 * n, gt.shape[0]), dtype=np.int8)
        hap[:n] = gt[:, :, 0].T
        hap[n:] = gt[:, :, 1].T
        pos = d["pos"]
        chrom_start, chrom_end = int(pos[0]), int(pos[-1])
        chrom = d["chrom"][0]
        if region is not None:
            chrom, begin, end = parse_region(region)
            chrom_start = begin if begin is not None else chrom_start
            chrom_end = end if end is not None else chrom_end
        hm = cls(hap, pos, chrom_start, chrom_end,
                 n_total_sites=gt.shape[0] if include_invariant else None,
                 samples=list(d["samples"]))
        if accessible_bed is not None:
            hm.set_accessible_mask(accessible_bed, chrom=chrom)
        return hm

    def load_pop_file(self, pop_file, pops=None):
        """sample -> pop file; sample i maps to rows i and i + n_samples."""
        from .vcf import read_pop_file

        if self.samples is None:
            raise ValueError("No sample names stored. Use from_vcf() to load data.")
        pm = read_pop_file(pop_file)
        if pops is None:
            pops = sorted(set(pm.values()))
        n = len(self.samples)
        sets = {p: [] for p in pops}
        for i, name in enumerate(self.samples):
            p = pm.get(name)
            if p in sets:
                sets[p].append(i)
                sets[p].append(i + n)
        self.sample_sets = sets

    def apply_biallelic_filter(self):
        """Keep variants with a 0 and a 1 and no allele >= 2 across all
        haplotypes (haplotype_matrix.py:724-789)."""
        from . import filters

        h = self.haplotypes
        idx = filters.keep_index(filters.biallelic_keep(h, _lib.FILTER_HAP_BIALLELIC))
        if idx.numel() == 0:
            # the reference cannot represent an empty matrix either; callers get
            # a placeholder whose binned statistics are all zero
            return _EmptyHaplotypes(int(h.shape[0]), self._sample_sets)
        hk = filters.take_columns(h, idx)
        pos = filters.take_columns(self.positions, idx)
        p_np = dv.to_numpy(pos[[0, -1]]) if not isinstance(pos, np.ndarray) else pos[[0, -1]]
        return HaplotypeMatrix(hk, pos, int(p_np[0]), int(p_np[1]),
                               sample_sets=self._sample_sets, n_total_sites=self.n_total_sites,
                               samples=self.samples, accessible_mask=self.accessible_mask)

    # ------------------------------------------------------------ pair tallies
    def _pop_rows(self, pop):
        if pop is None:
            return list(range(self.num_haplotypes))
        if self._sample_sets is None:
            raise ValueError("sample_sets must be defined to use pop parameter")
        if pop not in self._sample_sets:
            raise KeyError(f"Population key {pop} must exist in sample_sets")
        return [int(i) for i in self._sample_sets[pop]]

    def _upper_counts(self, pan, p):
        """int32 [V(V-1)/2, 4] (n11, n10, n01, n00) of population p over all
        pairs i < j in np.triu_indices(V, 1) order, from MFMA counts of
        32-row stripes of the upper triangle (pgld_count_block)."""
        t = dv.torch()
        V = pan.V
        parts = []
        stripe = max(32, min(4096, ((1 << 24) // max(V, 1)) // 32 * 32))
        for i0 in range(0, max(V - 1, 0), stripe):
            ni = min(stripe, V - i0)
            blk = pan.count_block(p, i0, ni, i0, V - i0)          # [4, ni, V - i0]
            upper = t.ones((ni, V - i0), dtype=t.bool, device=pan.device).triu(1)
            parts.append(blk[:, upper].t())
        if not parts:
            return t.zeros((0, 4), dtype=t.int32, device=pan.device)
        return t.cat(parts).contiguous()

    def tally_gpu_haplotypes(self, pop=None):
        """(counts int32 [V(V-1)/2, 4] = n11, n10, n01, n00 for every pair of
        variants i < j, n_valid int32 [V(V-1)/2] or None without missing
        calls) over the haplotypes of ``pop`` (all if None); pairs in
        triu_indices(V, 1) order (haplotype_matrix.py:1320-1375).  Counts
        come from the MFMA counting path and are exact; the reference's
        no-missing branch multiplies int8 matrices, whose n11 wraps above 127
        carriers, and that wrap is not reproduced."""
        from .engine import LDPanels

        if self._device == "CPU":
            self.transfer_to_gpu()
        rows = self._pop_rows(pop)
        pan = LDPanels(self.haplotypes, [rows], "hap")
        counts = self._upper_counts(pan, 0)
        n_valid = counts.sum(dim=1, dtype=dv.torch().int32) if pan.has_missing[0] else None
        return counts, n_valid

    def _tally_gpu_haplotypes_with_missing_impl(self, X):
        """(counts, n_valid) of all variant pairs over the rows of X, counting
        only haplotypes valid at both variants (haplotype_matrix.py:1377-1433)."""
        from .engine import LDPanels

        t = dv.torch()
        X = dv.as_tensor(X, t.int8)
        pan = LDPanels(X, [list(range(X.shape[0]))], "hap")
        counts = self._upper_counts(pan, 0)
        return counts, counts.sum(dim=1, dtype=t.int32)

    def tally_gpu_haplotypes_two_pops(self, pop1, pop2):
        """(counts int32 [V(V-1)/2, 8] = (n11, n10, n01, n00) of pop1 then of
        pop2, n_valid1, n_valid2), the n_valid arrays None when the matrix has
        no missing call (haplotype_matrix.py:1511-1589)."""
        from .engine import LDPanels

        if self._device == "CPU":
            self.transfer_to_gpu()
        if self._sample_sets is None:
            raise ValueError("sample_sets must be defined to use this function")
        if pop1 not in self._sample_sets or pop2 not in self._sample_sets:
            raise KeyError(f"Population keys {pop1} and {pop2} must exist in sample_sets")
        t = dv.torch()
        h = self.haplotypes
        pan = LDPanels(h, [self._pop_rows(pop1), self._pop_rows(pop2)], "hap")
        c1, c2 = self._upper_counts(pan, 0), self._upper_counts(pan, 1)
        counts = t.cat([c1, c2], dim=1)
        whole = LDPanels(h, [list(range(h.shape[0]))], "hap", pack=False)
        if not whole.has_missing[0]:
            return counts, None, None
        return counts, c1.sum(dim=1, dtype=t.int32), c2.sum(dim=1, dtype=t.int32)

    def tally_gpu_haplotypes_two_pops_with_missing(self, pop1, pop2):
        """As tally_gpu_haplotypes_two_pops, always returning the valid-pair
        counts (haplotype_matrix.py:1435-1509)."""
        counts, n1, n2 = self.tally_gpu_haplotypes_two_pops(pop1, pop2)
        if n1 is None:
            t = dv.torch()
            n1 = counts[:, 0:4].sum(dim=1, dtype=t.int32)
            n2 = counts[:, 4:8].sum(dim=1, dtype=t.int32)
        return counts, n1, n2

    # ------------------------------------------------------------ binned LD
    def _binned(self, bp_bins, pops, raw):
        from .engine import LDPanels

        if self._device == "CPU":
            self.transfer_to_gpu()
        bins = np.asarray(bp_bins, np.float64)
        rows = [list(p) for p in pops]
        pan = LDPanels(self.haplotypes, rows, "hap")
        res = pan.binned(dv.to_numpy(self.positions), bins, float(bins[-1]))
        out = []
        for b in range(len(bins) - 1):
            c = int(res.counts[b])
            if raw:
                vals = res.sums[b]
            else:
                vals = res.sums[b] / c if c > 0 else np.zeros_like(res.sums[b])
            out.append(((float(bins[b]), float(bins[b + 1])), vals))
        return out

    def pairwise_r2(self, estimator="r2"):
        """Pairwise r^2 of all variant pairs (haplotype_matrix.py:1138-1170)."""
        from .ld_tiles import pairwise_r2

        return pairwise_r2(self, estimator)

    def windowed_r_squared(self, bp_bins, percentile=50, pop=None, estimator="r2"):
        """Percentiles of r^2 in distance bins (haplotype_matrix.py:1235-1317)."""
        from .ld_tiles import windowed_r_squared

        return windowed_r_squared(self, bp_bins, percentile, pop, estimator)

    def compute_ld_statistics_gpu_single_pop(self, bp_bins, raw=False, ac_filter=True,
                                             chunk_size="auto"):
        """{(lo, hi): (DD, Dz, pi2)} over all haplotypes; sums if ``raw``,
        else means per pair, zeros for empty bins (haplotype_matrix.py:1592-1716)."""
        if ac_filter:
            f = self.apply_biallelic_filter()
            if isinstance(f, _EmptyHaplotypes):
                return {(float(a), float(b)): (0.0, 0.0, 0.0) for a, b in zip(bp_bins[:-1], bp_bins[1:])}
            return f.compute_ld_statistics_gpu_single_pop(bp_bins, raw, False, chunk_size)
        rows = list(range(self.num_haplotypes))
        return {k: tuple(float(x) for x in v) for k, v in self._binned(bp_bins, [rows], raw)}

    def compute_ld_statistics_gpu_two_pops(self, bp_bins, pop1, pop2, raw=False, ac_filter=True,
                                           chunk_size="auto"):
        """{(lo, hi): OrderedDict of the 15 two-population statistics}
        (haplotype_matrix.py:1719-1866)."""
        from .ld_pipeline import ld_names

        if ac_filter:
            f = self.apply_biallelic_filter()
            if isinstance(f, _EmptyHaplotypes):
                return {(float(a), float(b)): OrderedDict((n, 0.0) for n in ld_names(2))
                        for a, b in zip(bp_bins[:-1], bp_bins[1:])}
            return f.compute_ld_statistics_gpu_two_pops(bp_bins, pop1, pop2, raw, False,
                                                        chunk_size)
        pops = [self._sample_sets[pop1], self._sample_sets[pop2]]
        names = ld_names(2)
        return {k: OrderedDict(zip(names, (float(x) for x in v)))
                for k, v in self._binned(bp_bins, pops, raw)}


class _EmptyHaplotypes:
    """Result of a filter that removed every variant."""

    def __init__(self, n_hap, sample_sets):
        self.num_haplotypes = n_hap
        self._sample_sets = sample_sets
        self.num_variants = 0
