"""GenotypeMatrix container (drop-in for the LD-relevant part of
pg_gpu/genotype_matrix.py): unphased diploid genotypes (n_individuals,
n_variants) int8 with 0/1/2 alt-allele counts and -1 for missing."""
from __future__ import annotations

import numpy as np

from . import _device as dv
from .accessible import read_bed_mask


class GenotypeMatrix:
    def __init__(self, genotypes, positions, chrom_start=None, chrom_end=None, sample_sets=None,
                 n_total_sites=None, samples=None, accessible_mask=None):
        if genotypes is None or (hasattr(genotypes, "size") and not callable(genotypes.size)
                                 and genotypes.size == 0) or \
                (hasattr(genotypes, "numel") and genotypes.numel() == 0):
            raise ValueError("genotypes cannot be empty")
        if positions is None or len(positions) == 0:
            raise ValueError("positions cannot be empty")
        t = dv.torch()
        if isinstance(genotypes, t.Tensor) and genotypes.is_cuda:
            self._device = "GPU"
            if not (isinstance(positions, t.Tensor) and positions.is_cuda):
                positions = dv.as_tensor(positions, t.int64, genotypes.device)
        else:
            self._device = "CPU"
            genotypes = dv.to_numpy(genotypes)
            positions = dv.to_numpy(positions)
        self._genotypes = genotypes
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

    def _take_idx(self):
        if self._device == "GPU":
            return dv.as_tensor(self._accessible_idx, dv.torch().int64, self._genotypes.device)
        return self._accessible_idx

    @property
    def genotypes(self):
        if self._accessible_idx is None:
            return self._genotypes
        return self._genotypes[:, self._take_idx()]

    @genotypes.setter
    def genotypes(self, value):
        self._genotypes = value

    @property
    def positions(self):
        if self._accessible_idx is None:
            return self._positions
        return self._positions[self._take_idx()]

    @positions.setter
    def positions(self, value):
        self._positions = value

    @property
    def accessible_mask(self):
        return self._accessible_mask

    @accessible_mask.setter
    def accessible_mask(self, mask):
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
        self.accessible_mask = read_bed_mask(mask_or_path, chrom) \
            if isinstance(mask_or_path, (str, bytes)) or hasattr(mask_or_path, "__fspath__") \
            else mask_or_path
        return self

    @property
    def device(self):
        return self._device

    @property
    def num_individuals(self):
        return int(self.genotypes.shape[0])

    @property
    def num_variants(self):
        return int(self.genotypes.shape[1])

    @property
    def sample_sets(self):
        if self._sample_sets is None:
            return {"all": list(range(self.genotypes.shape[0]))}
        return self._sample_sets

    @sample_sets.setter
    def sample_sets(self, sample_sets):
        if not isinstance(sample_sets, dict):
            raise ValueError("sample_sets must be a dictionary")
        self._sample_sets = sample_sets

    def transfer_to_gpu(self):
        if self._device == "CPU":
            t = dv.torch()
            dev = dv.device()
            self._genotypes = t.as_tensor(np.ascontiguousarray(self._genotypes, np.int8), device=dev)
            self._positions = t.as_tensor(np.asarray(self._positions, np.int64), device=dev)
            self._device = "GPU"

    def transfer_to_cpu(self):
        if self._device == "GPU":
            self._genotypes = dv.to_numpy(self._genotypes)
            self._positions = dv.to_numpy(self._positions)
            self._device = "CPU"

    @classmethod
    def from_haplotype_matrix(cls, hap_matrix):
        """Pair haplotype rows (0,1), (2,3), ...; a genotype is missing if
        either haplotype is (genotype_matrix.py:247-305).  The conversion runs
        on the GPU (pgld_hap_to_geno); the result stays on the input's device."""
        from . import filters

        g = filters.hap_to_geno(hap_matrix.haplotypes)
        sets = None
        if hap_matrix._sample_sets is not None:
            sets = {k: sorted(set(int(i) // 2 for i in v))
                    for k, v in hap_matrix._sample_sets.items()}
        return cls(g, hap_matrix.positions, hap_matrix.chrom_start, hap_matrix.chrom_end,
                   sample_sets=sets, n_total_sites=hap_matrix.n_total_sites,
                   accessible_mask=hap_matrix.accessible_mask)

    def to_haplotype_matrix(self):
        """Each individual becomes two haplotypes: rows 2i = [g >= 1] and
        2i+1 = [g >= 2], so hets become (1, 0); missing stays -1 in both
        (genotype_matrix.py:307-340; on the GPU by pgld_geno_to_hap)."""
        from . import filters
        from .haplotype_matrix import HaplotypeMatrix

        hap = filters.geno_to_hap(self.genotypes)
        return HaplotypeMatrix(hap, self.positions, self.chrom_start, self.chrom_end,
                               n_total_sites=self.n_total_sites,
                               accessible_mask=self.accessible_mask)

    @classmethod
    def from_vcf(cls, path, include_invariant=False, accessible_bed=None):
        """Keep sites whose alleles are only 0/1 with both present
        (is_biallelic_01); genotype = allele sum, -1 if either allele is
        missing (genotype_matrix.py:342-390)."""
        from .vcf import read_vcf_gt

        d = read_vcf_gt(path)
        gt = d["gt"]
        called = gt >= 0
        n0 = np.sum(gt == 0, axis=(1, 2))
        n1 = np.sum(gt == 1, axis=(1, 2))
        nother = np.sum(gt >= 2, axis=(1, 2))
        keep = (n0 > 0) & (n1 > 0) & (nother == 0)
        gt = gt[keep]
        pos = d["pos"][keep]
        g = np.where(called[keep].all(axis=2), gt.sum(axis=2), -1).astype(np.int8).T
        gm = cls(g, pos, int(pos[0]), int(pos[-1]),
                 n_total_sites=g.shape[1] if include_invariant else None,
                 samples=list(d["samples"]))
        if accessible_bed is not None:
            gm.set_accessible_mask(accessible_bed, chrom=d["chrom"][0])
        return gm

    def load_pop_file(self, pop_file, pops=None):
        from .vcf import read_pop_file

        if self.samples is None:
            raise ValueError("No sample names stored. Use from_vcf() to load data.")
        pm = read_pop_file(pop_file)
        if pops is None:
            pops = sorted(set(pm.values()))
        sets = {p: [] for p in pops}
        for i, name in enumerate(self.samples):
            p = pm.get(name)
            if p in sets:
                sets[p].append(i)
        self.sample_sets = sets

    def apply_biallelic_filter(self):
        """Keep sites with 0 < alt count < 2 * n_valid and n_valid >= 2
        (genotype_matrix.py:506-539); the site test runs on the GPU
        (pgld_biallelic_mask)."""
        from . import _lib, filters

        g = self.genotypes
        idx = filters.keep_index(filters.biallelic_keep(g, _lib.FILTER_GENO_BIALLELIC))
        return GenotypeMatrix(filters.take_columns(g, idx),
                              filters.take_columns(self.positions, idx), self.chrom_start,
                              self.chrom_end, sample_sets=self._sample_sets,
                              n_total_sites=self.n_total_sites, samples=self.samples,
                              accessible_mask=self.accessible_mask)
