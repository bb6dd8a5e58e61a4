"""End-to-end oracle check against the moments golden fixture (CPU only).

tests/golden/moments_im_2pop.json holds moments' output for the reference's
IM example VCF (2 pops, bp_bins = logspace(2, 5, 4)); see
tests/golden/make_moments_fixture.py for how it was decoded.
"""
import json
from pathlib import Path

import numpy as np

from oracle import ld_oracle as o
from pg_gpu_amd.vcf import read_pop_file, read_vcf_gt

GOLD = Path(__file__).resolve().parent / "golden"


def load_im():
    d = read_vcf_gt(GOLD / "im-parsing-example.vcf.gz")
    gt = d["gt"]
    n = len(d["samples"])
    hap = np.concatenate([gt[:, :, 0].T, gt[:, :, 1].T], axis=0)
    keep = o.biallelic_filter_hap(hap)
    pm = read_pop_file(GOLD / "im_pop.txt")
    pops = []
    for name in ("deme0", "deme1"):
        rows = []
        for i, s in enumerate(d["samples"]):
            if pm.get(s) == name:
                rows += [i, i + n]
        pops.append(sorted(rows))
    return hap[:, keep], d["pos"][keep], pops


def test_golden_moments_two_pops():
    fx = json.loads((GOLD / "moments_im_2pop.json").read_text())
    hap, pos, pops = load_im()
    sums, counts, _ = o.binned_sums_hap(hap, pos, pops, fx["bp_bins"])
    ref = np.array(fx["sums"])
    assert np.max(np.abs(sums - ref) / np.abs(ref)) < 1e-12
    het = o.heterozygosity_hap(hap, pops)
    np.testing.assert_allclose([het[k] for k in fx["het_names"]], fx["het"], rtol=1e-14)
    assert fx["ld_names"] == o.ld_names(2)
