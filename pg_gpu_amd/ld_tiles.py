"""Dense LD tile consumers: Kelly's ZnS, Kim & Nielsen's omega and pairwise
r^2 matrices (drop-in for pg_gpu/ld_statistics.py:249-1021 and
HaplotypeMatrix.pairwise_r2, haplotype_matrix.py:1092-1170).

Every per-pair value comes from pgld_pair_block (FP4 MFMA counts of 32 x 32
site blocks, one value per pair); the reductions over the pairs (the upper
triangle sums of ZnS, the 2-D prefix sums of omega) run as torch operations
on the device.  Results are float64.
"""
from __future__ import annotations

import ctypes

from . import _device as dv
from . import _lib
from .engine import LDPanels

_BLOCK = 4096  # sites per pgld_pair_block call side (the output is [2][B][B] float64)


def _pair_matrix(data, kind, mode):
    """[m, m] float64 values and 0/1 flags of pgld_pair_block over every pair
    of the columns of ``data`` (int8 CUDA tensor [rows, m]); diagonal zero."""
    t = dv.torch()
    m = data.shape[1]
    pan = LDPanels(data.contiguous(), [list(range(data.shape[0]))], mode)
    val = t.zeros((m, m), dtype=t.float64, device=data.device)
    flag = t.zeros((m, m), dtype=t.float64, device=data.device)
    L = _lib.lib()
    for i0 in range(0, m, _BLOCK):
        ni = min(_BLOCK, m - i0)
        for j0 in range(i0, m, _BLOCK):
            nj = min(_BLOCK, m - j0)
            out = t.empty((2, ni, nj), dtype=t.float64, device=data.device)
            _lib.check(L.pgld_pair_block(ctypes.byref(pan.desc), _lib.ptr(pan.panels),
                                         _lib.ptr(pan.site_stats), 0, kind, i0, ni, j0, nj,
                                         _lib.ptr(out), _lib.stream_handle()))
            val[i0:i0 + ni, j0:j0 + nj] = out[0]
            flag[i0:i0 + ni, j0:j0 + nj] = out[1]
            if j0 != i0:
                val[j0:j0 + nj, i0:i0 + ni] = out[0].T
                flag[j0:j0 + nj, i0:i0 + ni] = out[1].T
    val.fill_diagonal_(0.0)
    flag.fill_diagonal_(0.0)
    return val, flag


def windowed_r_squared(hap_matrix, bp_bins, percentile=50, pop=None, estimator="r2"):
    """Percentiles of r^2 in genomic-distance bins over every variant pair
    (haplotype_matrix.py:1235-1317).  'r2': r^2 from each pair's counts over
    the samples valid at both sites (of ``pop`` if given); 'rogers_huff':
    r^2 of the dosages of adjacent haplotype pairs.  Returns (result
    [n_bins] or [n_bins, n_percentiles], pair counts per bin); pairs with
    undefined r^2 are left out, as in the reference."""
    import numpy as np

    t = dv.torch()
    h = _device_data(hap_matrix, True)
    if estimator == "rogers_huff":
        if pop is not None:
            raise NotImplementedError(
                "windowed_r_squared(estimator='rogers_huff', pop=...) is not yet implemented; "
                "pass the full matrix or subset to the desired haplotypes first.")
        if bool((h < 0).any()):
            raise ValueError("rogers_huff_r does not support missing values (-1)")
        if h.shape[0] % 2:
            raise ValueError("Need even number of haplotypes for diploid conversion")
        g = (h[0::2].to(t.int16) + h[1::2]).to(t.int8)
        r2, flag = _pair_matrix(g, _lib.PAIR_R2_DOSAGE, "geno")
        r2 = t.where(flag > 0, r2, t.full_like(r2, float("nan")))  # monomorphic: r undefined
    elif estimator == "r2":
        if pop is not None:
            rows = hap_matrix.sample_sets[pop] if isinstance(pop, str) else pop
            h = h[dv.as_tensor(list(rows), t.int64, h.device)]
        r2, _ = _pair_matrix(h, _lib.PAIR_R2, "hap")
    else:
        raise ValueError(f"Unknown estimator: {estimator!r} (expected 'r2' or 'rogers_huff')")
    m = r2.shape[0]
    ii, jj = t.triu_indices(m, m, offset=1, device=r2.device)
    vals = r2[ii, jj]
    pos = dv.as_tensor(hap_matrix.positions, t.float64, r2.device)
    dist = pos[jj] - pos[ii]
    edges = t.as_tensor(np.asarray(bp_bins, np.float64), device=r2.device)
    b = t.bucketize(dist, edges, right=True) - 1          # np.digitize(d, bins) - 1
    nb = len(bp_bins) - 1
    ok = (b >= 0) & (b < nb) & ~vals.isnan()
    v_cpu = vals[ok].cpu().numpy()
    b_cpu = b[ok].cpu().numpy()
    pct = np.atleast_1d(percentile)
    result = np.full((nb, len(pct)), np.nan)
    counts = np.zeros(nb, dtype=int)
    for k in range(nb):
        sel = v_cpu[b_cpu == k]
        counts[k] = len(sel)
        if len(sel):
            result[k] = np.percentile(sel, pct)
    if result.shape[1] == 1:
        result = result[:, 0]
    return result, counts


def _is_hap(obj):
    from .haplotype_matrix import HaplotypeMatrix

    return isinstance(obj, HaplotypeMatrix)


def _is_geno(obj):
    from .genotype_matrix import GenotypeMatrix

    return isinstance(obj, GenotypeMatrix)


def _device_data(mat, hap):
    t = dv.torch()
    if getattr(mat, "device", "GPU") == "CPU":
        mat.transfer_to_gpu()
    d = mat.haplotypes if hap else mat.genotypes
    return dv.as_tensor(d, t.int8)


def _segregating(h, missing_data):
    """Columns of the haplotype tensor kept by _prepare_segregating
    (ld_statistics.py:249-278): no missing call with 'exclude', and
    0 < derived count < valid count."""
    t = dv.torch()
    if missing_data == "exclude":
        h = h[:, (h < 0).sum(dim=0) == 0]
    valid = (h >= 0).sum(dim=0)
    dac = h.clamp(min=0).to(t.int32).sum(dim=0)
    seg = (dac > 0) & (dac < valid)
    return h[:, seg].contiguous()


def resolve_ld_estimator(estimator, is_hap_matrix):
    """'auto' -> 'sigma_d2' for haplotypes, 'rogers_huff' otherwise
    (ld_statistics.py:344-363)."""
    if estimator == "auto":
        return "sigma_d2" if is_hap_matrix else "rogers_huff"
    if estimator not in ("r2", "sigma_d2", "rogers_huff"):
        raise ValueError(f"Unknown estimator: {estimator!r} "
                         "(expected one of 'auto', 'r2', 'sigma_d2', 'rogers_huff')")
    return estimator


def pairwise_r2(hap_matrix, estimator="r2"):
    """Pairwise r^2 of all variant pairs, diagonal 0 (haplotype_matrix.py:1138-1170):
    'r2' the frequency r^2; 'rogers_huff' r^2 of the diploid dosages of
    adjacent haplotype pairs (missing calls raise)."""
    t = dv.torch()
    h = _device_data(hap_matrix, True)
    if estimator == "rogers_huff":
        if bool((h < 0).any()):
            raise ValueError("rogers_huff_r does not support missing values (-1)")
        if h.shape[0] % 2:
            raise ValueError("Need even number of haplotypes for diploid conversion")
        g = (h[0::2].to(t.int16) + h[1::2]).to(t.int8)
        return _pair_matrix(g, _lib.PAIR_R2_DOSAGE, "geno")[0]
    if estimator != "r2":
        raise ValueError(f"Unknown estimator: {estimator!r} (expected 'r2' or 'rogers_huff')")
    return _pair_matrix(h, _lib.PAIR_R2_NAIVE, "hap")[0]


def r2_matrix_diploid(genotype_matrix):
    """r^2 of mean-centred dosages with missing calls zeroed (ld_statistics.py:973-1021)."""
    g = _device_data(genotype_matrix, False) if _is_geno(genotype_matrix) else \
        dv.as_tensor(genotype_matrix, dv.torch().int8)
    return _pair_matrix(g, _lib.PAIR_R2_DOSAGE, "geno")[0]


def _r2_for(obj, missing_data, estimator):
    """The r^2 (or sigma_D^2) matrix zns / omega work on (ld_statistics.py:926-970)."""
    t = dv.torch()
    if _is_hap(obj):
        h = _segregating(_device_data(obj, True), missing_data)
        if h.shape[1] < 2:
            return t.zeros((h.shape[1], h.shape[1]), dtype=t.float64, device=h.device), None
        if estimator == "sigma_d2":
            return _pair_matrix(h, _lib.PAIR_SIGMA_D2, "hap")
        return _pair_matrix(h, _lib.PAIR_R2_NAIVE, "hap")[0], None
    if _is_geno(obj):
        g = _device_data(obj, False)
        if missing_data == "exclude":
            g = g[:, (g < 0).sum(dim=0) == 0].contiguous()
        if g.shape[1] < 1:
            return t.zeros((0, 0), dtype=t.float64, device=g.device), None
        return _pair_matrix(g, _lib.PAIR_R2_DOSAGE, "geno")[0], None
    return dv.as_tensor(obj, t.float64), None


def zns(r2_matrix_or_matrix, missing_data="include", estimator="auto"):
    """Kelly's ZnS: mean pairwise r^2 (or mean sigma_D^2 with 'sigma_d2', the
    default for a HaplotypeMatrix) over all pairs of segregating sites
    (ld_statistics.py:534-731)."""
    is_hm = _is_hap(r2_matrix_or_matrix)
    estimator = resolve_ld_estimator(estimator, is_hm)
    if estimator == "sigma_d2" and not is_hm:
        raise ValueError("estimator='sigma_d2' requires a HaplotypeMatrix, "
                         "not a pre-computed r² array")
    # the sigma_d2 ZnS keeps sites with missing calls whatever missing_data
    # says (ld_statistics.py:712-718 maps it to the internal "project" mode)
    md = "include" if estimator == "sigma_d2" else missing_data
    r2, valid = _r2_for(r2_matrix_or_matrix, md, estimator)
    m = r2.shape[0]
    if m < 2:
        return 0.0
    if valid is not None:  # sigma_d2: mean over pairs with n >= 4
        n_pairs = float(valid.sum().item())
        return float(r2.sum().item()) / n_pairs if n_pairs > 0 else 0.0
    total = r2.sum() - r2.diagonal().sum()
    return float(total.item()) / (m * (m - 1))


def omega(r2_matrix_or_matrix, missing_data="include", estimator="auto"):
    """Kim & Nielsen's omega: max over split points l = 3..m-2 of the mean
    within-block r^2 over the mean cross-block r^2, upper-triangle pairs
    (ld_statistics.py:761-862)."""
    t = dv.torch()
    is_hm = _is_hap(r2_matrix_or_matrix)
    estimator = resolve_ld_estimator(estimator, is_hm)
    if estimator == "sigma_d2" and not is_hm:
        raise ValueError("estimator='sigma_d2' requires a HaplotypeMatrix")
    r2, _ = _r2_for(r2_matrix_or_matrix, missing_data, estimator)
    m = r2.shape[0]
    if m < 5:
        return 0.0
    S = t.cumsum(t.cumsum(t.triu(r2, diagonal=1), dim=0), dim=1)
    l = t.arange(3, m - 1, device=r2.device)
    left = S[l - 1, l - 1]
    total = S[m - 1, m - 1]
    cross = S[l - 1, m - 1] - left
    right = total - left - cross
    lf = l.to(t.float64)
    n_within = lf * (lf - 1) / 2 + (m - lf) * (m - lf - 1) / 2
    n_cross = lf * (m - lf)
    valid = (n_within > 0) & (n_cross > 0) & (cross > 0)
    mean_within = t.where(n_within > 0, (left + right) / n_within, t.zeros_like(left))
    mean_cross = t.where(n_cross > 0, cross / n_cross, t.ones_like(cross))
    om = t.where(valid, mean_within / mean_cross, t.zeros_like(left))
    return float(om.max().item())
