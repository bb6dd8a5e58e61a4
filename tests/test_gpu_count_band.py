"""The counting stage alone (pgld_count_band) against the oracle's
closed-form checksum of the band's count fields: at small sizes for one to
four populations with and without missing calls, and at the full size of the
two-population workload (2 x 500 haplotypes, 200,000 sites, pairs within
200 kb)."""
import ctypes

import numpy as np
import pytest

from oracle import ld_oracle as o

pytestmark = pytest.mark.gpu


def count_band(pan, tiles, dynamic=True):
    import torch

    from pg_gpu_amd import _lib

    out = torch.zeros((1, 1, 3), dtype=torch.int64, device=pan.device)
    plan = pan.plan(np.arange(pan.V, dtype=np.float64), [0.0, 1.0], 1.0, tiles=tiles)
    a = plan.args(out)
    sched = None
    if dynamic:
        sched = torch.zeros(_lib.SCHED_WORDS, dtype=torch.int32, device=pan.device)
        a.d_sched = _lib.ptr(sched).value
    sums = torch.zeros((pan.P, 4), dtype=torch.int64, device=pan.device)
    _lib.check(_lib.lib().pgld_count_band(ctypes.byref(pan.desc), ctypes.byref(a),
                                          _lib.ptr(sums), _lib.stream_handle()))
    return sums.cpu().numpy()


@pytest.mark.parametrize("P,missing", [(1, 0.0), (1, 0.05), (2, 0.0), (2, 0.05), (3, 0.03),
                                       (4, 0.03), (4, 0.0)])
def test_count_band_checksum_small(cuda, P, missing):
    import torch

    from pg_gpu_amd.engine import LDPanels, band_tiles

    rng = np.random.default_rng(10 * P + int(100 * missing))
    sizes = [40, 120, 300, 64][:P]
    V = 1000 + 17 * P
    hap = (rng.random((sum(sizes), V)) < rng.beta(0.5, 0.5, V)).astype(np.int8)
    if missing:
        hap[rng.random(hap.shape) < missing] = -1
    offs = np.cumsum([0] + sizes)
    pops = [list(range(offs[k], offs[k + 1])) for k in range(P)]
    pos = np.sort(rng.choice(np.arange(1, 40 * V), V, replace=False)).astype(np.float64)
    pan = LDPanels(torch.as_tensor(hap, device=cuda), pops, "hap")
    tiles = band_tiles(pos, 3000.0)
    want = o.band_count_checksum(hap, pops, tiles, marginals=P > 1)
    for dyn in (True, False):
        np.testing.assert_array_equal(count_band(pan, tiles, dyn), want)


def test_count_band_checksum_full_two_population_workload(cuda):
    import torch

    from pg_gpu_amd import synth
    from pg_gpu_amd.engine import LDPanels, band_tiles

    w = synth.make("c1")
    pan = LDPanels(torch.as_tensor(w.mat, device=cuda), w.pop_rows, "hap")
    tiles = band_tiles(w.positions, w.max_dist)
    got = count_band(pan, tiles)
    want = o.band_count_checksum(w.mat, w.pop_rows, tiles, marginals=True)
    np.testing.assert_array_equal(got, want)
