"""pgld_binned_fused_noalloc against pgld_binned_fused through the C ABI:
bit-identical fixed-point limbs on two consecutive calls, the second on a
smaller input (the shared scheduler words and the output must be cleared by
the call itself)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def run_pair(cuda, scale, seed, dirty):
    import torch

    from pg_gpu_amd import _lib, synth
    from pg_gpu_amd.engine import LDPanels

    w = synth.make("c1", scale=scale, seed=seed)
    pan = LDPanels(torch.as_tensor(w.mat, device=cuda), w.pop_rows, "hap")
    plan = pan.plan(w.positions, w.bins, w.max_dist)
    assert len(plan.tiles) > 0
    ref = plan.run("fused").finish()
    width = pan.n_stats
    out = torch.full((plan.nb, width + 1, 3), dirty, dtype=torch.int64, device=cuda)
    a = plan.args(out)
    a.d_sched = None
    rc = _lib.lib().pgld_binned_fused_noalloc(ctypes.byref(pan.desc), ctypes.byref(a),
                                              _lib.stream_handle())
    _lib.check(rc)
    torch.cuda.synchronize()
    got_raw = out.cpu().numpy()
    exp = torch.zeros_like(out)
    b = plan.args(exp)
    sched = torch.zeros(_lib.SCHED_WORDS, dtype=torch.int32, device=cuda)
    b.d_sched = _lib.ptr(sched).value
    _lib.check(_lib.lib().pgld_binned_fused(ctypes.byref(pan.desc), ctypes.byref(b),
                                            _lib.stream_handle()))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(got_raw, exp.cpu().numpy())
    from pg_gpu_amd.engine import decode_limbs

    sums, counts = decode_limbs(got_raw, width)
    np.testing.assert_array_equal(counts, ref.counts)
    np.testing.assert_array_equal(sums, ref.sums)
    return len(plan.tiles)


def test_noalloc_matches_allocating_entry_point_twice(cuda):
    n1 = run_pair(cuda, 0.01, 7, 12345)
    n2 = run_pair(cuda, 0.003, 8, -1)      # smaller input, output pre-filled with garbage
    assert n2 < n1


def test_noalloc_rejects_a_scheduler_buffer(cuda):
    import torch

    from pg_gpu_amd import _lib, synth
    from pg_gpu_amd.engine import LDPanels

    w = synth.make("c1", scale=0.002, seed=9)
    pan = LDPanels(torch.as_tensor(w.mat, device=cuda), w.pop_rows, "hap")
    plan = pan.plan(w.positions, w.bins, w.max_dist)
    out = torch.zeros((plan.nb, pan.n_stats + 1, 3), dtype=torch.int64, device=cuda)
    a = plan.args(out)
    sched = torch.zeros(_lib.SCHED_WORDS, dtype=torch.int32, device=cuda)
    a.d_sched = _lib.ptr(sched).value
    rc = _lib.lib().pgld_binned_fused_noalloc(ctypes.byref(pan.desc), ctypes.byref(a),
                                              _lib.stream_handle())
    assert rc == _lib.PGLD_EINVAL
    with pytest.raises(ValueError, match="d_sched must be NULL"):
        _lib.check(rc)
