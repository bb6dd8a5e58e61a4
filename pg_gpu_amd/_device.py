"""Tensor plumbing shared by the API modules: moving array-likes to the GPU
in the dtype a libpgld entry point expects, and device selection."""
from __future__ import annotations

import numpy as np


def torch():
    import torch as _t

    return _t


def device(hint=None):
    t = torch()
    if hint is not None and getattr(hint, "is_cuda", False):
        return hint.device
    if not t.cuda.is_available():
        raise RuntimeError("pg_gpu_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return t.device("cuda", t.cuda.current_device())


def as_tensor(x, dtype, dev=None):
    """Contiguous CUDA tensor of ``dtype`` from a tensor / ndarray / list."""
    t = torch()
    if isinstance(x, t.Tensor):
        dev = dev or (x.device if x.is_cuda else device())
        return x.to(device=dev, dtype=dtype).contiguous()
    return t.as_tensor(np.asarray(x), dtype=dtype, device=dev or device()).contiguous()


def to_numpy(x):
    t = torch()
    if isinstance(x, t.Tensor):
        return x.detach().cpu().numpy()
    return np.asarray(x)


def is_gpu_array(x):
    t = torch()
    return isinstance(x, t.Tensor) and x.is_cuda
