"""Host/device plumbing for the drop-in API.

The reference API takes and returns numpy arrays.  Every public function
here accepts numpy arrays (copied to the GPU, result copied back) or torch
tensors already on the GPU (no copy; the result stays on the GPU).
"""
from __future__ import annotations

import numpy as np

from . import _native


def torch():
    return _native.torch_mod()


def is_device_tensor(x) -> bool:
    try:
        import torch as _t
    except ImportError:  # pragma: no cover
        return False
    return isinstance(x, _t.Tensor) and x.is_cuda


def to_device(x, dtype=np.float64):
    """Contiguous GPU tensor of `dtype` holding x (numpy or torch)."""
    t = torch()
    tdt = {np.float64: t.float64, np.float32: t.float32, np.int32: t.int32}[dtype]
    if isinstance(x, t.Tensor):
        return x.to(device="cuda", dtype=tdt).contiguous()
    a = np.ascontiguousarray(np.asarray(x), dtype=dtype)
    return t.from_numpy(a).to("cuda", non_blocking=False)


def empty(shape, dtype=np.float64):
    t = torch()
    tdt = {np.float64: t.float64, np.float32: t.float32, np.int32: t.int32, np.uint8: t.uint8}[dtype]
    return t.empty(shape, dtype=tdt, device="cuda")


def workspace(nbytes: int):
    t = torch()
    return t.empty(max(int(nbytes), 1), dtype=t.uint8, device="cuda")


def to_host(t_):
    return t_.detach().cpu().numpy()
