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


def check_offsets(offsets, natoms: int) -> None:
    """Molecule m owns atoms offsets[m]:offsets[m+1]: the offsets must start
    at 0 or more, never decrease and end within the natoms atoms.  Device
    tensors are taken as given (checking them would wait on the GPU)."""
    if is_device_tensor(offsets):
        return
    off = np.asarray(offsets).reshape(-1)
    if off.size == 0:
        raise ValueError("offsets needs nmol + 1 entries")
    if off[0] < 0 or np.any(np.diff(off) < 0) or off[-1] > natoms:
        raise ValueError(f"offsets must be non-decreasing within 0..{natoms}")


def empty(shape, dtype=np.float64):
    t = torch()
    tdt = {np.float64: t.float64, np.float32: t.float32, np.int32: t.int32, np.uint8: t.uint8}[dtype]
    return t.empty(shape, dtype=tdt, device="cuda")


def workspace(nbytes: int):
    t = torch()
    return t.empty(max(int(nbytes), 1), dtype=t.uint8, device="cuda")


def to_host(t_):
    return t_.detach().cpu().numpy()
