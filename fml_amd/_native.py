"""ctypes binding of libfmlhip.so (the C ABI declared in include/fmlh.h).

There is no CPU fallback: importing works without a GPU (so the package can
be inspected and the symbol table checked), but every compute call needs the
built library and a visible HIP device, and raises otherwise.

Device memory and streams come from PyTorch (ROCm build): buffers are torch
tensors on ``cuda`` devices, passed as raw pointers, and each call is
enqueued on ``torch.cuda.current_stream()``.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfmlhip.so")

c_void_p = ctypes.c_void_p
c_int = ctypes.c_int
c_int64 = ctypes.c_int64
c_double = ctypes.c_double
c_size_t = ctypes.c_size_t
c_dbl_p = ctypes.POINTER(ctypes.c_double)

#: name -> (restype, argtypes); mirrors include/fmlh.h
SIGNATURES = {
    "fmlh_abi_version": (c_int, []),
    "fmlh_last_error": (ctypes.c_char_p, []),
    "fmlh_device_cu_count": (c_int, []),
    "fmlh_init": (c_int, []),
    "fmlh_stream_create": (c_int, [c_int, c_void_p]),
    "fmlh_stream_destroy": (c_int, [c_void_p]),
    "fmlh_graph_begin": (c_int, [c_void_p]),
    "fmlh_graph_end": (c_int, [c_void_p, c_void_p]),
    "fmlh_graph_launch": (c_int, [c_void_p, c_void_p]),
    "fmlh_graph_destroy": (c_int, [c_void_p]),
    "fmlh_coulomb_sorted": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p,
                                    c_void_p, c_void_p, c_void_p]),
    "fmlh_gaussian_kernel": (c_int, [c_void_p, c_int64, c_int, c_void_p, c_int64, c_int, c_int, c_double,
                                     c_void_p, c_int64, c_void_p]),
    "fmlh_gaussian_kernel_f32_ws_bytes": (c_size_t, [c_int, c_int, c_int, c_int, c_int]),
    "fmlh_gaussian_kernel_f32": (c_int, [c_void_p, c_int64, c_int, c_void_p, c_int64, c_int, c_int,
                                         c_double, c_int, c_void_p, c_void_p, c_int64, c_void_p,
                                         c_size_t, c_void_p, c_void_p]),
    "fmlh_laplacian_kernel": (c_int, [c_void_p, c_int64, c_int, c_void_p, c_int64, c_int, c_int, c_double,
                                      c_void_p, c_int64, c_void_p]),
    "fmlh_laplacian_kernel_f32": (c_int, [c_void_p, c_int64, c_int, c_void_p, c_int64, c_int, c_int,
                                          c_double, c_void_p, c_void_p, c_int64, c_void_p]),
    "fmlh_arad_describe": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_double, c_void_p, c_int,
                                   c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "fmlh_coulomb_form": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p,
                                  c_void_p]),
    "fmlh_vector_kernels": (c_int, [c_int, c_void_p, c_int64, c_void_p, c_int, c_int, c_void_p, c_int64,
                                    c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_int, c_void_p,
                                    c_int, c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "fmlh_distance_mol_kernel": (c_int, [c_void_p, c_int64, c_void_p, c_int, c_int, c_int, c_double,
                                         c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "fmlh_atomic_predict_ws_bytes": (c_size_t, [c_int, c_int]),
    "fmlh_atomic_predict": (c_int, [c_void_p, c_int64, c_void_p, c_int, c_int, c_void_p, c_int64, c_int,
                                    c_int, c_int, c_void_p, c_double, c_void_p, c_int, c_void_p,
                                    c_void_p, c_void_p, c_size_t, c_void_p]),
    "fmlh_distance_predict": (c_int, [c_void_p, c_int64, c_void_p, c_int, c_int, c_int, c_int, c_void_p,
                                      c_double, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_size_t,
                                      c_void_p]),
    "fmlh_aras_ws_bytes": (c_size_t, [c_int, c_int, c_int, c_int, c_int]),
    "fmlh_aras_kernels": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                                  c_int, c_int, c_void_p, c_int, c_double, c_double, c_double, c_int,
                                  c_void_p, c_int, c_double, c_void_p, c_void_p, c_int, c_void_p, c_void_p,
                                  c_void_p, c_size_t, c_void_p]),
    "fmlh_arad_ws_bytes": (c_size_t, [c_int, c_int, c_int, c_int]),
    "fmlh_arad_kernels": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                                  c_int, c_int, c_dbl_p, c_int, c_double, c_double, c_double, c_double,
                                  c_int, c_void_p, c_void_p, c_size_t, c_void_p, c_void_p]),
    "fmlh_arad_autotune": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                                   c_int, c_int, c_dbl_p, c_int, c_double, c_double, c_double, c_double,
                                   c_int, c_void_p, c_void_p, c_size_t, c_int, c_void_p, c_dbl_p, c_void_p]),
    "fmlh_arad_tiling_get": (c_int, [c_int, c_int, c_int, c_int, c_void_p]),
    "fmlh_arad_tiling_set": (c_int, [c_int, c_int, c_int, c_int, c_void_p]),
    "fmlh_arad_atomic_ws_bytes": (c_size_t, [c_int, c_int, c_int]),
    "fmlh_arad_atomic": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                                 c_int, c_int, c_dbl_p, c_int, c_double, c_double, c_double, c_double,
                                 c_void_p, c_void_p, c_size_t, c_void_p]),
    "fmlh_potrf_ws_bytes": (c_size_t, [c_int]),
    "fmlh_potrf": (c_int, [c_void_p, c_int64, c_int, c_void_p, c_void_p, c_size_t, c_void_p]),
    "fmlh_potrf_ozaki_ws_bytes": (c_size_t, [c_int]),
    "fmlh_potrf_ozaki": (c_int, [c_void_p, c_int64, c_int, c_void_p, c_void_p, c_size_t, c_void_p]),
    "fmlh_potri_ws_bytes": (c_size_t, [c_int]),
    "fmlh_potri": (c_int, [c_void_p, c_int64, c_int, c_void_p, c_void_p, c_size_t, c_void_p]),
    "fmlh_potrs_ws_bytes": (c_size_t, [c_int, c_int]),
    "fmlh_potrs": (c_int, [c_void_p, c_int64, c_int, c_void_p, c_int64, c_int, c_void_p, c_size_t,
                           c_void_p]),
    "fmlh_potrs_correction": (c_int, [c_void_p, c_int64, c_int, c_void_p, c_void_p, c_size_t, c_void_p]),
    "fmlh_potrs_noalloc_max_bytes": (c_size_t, []),
    "fmlh_potrs_noalloc": (c_int, [c_void_p, c_int64, c_int, c_void_p, c_int64, c_int, c_void_p]),
    "fmlh_spotrf_ws_bytes": (c_size_t, [c_int]),
    "fmlh_spotrf": (c_int, [c_void_p, c_int64, c_int, c_void_p, c_void_p, c_size_t, c_void_p]),
    "fmlh_spotrs_ws_bytes": (c_size_t, [c_int, c_int]),
    "fmlh_spotrs": (c_int, [c_void_p, c_int64, c_int, c_void_p, c_int64, c_int, c_void_p, c_size_t,
                            c_void_p]),
    "fmlh_add_diagonal": (c_int, [c_void_p, c_int64, c_int, c_double, c_void_p]),
    "fmlh_sym_residual": (c_int, [c_void_p, c_int64, c_int, c_double, c_void_p, c_void_p, c_void_p,
                                  c_void_p]),
    "fmlh_gemv_t": (c_int, [c_void_p, c_int64, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "fmlh_residual_dd": (c_int, [c_void_p, c_int64, c_int, c_double, c_void_p, c_void_p, c_void_p,
                                 c_void_p]),
    "fmlh_residual_dd_lower_ws_bytes": (c_size_t, [c_int]),
    "fmlh_residual_dd_lower": (c_int, [c_void_p, c_int64, c_int, c_void_p, c_double, c_void_p, c_void_p,
                                       c_void_p, c_void_p, c_size_t, c_void_p]),
    "fmlh_krr_predict_ws_bytes": (c_size_t, [c_int, c_int, c_int]),
    "fmlh_krr_predict": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p, c_int64, c_int,
                                 c_void_p, c_double, c_void_p, c_void_p, c_void_p, c_size_t, c_void_p]),
}


class AradTiling(ctypes.Structure):
    """fmlh_arad_tiling: tile shape and launch shape of the ARAD tile kernel."""
    _fields_ = [("pairs_per_tile", c_int), ("waves_per_simd", c_int), ("groups_per_cu", c_int),
                ("pair_order", c_int)]

    def as_tuple(self):
        return (self.pairs_per_tile, self.waves_per_simd, self.groups_per_cu, self.pair_order)


ARAD_GLOBAL, ARAD_SYMMETRIC, ARAD_KERNEL, ARAD_DISTANCE = 0, 1, 2, 3
ARAD_ATOMIC_KERNELS, ARAD_ATOMIC_DISTANCE = 4, 5
FMLH_EHIP = -1000


class NativeError(RuntimeError):
    pass


_lib = None


def load(path: str = LIB_PATH):
    """Load libfmlhip.so and bind every symbol of the C ABI."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise NativeError(
            f"{path} is missing: build it with `python -m fml_amd.build` "
            "(there is no CPU fallback)")
    # PyTorch ships its own HIP runtime under the same soname; load it first
    # so this library binds to the runtime that owns the device memory.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, fn: str):
    if rc == 0:
        return
    if rc == FMLH_EHIP:
        raise NativeError(f"{fn}: HIP error: {load().fmlh_last_error().decode()}")
    raise ValueError(f"{fn}: invalid argument #{-rc}")


def call(fn: str, *args):
    rc = getattr(load(), fn)(*args)
    check(rc, fn)


def torch_mod():
    import torch
    if not torch.cuda.is_available():
        raise NativeError("no HIP device is visible; fml_amd runs on MI355X (gfx950) only")
    load()
    return torch


def stream_ptr():
    import torch
    return c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return c_void_p(0)
    return c_void_p(t.data_ptr())
