"""Kernels of the fml KRR path on the GPU (fml/kernels/__init__.py:31-41)."""
from .kernels import (PTP, gaussian_kernel, laplacian_kernel, gaussian_kernel_rows,
                      laplacian_kernel_rows, get_atomic_kernels_arad,
                      get_atomic_symmetric_kernels_arad, fget_kernels_arad,
                      fget_symmetric_kernels_arad, fget_kernel_arad, fget_atomic_kernels_arad,
                      fget_atomic_distance_arad, fget_alpha_arad, arad_types)

__all__ = ["gaussian_kernel", "laplacian_kernel", "get_atomic_kernels_arad",
           "get_atomic_symmetric_kernels_arad", "fget_kernels_arad", "fget_symmetric_kernels_arad",
           "fget_kernel_arad", "fget_atomic_kernels_arad", "fget_atomic_distance_arad",
           "fget_alpha_arad"]
