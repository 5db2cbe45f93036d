"""Kernels of the fml KRR path on the GPU (fml/kernels/__init__.py:31-41,
fml/kernels/kernels.py)."""
from .kernels import (PTP, gaussian_kernel, laplacian_kernel, gaussian_kernel_rows,
                      laplacian_kernel_rows, get_atomic_kernels_arad,
                      get_atomic_symmetric_kernels_arad, fget_kernels_arad,
                      fget_symmetric_kernels_arad, fget_kernel_arad, fget_atomic_kernels_arad,
                      fget_atomic_distance_arad, fget_alpha_arad, arad_types)
from .atomic import (get_atomic_kernels_gaussian, get_atomic_kernels_laplacian,
                     fget_vector_kernels_gaussian, fget_vector_kernels_laplacian, get_alpha,
                     get_prediction, get_alpha_from_distance, get_prediction_from_distance,
                     fget_alpha, fget_prediction, fget_alpha_from_distance,
                     fget_prediction_from_distance)
from .aras_kernels import (get_atomic_kernels_aras, get_atomic_symmetric_kernels_aras,
                           fget_kernels_aras, fget_symmetric_kernels_aras, gen_pd, periodic_distance)

__all__ = ["gaussian_kernel", "laplacian_kernel", "get_atomic_kernels_arad",
           "get_atomic_symmetric_kernels_arad", "fget_kernels_arad", "fget_symmetric_kernels_arad",
           "fget_kernel_arad", "fget_atomic_kernels_arad", "fget_atomic_distance_arad",
           "fget_alpha_arad", "get_atomic_kernels_gaussian", "get_atomic_kernels_laplacian",
           "fget_vector_kernels_gaussian", "fget_vector_kernels_laplacian", "get_alpha",
           "get_prediction", "get_alpha_from_distance", "get_prediction_from_distance",
           "fget_alpha", "fget_prediction", "fget_alpha_from_distance",
           "fget_prediction_from_distance", "get_atomic_kernels_aras",
           "get_atomic_symmetric_kernels_aras", "fget_kernels_aras", "fget_symmetric_kernels_aras",
           "gen_pd", "periodic_distance"]
