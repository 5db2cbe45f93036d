"""ARAD atomic distance matrices (fml/math/distance.py:120-189 ->
fml/math/farad.f90:452-600) on the GPU."""
from __future__ import annotations

from .. import _native
from ..device import is_device_tensor, to_host
from ..kernels.kernels import _arad_call, _check_arad_shapes, arad_types


def atomic_arad_l2_distance_all(q1, q2, z1, z2, n1, n2, nm1=None, nm2=None, width=0.2,
                                cut_distance=6.0, r_width=1.0, c_width=0.5, amax=None):
    """farad.f90:452 -> dmatrix (nm1, nm2, amax, amax) Fortran order."""
    dev = is_device_tensor(q1)
    out = _arad_call(_native.ARAD_DISTANCE, q1, z1, n1, q2, z2, n2, None, width, cut_distance,
                     r_width, c_width)
    return out.permute(3, 2, 1, 0) if dev else to_host(out).transpose(3, 2, 1, 0)


def get_l2_distance_arad(X1, X2, Z1, Z2, width=0.2, cut_distance=6.0, r_width=1.0, c_width=0.5):
    """Atomic ARAD distance matrices for two molecule sets
    (distance.py:120-189) -> (N1, N2, amax, amax)."""
    amax = _check_arad_shapes(X1, X2, Z1, Z2)
    z1, n1 = arad_types(Z1, amax)
    z2, n2 = arad_types(Z2, amax)
    return atomic_arad_l2_distance_all(X1, X2, z1, z2, n1, n2, width=width,
                                       cut_distance=cut_distance, r_width=r_width, c_width=c_width)
