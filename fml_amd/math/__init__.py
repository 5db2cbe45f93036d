"""fml.math drop-ins on the GPU: cho_solve and cho_invert
(fml/math/cho_solve.py:27-58) and
the ARAD atomic distance matrices (fml/math/distance.py:120-189,
fml/math/farad.f90:452-600)."""
from .cho_solve import cho_solve, cho_invert, cho_factor, cho_solve_factored
from .distance import get_l2_distance_arad, atomic_arad_l2_distance_all

__all__ = ["cho_solve", "cho_invert", "get_l2_distance_arad", "atomic_arad_l2_distance_all"]
