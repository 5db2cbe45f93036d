"""fml_amd -- the fml kernel-ridge-regression hot path on AMD Instinct MI355X.

Drop-in for the reference package's API on that path (fml/__init__.py,
fml/fml.py, fml/kernels, fml/math, fml/representations):

    import fml_amd as fml
    mol = fml.Molecule(); mol.read_xyz("x.xyz"); mol.generate_coulomb_matrix(23)
    K = fml.kernels.gaussian_kernel(A, B, sigma)
    alpha = fml.math.cho_solve(K, y)

Compute runs in hand-written HIP kernels for gfx950 (libfmlhip.so, C ABI in
include/fmlh.h) called through ctypes; there is no CPU fallback.
"""
from .data import NUCLEAR_CHARGE
from .arad import ARAD, generate_arad_descriptors
from .aras import ARAS
from .fml import Molecule
from . import kernels, math, representations

__version__ = "0.1.0"
__all__ = ["Molecule", "ARAD", "ARAS", "generate_arad_descriptors", "NUCLEAR_CHARGE", "kernels", "math", "representations"]
