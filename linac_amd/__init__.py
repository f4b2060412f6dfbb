"""linac_amd: linac's dense row-reduction hot path, native on MI355X (gfx950).

Drop-in names (``linac/__init__.py:1-24``): ``cuda_row_reduce``,
``cuda_load_matrix``, ``load_matrices``, ``CUDA_FOUND``, ``GALOIS_FOUND``,
``timeit``, ``iterative_gaussian_solver``.  The compute runs in ``liblinac_amd.so`` (HIP C ABI,
``include/linac_amd.h``) through ctypes.
"""
from ._optional import CUDA_FOUND, GALOIS_FOUND, GPU_FOUND
from .hip_row_reduce import cuda_row_reduce, rref_with_pivots, rref_device
from .hip_matrix_loader import cuda_load_matrix, load_matrices, load_matrices_stacked
from .batch import Batch, rref_batch
from .timeit_decorator import timeit
from .linear_algebra_tools import pivot_columns_from_row_reduced_echelon_form, drop_bottom_zero_rows
from .linear_system_solver import iterative_gaussian_solver

__version__ = "0.1.0"

__all__ = [
    "__version__",
    "CUDA_FOUND",
    "GPU_FOUND",
    "GALOIS_FOUND",
    "cuda_row_reduce",
    "rref_with_pivots",
    "rref_device",
    "cuda_load_matrix",
    "load_matrices",
    "load_matrices_stacked",
    "Batch",
    "rref_batch",
    "timeit",
    "pivot_columns_from_row_reduced_echelon_form",
    "drop_bottom_zero_rows",
    "iterative_gaussian_solver",
]
