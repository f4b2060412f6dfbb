"""Availability flags (the role of ``linac/_optional.py:4-23``).

``CUDA_FOUND`` keeps the reference's name so that callers gating on it
(``linac/vector_spaces/over_number_fields.py:42,54``) keep working; here it
means "liblinac_amd.so is built and a GPU is visible".  ``GPU_FOUND`` is the
same flag under a neutral name.  Like the reference, computing it touches the
GPU runtime in this process, so a process forked afterwards cannot use the GPU.
"""
from . import _lib

try:
    import galois  # noqa: F401
    GALOIS_FOUND = True
except ImportError:
    GALOIS_FOUND = False

GPU_FOUND = _lib.available()
CUDA_FOUND = GPU_FOUND
