"""CPU oracle for the dense row-reduction path.

This package is TEST INFRASTRUCTURE. It restates, on the CPU, the algorithms the
GPU library must agree with, and is imported only by the test-suite, by the
smoke check in ``__graft_entry__`` and by the CPU-baseline leg of ``bench.py``.
The product package ``linac_amd`` never imports it and never falls back to it.

Modules
-------
rref      numpy restatement of ``linac/row_reduce.py:16-78`` (all pivoting and
          scaling modes, exact int64 mod-p path, float/complex path).
crref     ctypes binding of ``rref.c``, a C restatement of the same prime path
          (uint64 storage, unsigned __int128 products, optional OpenMP).
modp      exact mod-p helpers used to build and certify test inputs
          (limb-split matmul, Freivalds check, planted solutions, sympy check).
loader    restatement of ``linac/pycuda_matrix_loader.py:127-230`` and of the
          ``LoadMatrix`` kernel's arithmetic (``linac/matrix_loader.cu:88-112``).
pivots    restatement of ``linac/linear_algebra_tools.py:10-36,48-78`` (pivots, canonical kernels).
coo       restatement of ``linac/sparse_matrix_tools.py:20-42`` (COO dictionaries).
solver    restatement of ``linac/linear_system_solver.py:24-79`` (iterative_gaussian_solver).
"""
