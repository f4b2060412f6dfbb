"""Drop-in ``iterative_gaussian_solver`` (``linac/linear_system_solver.py:24-79``)
with the system kept in HBM.

The reference re-reduces, on every iteration, the system restricted to the
unknowns that survived the previous one (rows and columns of the same
indices, plus the augmented column), copying each matrix to the GPU and its
full RREF back.  Here the system is uploaded once; each iteration gathers the
surviving rows and columns on the device into an aligned buffer, reduces it
in place (``rref_device``: pivot columns and rank come from the device), and
copies back only the first ``rank`` entries of the solution column.  The
dropping rules are the reference's:

* unknowns whose columns are not pivots are redundant (the augmented column
  never counts, ``:47-49``);
* unknowns whose solution entry is close to 0 (``numpy.isclose(complex(v), 0)``:
  ``|v| <= 1e-8``, exactly 0 over GF(p)) are zero (``:51-52``);
* the loop ends when an iteration drops nothing new or drops every unknown,
  or after ``max_iterations`` (``:74-75``); dropped unknowns are 0 in the
  returned list, the others take the surviving solution values in order.

``use_gpu``, ``pivoting`` and ``scaling`` select the reference's CPU
elimination (``linac/row_reduce.py``), which this library does not provide:
every call runs on the device, whose RREF equals the CPU one over GF(p) and
matches it within the float tolerance (partial pivoting on row-scaled input,
the GPU path of the reference).  Like the reference, each iteration prints one
progress line.
"""
import numpy

from .hip_row_reduce import _prepare, rref_device


def _torch_dtype(torch, dt):
    return {numpy.dtype(numpy.uint32): torch.int32, numpy.dtype(numpy.uint64): torch.int64,
            numpy.dtype(numpy.float64): torch.float64, numpy.dtype(numpy.complex128): torch.complex128}[dt]


def iterative_gaussian_solver(matrix, use_gpu=True, max_iterations=1, pivoting=1, scaling=True, field_characteristic=0,
                              verbose=True):
    import torch
    shape = numpy.shape(matrix)
    if len(shape) != 2 or shape[0] + 1 != shape[1]:
        raise ValueError(f"expected an n x (n + 1) augmented system, got shape {shape}")
    n = shape[0]
    p, host = _prepare(matrix, field_characteristic, None, None)
    dt = host.dtype
    tdt = _torch_dtype(torch, dt)
    per = 16 // dt.itemsize  # elements per 16 bytes: device rows must be 16-byte aligned
    raw = host.view(numpy.int32) if dt == numpy.uint32 else host.view(numpy.int64) if dt == numpy.uint64 else host
    system = torch.from_numpy(numpy.ascontiguousarray(raw)).to("cuda")
    live = list(range(n))
    redundant, zero = set(), set()
    solution = [0] * n
    for iteration in range(max_iterations):
        before = redundant | zero
        m = len(live)
        if m == 0:
            break
        rows = torch.tensor(live, dtype=torch.int64, device="cuda")
        cols = torch.tensor(live + [n], dtype=torch.int64, device="cuda")
        ld = (m + 1 + per - 1) // per * per
        sub = torch.zeros((m, ld), dtype=tdt, device="cuda")
        sub[:, :m + 1] = system.index_select(0, rows).index_select(1, cols)
        torch.cuda.current_stream().synchronize()  # the library works on its own stream
        piv, rank = rref_device(sub.data_ptr(), m, m + 1, ld, p, dt)
        pivot_set = set(piv)
        local_redundant = [j for j in range(m) if j not in pivot_set]
        col = sub[:rank, m].cpu().numpy()
        if dt in (numpy.uint32, numpy.uint64):
            col = col.view(dt)  # the device buffer holds the unsigned values' bits
        col = col.tolist()
        zero_pos = [i for i, v in enumerate(col) if numpy.isclose(complex(v), 0)]
        pivot_unknowns = [live[j] for j in range(m) if j in pivot_set]
        redundant |= {live[j] for j in local_redundant}
        zero |= {pivot_unknowns[i] for i in zero_pos}
        values = iter([v for i, v in enumerate(col) if i not in zero_pos])
        dropped = redundant | zero
        solution = [0 if i in dropped else next(values) for i in range(n)]
        print("Iteration number {}: dropped_redundant: {}, dropped_zero: {}, dropped_total: {}.".format(
            iteration + 1, len(redundant), len(zero), len(dropped)))
        live = [u for u in live if u not in dropped]
        if dropped == before or len(dropped) == n:
            break
    return solution
