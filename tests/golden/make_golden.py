"""Writes the golden fixtures under tests/golden/ (run once, in a container that has
the reference tree; the outputs are committed and the tests only read them).

* ``known_result_4096_4097_2to32m19.npy`` -- data copied from the reference's
  tests/test_data (uint32[4096]): the last column of the RREF of
  ``numpy.random.seed(0); randint(2**31-19, size=(4097, 4097), dtype=uint32)[:-1, :]``
  (reference tests/test_row_reduce_random_cuda.py:52-63).  Loaded with
  ``allow_pickle=False``.
* ``gf_small_cases.npz`` -- seeded small matrices over several primes and their
  RREF and pivot columns computed by sympy's ``DomainMatrix(..., GF(p)).rref()``
  (a library routine independent of the oracle).
* ``c0_config.json`` -- the CPU-runnable config: a 500x500 matrix mod 2**31-1,
  the mod-p product of a seed-0 500x400 and 400x500 uniform matrix (rank 400);
  sha256 of its RREF (uint32, C order) and its pivot columns, from sympy.
"""
import hashlib
import json
import os
import sys

import numpy

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import modp  # noqa: E402

REFERENCE_DATA = "/root/reference/tests/test_data/known_result_4096_4097_2to32m19.npy"

PRIMES = [2, 3, 65537, 2 ** 31 - 19, 2 ** 31 - 1, 2 ** 31 + 11, 2 ** 32 - 5, 2 ** 61 - 1]


def small_cases():
    rs = numpy.random.RandomState(20240601)
    cases = []
    shapes = [(7, 7, None), (12, 20, 5), (30, 18, 11), (16, 16, 16), (9, 25, None), (20, 8, 3), (1, 5, None),
              (6, 1, None), (24, 40, 12)]
    for p in PRIMES:
        for (r, c, rank) in shapes:
            if rank is None:
                A = rs.randint(0, min(p, 2 ** 62), size=(r, c), dtype=numpy.int64)
            else:
                L = rs.randint(0, min(p, 2 ** 62), size=(r, rank), dtype=numpy.int64).astype(object)
                U = rs.randint(0, min(p, 2 ** 62), size=(rank, c), dtype=numpy.int64).astype(object)
                A = numpy.array((L @ U) % p, dtype=numpy.int64) if p < 2 ** 62 else None
            if c > 3:
                A[:, 2] = 0  # an all-zero column
            R, piv = modp.sympy_rref(A, p)
            Rfull = numpy.zeros((r, c), dtype=numpy.uint64)
            Rfull[:R.shape[0]] = numpy.array(R, dtype=object).astype(numpy.uint64)
            cases.append((p, A.astype(numpy.uint64), Rfull, numpy.array(piv, dtype=numpy.int64)))
    return cases


def main():
    known = numpy.load(REFERENCE_DATA, allow_pickle=False)
    assert known.dtype == numpy.uint32 and known.shape == (4096,)
    numpy.save(os.path.join(HERE, "known_result_4096_4097_2to32m19.npy"), known)

    cases = small_cases()
    blob = {}
    for k, (p, A, R, piv) in enumerate(cases):
        blob[f"p_{k}"] = numpy.array([p], dtype=numpy.uint64)
        blob[f"A_{k}"] = A
        blob[f"R_{k}"] = R
        blob[f"piv_{k}"] = piv
    blob["count"] = numpy.array([len(cases)])
    numpy.savez_compressed(os.path.join(HERE, "gf_small_cases.npz"), **blob)

    p = 2 ** 31 - 1
    rs = numpy.random.RandomState(0)
    L = rs.randint(p, size=(500, 400), dtype=numpy.uint32)
    U = rs.randint(p, size=(400, 500), dtype=numpy.uint32)
    A = modp.matmul_modp(L, U, p)
    R, piv = modp.sympy_rref(A, p)
    Rfull = numpy.zeros((500, 500), dtype=numpy.uint32)
    Rfull[:R.shape[0]] = numpy.array(R, dtype=object).astype(numpy.uint32)
    with open(os.path.join(HERE, "c0_config.json"), "w") as fh:
        json.dump({"p": p, "shape": [500, 500], "rank": len(piv), "pivots": [int(x) for x in piv],
                   "input_sha256": hashlib.sha256(numpy.ascontiguousarray(A.astype(numpy.uint32)).tobytes()).hexdigest(),
                   "rref_sha256": hashlib.sha256(numpy.ascontiguousarray(Rfull).tobytes()).hexdigest(),
                   "generator": "RandomState(0): L=randint(p,(500,400),uint32), U=randint(p,(400,500),uint32), A=L@U mod p",
                   "rref_by": "sympy DomainMatrix GF(p) rref"}, fh, indent=1)


if __name__ == "__main__":
    main()
