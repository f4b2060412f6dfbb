"""Writes the oracle fixtures of the full-size configurations (run once here,
on the CPU; the outputs are committed and the GPU tests only read them).

* ``lu5000.json`` -- the N = 5000 GF(2^31-1) input of rank 4500, the seeded
  mod-p product modp.low_rank(5000, 5000, 4500, p, 11): sha256 of the input,
  the C oracle's pivot columns and the sha256 of its RREF (uint32, C order).
* ``float_configs.npz`` -- for the float configurations whose inputs are pure
  RandomState draws, [A | b] (tests/certificates.float_system: f64 8000, c128
  6000, seed 0; the ill-conditioned one is built with LAPACK and its oracle runs
  in the test): sha256 of the input, the C oracle's FP64 RREF solution column
  ``x_oracle`` (linac's float path: row scaling, partial pivoting, 1e-9
  threshold), its max |R[:, :n] - I|, and ``x_star``: the same solution
  refined to extended accuracy (LAPACK LU corrections, residuals b - A x in
  80-bit long double) -- the reference point for errors, since the FP64 oracle
  itself carries an error of order cond(A) * 2^-53.  Also cond_2(A).

Usage: python tests/golden/make_config_fixtures.py [lu5000] [f64] [c128]
"""
import hashlib
import json
import os
import sys
import time

import numpy

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from oracle import crref, modp  # noqa: E402
import certificates  # noqa: E402

P31 = 2 ** 31 - 1
FLOAT_CASES = {"f64": 8000, "c128": 6000}


def sha(a):
    return hashlib.sha256(numpy.ascontiguousarray(a).tobytes()).hexdigest()


def lu5000():
    A = modp.low_rank(5000, 5000, 4500, P31, 11).astype(numpy.uint32)
    t = time.time()
    R, piv = crref.rref_modp(A, P31)
    print(f"lu5000: oracle {time.time() - t:.1f} s, rank {len(piv)}", flush=True)
    out = {"p": P31, "rows": 5000, "cols": 5000, "rank": len(piv), "seed": 11, "input_sha256": sha(A),
           "pivots": piv, "rref_sha256": sha(R)}
    with open(os.path.join(HERE, "lu5000.json"), "w") as f:
        json.dump(out, f)


def float_case(kind):
    n = FLOAT_CASES[kind]
    Ab = certificates.float_system(kind, n, seed=0)
    t = time.time()
    R, piv = crref.rref_c128(Ab) if kind == "c128" else crref.rref_f64(Ab)
    t_or = time.time() - t
    assert piv == list(range(n)), (kind, len(piv))
    x_or = R[:, n].copy()
    eye_err = float(numpy.max(numpy.abs(R[:, :n] - numpy.eye(n))))
    del R
    A, b = Ab[:, :n], Ab[:, n]
    x_star = certificates.refine_solution(A, b, numpy.linalg.solve(A, b))
    cond = float(numpy.linalg.cond(A))
    err = float(numpy.max(numpy.abs(x_or - x_star)) / numpy.max(numpy.abs(x_star)))
    print(f"{kind} {n}: oracle {t_or:.1f} s, cond {cond:.3g}, oracle vs refined {err:.3g}, |R-I| {eye_err:.3g}",
          flush=True)
    return {f"{kind}_input_sha256": numpy.array(sha(Ab)), f"{kind}_x_oracle": x_or, f"{kind}_x_star": x_star,
            f"{kind}_cond": numpy.array(cond), f"{kind}_oracle_eye_err": numpy.array(eye_err),
            f"{kind}_oracle_vs_star": numpy.array(err)}


def main(which):
    if "lu5000" in which:
        lu5000()
    kinds = [k for k in FLOAT_CASES if k in which]
    if kinds:
        path = os.path.join(HERE, "float_configs.npz")
        data = dict(numpy.load(path, allow_pickle=False)) if os.path.exists(path) else {}
        for k in kinds:
            data.update(float_case(k))
            numpy.savez(path, **data)


if __name__ == "__main__":
    main(sys.argv[1:] or ["lu5000", "f64", "c128"])
