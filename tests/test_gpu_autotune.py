"""The autotuner measures on the device, sets and saves its choices, and the
knobs it moves leave GF(p) results bit-exact."""
import json

import numpy
import pytest

from oracle import crref, modp

pytestmark = pytest.mark.gpu


def test_quick_tune_sets_and_saves(gpu, tmp_path):
    from linac_amd import autotune
    before = autotune.current()
    try:
        data = autotune.tune(quick=True, path=str(tmp_path / "tune.json"))
        saved = json.load(open(data["file"]))
        assert saved["knobs"] == data["knobs"] == {k: autotune.get(k) for k in data["knobs"]}
        assert data["measurements"]["gf_crossover"] and data["measurements"]["float_crossover"]
        assert "gfx" in data["device"]
    finally:
        autotune.apply(before)


@pytest.mark.parametrize("group, lookahead", [(1, 0), (4, 1), (16, 0), (64, 1)])
def test_knobs_keep_gf_results_exact(gpu, group, lookahead):
    from linac_amd import autotune
    before = autotune.current()
    p = 2 ** 31 - 1
    A = modp.low_rank(700, 900, 500, p, 21)
    try:
        autotune.set_knob("gemm_group", group)
        autotune.set_knob("lookahead", lookahead)
        R, piv, rank = gpu.rref_with_pivots(A, p, method="blocked")
    finally:
        autotune.apply(before)
    Ro, po = crref.rref_modp(A, p)
    assert piv == po and numpy.array_equal(R, Ro)
