"""COO files and dictionaries (CPU): the reference's own tests
(tests/test_sparse_matrix_tools.py, on its fixture test_data/test_coo.json,
copied to tests/golden), and agreement with the oracle's loop restatement of
linac/sparse_matrix_tools.py:20-42."""
import os
from fractions import Fraction

import numpy
import pytest

from conftest import GOLDEN
from linac_amd import sparse_matrix_tools as smt
from oracle import coo as ocoo

COO_JSON = os.path.join(GOLDEN, "test_coo.json")
COO_LARGE = os.path.join(GOLDEN, "test_coo_large_fractions.json")


def test_empty_coo():
    """reference tests/test_sparse_matrix_tools.py:16-20 (seeded shapes here)."""
    rs = numpy.random.RandomState(0)
    for _ in range(5):
        zero = numpy.zeros((rs.randint(5, 11), rs.randint(5, 11)))
        coo = smt.coo_from_matrix(zero)
        back = smt.matrix_from_coo(coo)
        assert numpy.all(zero == back) and zero.shape == back.shape
        assert coo == ocoo.coo_from_matrix(zero)


def test_coo_fixture_sparsity_and_round_trip():
    """reference tests/test_sparse_matrix_tools.py:23-26: the fixture's sparsity is 95.180%."""
    matrix = smt.matrix_from_json_coo(COO_JSON)
    assert smt.sparsity(matrix) == '95.180%'
    assert numpy.all(smt.matrix_from_coo(smt.coo_from_matrix(matrix)) == matrix)
    assert all(isinstance(v, Fraction) for v in matrix[matrix != 0])


def test_json_and_plain_txt_round_trips(tmp_path):
    """reference tests/test_sparse_matrix_tools.py:29-41."""
    matrix = smt.matrix_from_json_coo(COO_JSON)
    smt.json_coo_from_matrix(matrix, str(tmp_path / "m.json"))
    assert numpy.all(smt.matrix_from_json_coo(str(tmp_path / "m.json")) == matrix)
    smt.plain_txt_coo_from_matrix(matrix, str(tmp_path / "m.txt"))
    assert numpy.all(smt.matrix_from_plain_txt_coo(str(tmp_path / "m.txt")) == matrix)


@pytest.mark.parametrize("kind", ["int", "float", "object", "uint32", "zero_last_row", "zero_last_col"])
def test_coo_matches_oracle(kind):
    """Same dictionary (keys, order, values) and the same dense matrix as the
    loop restatement, including the bottom-right entry that keeps the shape."""
    rs = numpy.random.RandomState(7)
    A = rs.randint(-3, 4, size=(9, 13)) * (rs.random_sample((9, 13)) < 0.3)
    if kind == "float":
        A = A / 4.0
    elif kind == "object":
        A = numpy.array([[Fraction(int(x), 3) for x in row] for row in A], dtype=object)
    elif kind == "uint32":
        A = numpy.abs(A).astype(numpy.uint32)
    elif kind == "zero_last_row":
        A[-1, :] = 0
    elif kind == "zero_last_col":
        A[:, -1] = 0
    got, want = smt.coo_from_matrix(A), ocoo.coo_from_matrix(A)
    assert list(got.items()) == list(want.items())
    back = smt.matrix_from_coo(got)
    assert back.shape == A.shape and back.dtype == object
    assert numpy.array_equal(back, ocoo.matrix_from_coo(want))


def test_large_fraction_values_survive_files(tmp_path):
    """Fractions with thousands of digits (reference fixture
    test_coo_large_fractions.json) round-trip through both formats exactly."""
    coo = smt.coo_from_json_coo(COO_LARGE)
    assert max(len(str(v)) for v in coo.values()) > 1000
    smt.json_coo_from_coo(coo, str(tmp_path / "l.json"))
    assert smt.coo_from_json_coo(str(tmp_path / "l.json")) == coo
    smt.plain_txt_coo_from_coo(coo, str(tmp_path / "l.txt"))
    assert smt.coo_from_plain_txt_coo(str(tmp_path / "l.txt")) == coo


def test_keys_are_parsed_not_evaluated(tmp_path):
    (tmp_path / "bad.json").write_text('{"__import__(\'os\')": "1"}')
    with pytest.raises(ValueError):
        smt.coo_from_json_coo(str(tmp_path / "bad.json"))
    (tmp_path / "short.txt").write_text("1 2\n")
    with pytest.raises(ValueError):
        smt.coo_from_plain_txt_coo(str(tmp_path / "short.txt"))


def test_key_arrays():
    r, c = smt._parse_keys(["(3, 4)", " (10,0) ", "(0, 7)"])
    assert r.tolist() == [3, 10, 0] and c.tolist() == [4, 0, 7]
    with pytest.raises(ValueError):
        smt._parse_keys(["(1, 2)", "(os.system('x'), 1)"])


def test_residues_mod_p():
    """Integers and fractions modulo p: a/b maps to a * b^(p-2) (Fermat), and
    the whole-array integer path agrees with Python's own modulo."""
    for p in (2 ** 31 - 1, 65537, 2 ** 61 - 1, 9223372036854775783):
        vals = ["-1", "0", "5", "123456789012345678", "-987654321098765432"]
        assert smt.residues_mod_p(vals, p).tolist() == [int(v) % p for v in vals]
        mixed = ["1/2", "-7/3", "5", "-12", "999999999999999999/7", "3/123456789012345678901"]
        got = smt.residues_mod_p(mixed, p).tolist()
        for g, v in zip(got, mixed):
            f = Fraction(v)
            assert g * (f.denominator % p) % p == f.numerator % p
        fr = ["1/2", "-7/3", "5", Fraction(10 ** 40 + 1, 3 ** 30)]
        got = smt.residues_mod_p(fr, p).tolist()
        for g, v in zip(got, fr):
            f = Fraction(v)
            assert g * (f.denominator % p) % p == f.numerator % p
    with pytest.raises(ValueError):
        smt.residues_mod_p(["1/7"], 7)
