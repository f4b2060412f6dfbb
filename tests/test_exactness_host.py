"""Exactness of the GF(p) arithmetic at its stated worst cases, on the CPU:
the library's own fold and field routines (linac_amd/csrc/modp_gemm.hpp,
field.hpp, compiled for the host) against Python integers.

* fold_sub (int8-limb GEMM, 4 digits): every diagonal sum S_s at the extremes
  its K = 512 products reach (npairs(s) * K * 2^14 and -npairs(s) * K * 16256)
  and at the bound its comment states (|S_s| <= 2^25), in all 2^7 sign
  patterns, for primes up to the largest the limb GEMM accepts.
* fold_sub_k32 (inner GEMM, K <= 32) likewise at |S_s| <= 2^21.
* fold_sub64 (8 digits, 15 diagonals, K = 512, |S_s| <= 2^26) for primes
  from just above 2^32 to just below 2^63, on random sign patterns.
* ff_update, mul, inverse (Montgomery, R = 2^32 / 2^64), binary_inverse
  (binary extended Euclid) and euclid_inverse_u32 (extended Euclid in double
  precision, the inner factorisation's one inversion) on boundary operands.
* balanced digits of centred values at the ends of their ranges.

These replace the reference's arbitrary-precision modular arithmetic
(linac/row_reduce.cu:120-199, linac/pycuda_row_reduce.py:61-92): any result
that differs from the integer one would break bit-exact parity."""
import itertools
import os
import random
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "linac_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
NPAIRS4 = [1, 2, 3, 4, 3, 2, 1]
NPAIRS8 = [min(s, 14 - s) + 1 for s in range(15)]


def is_prime(n):
    if n < 2:
        return False
    for q in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if n % q == 0:
            return n == q
    d, r = n - 1, 0
    while d % 2 == 0:
        d //= 2
        r += 1
    for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):  # deterministic below 3.3e24
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(r - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def prev_prime(n):
    while not is_prime(n):
        n -= 1
    return n


def next_prime(n):
    while not is_prime(n):
        n += 1
    return n


LIMB_P_MAX = 4278124287
P32 = [3, 5, 251, 257, 65521, 65537, 2 ** 31 - 1, next_prime(2 ** 31), 3221225473, prev_prime(LIMB_P_MAX)]
P64 = [next_prime(2 ** 32), 2 ** 61 - 1, prev_prime(2 ** 62), prev_prime(2 ** 63 - 1)]


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if not shutil.which(HIPCC) and not os.path.exists(HIPCC):
        pytest.skip("hipcc not found")
    exe = str(tmp_path_factory.mktemp("exact") / "exact_host")
    subprocess.run([HIPCC, "--offload-host-only", "-O2", "-std=c++17", "-I", CSRC,
                    os.path.join(HERE, "host", "exact_host.cpp"), "-o", exe], check=True)

    def run(lines):
        out = subprocess.run([exe], input="\n".join(lines) + "\n", capture_output=True, text=True, check=True)
        return out.stdout.split("\n")[:len(lines)]
    return run


def _fold_expected(p, c, s):
    return (c - sum(v * 256 ** i for i, v in enumerate(s))) % p


def _fold_cases(npairs, k, bound, primes, rng, patterns):
    hi = [n * k * 16384 for n in npairs]
    lo = [-n * k * 16256 for n in npairs]
    cases = []
    for p in primes:
        for c in (0, p - 1, rng.randrange(p)):
            for pat in patterns:
                cases.append((p, c, [hi[i] if b else lo[i] for i, b in enumerate(pat)]))
                cases.append((p, c, [bound if b else -bound for b in pat]))
            for _ in range(50):
                cases.append((p, c, [rng.randint(lo[i], hi[i]) for i in range(len(npairs))]))
    return cases


@pytest.mark.parametrize("op, k, bound", [("F32", 512, 2 ** 25), ("K32", 32, 2 ** 21)])
def test_fold_sub_u32_worst_case(harness, op, k, bound):
    rng = random.Random(1)
    cases = _fold_cases(NPAIRS4, k, bound, P32, rng, list(itertools.product((0, 1), repeat=7)))
    got = harness([f"{op} {p} {c} " + " ".join(map(str, s)) for p, c, s in cases])
    bad = [(p, c, s, g) for (p, c, s), g in zip(cases, got) if int(g) != _fold_expected(p, c, s)]
    assert not bad, bad[:3]


def test_fold_sub64_worst_case(harness):
    rng = random.Random(2)
    pats = [tuple([b] * 15) for b in (0, 1)] + [tuple(i % 2 for i in range(15)), tuple(1 - i % 2 for i in range(15))]
    pats += [tuple(rng.randint(0, 1) for _ in range(15)) for _ in range(300)]
    cases = _fold_cases(NPAIRS8, 512, 2 ** 26, P64, rng, pats)
    got = harness(["F64 %d %d " % (p, c) + " ".join(map(str, s)) for p, c, s in cases])
    bad = [(p, c, s, g) for (p, c, s), g in zip(cases, got) if int(g) != _fold_expected(p, c, s)]
    assert not bad, bad[:3]


def _operands(p, rng):
    edge = sorted({0, 1, 2, (p - 1) // 2, (p + 1) // 2, p - 2, p - 1})
    return edge + [rng.randrange(p) for _ in range(6)]


@pytest.mark.parametrize("bits", [32, 64])
def test_montgomery_update_mul_inverse(harness, bits):
    rng = random.Random(bits)
    primes = [q for q in P32 if q > 2] if bits == 32 else P64
    R = 2 ** bits
    lines, want = [], []
    for p in primes:
        rinv = pow(R, -1, p)
        ops = _operands(p, rng)
        for h, x, a, y in itertools.product(ops[:5], ops[3:8], ops[5:10], ops[-5:]):
            lines.append(f"U{bits} {p} {h} {x} {a} {y}")
            want.append((h * x - a * y) * rinv % p)
        for a, b in itertools.product(ops, ops):
            lines.append(f"M{bits} {p} {a} {b}")
            want.append(a * b % p)
        for h in ops:
            if h:
                lines.append(f"I{bits} {p} {h}")
                want.append(pow(h, -1, p))
                lines.append(f"B{bits} {p} {h}")
                want.append(pow(h, -1, p))
                if bits == 32:
                    lines.append(f"E32 {p} {h}")
                    want.append(pow(h, -1, p))
        # the binary inverse's control flow depends on the value: powers of
        # two (long halving runs), values next to p and to its half
        for h in [2 ** e for e in range(1, bits) if 2 ** e < p] + [p - 3, (p - 1) // 2 - 1, (p + 1) // 2 + 1] + [rng.randrange(1, p) for _ in range(40)]:
            if 0 < h < p:
                lines.append(f"B{bits} {p} {h}")
                want.append(pow(h, -1, p))
                if bits == 32:  # (long and short quotient chains alike)
                    lines.append(f"E32 {p} {h}")
                    want.append(pow(h, -1, p))
        if bits == 32:
            lines.append(f"E32 {p} 0")
            want.append(0)
            for h in {1, 2, 3, p - 1, p - 2} | {(p * k) // 13 for k in range(1, 13)}:
                if 0 < h < p:
                    lines.append(f"E32 {p} {h}")
                    want.append(pow(h, -1, p))
    got = harness(lines)
    bad = [(ln, g, w) for ln, g, w in zip(lines, got, want) if int(g) != w]
    assert not bad, bad[:3]


def test_mersenne61_inverse(harness):
    """The addition-chain inverse for p = 2^61 - 1 (the 64-bit inner
    factorisation's one inversion at that prime)."""
    p = 2 ** 61 - 1
    rng = random.Random(61)
    hs = [1, 2, 3, p - 1, p - 2, 2 ** 60, 2 ** 58 + 1] + [2 ** e for e in range(1, 61)] + \
        [rng.randrange(1, p) for _ in range(300)]
    got = harness([f"X61 {h}" for h in hs] + ["X61 0"])
    assert [int(g) for g in got] == [pow(h, -1, p) for h in hs] + [0]


def test_balanced_digits(harness):
    rng = random.Random(3)
    c32 = [0, 1, -1, 127, 128, -128, -129, 2139062143, -2139062143] + [rng.randint(-2139062143, 2139062143)
                                                                       for _ in range(200)]
    c64 = [0, 1, -1, 2 ** 62 - 1, -(2 ** 62 - 1), 2 ** 61, -(2 ** 61)] + [rng.randint(-(2 ** 62) + 1, 2 ** 62 - 1)
                                                                        for _ in range(200)]
    got = harness([f"D32 {c}" for c in c32] + [f"D64 {c}" for c in c64])
    for c, g in zip(c32 + c64, got):
        d = list(map(int, g.split()))
        assert sum(v * 256 ** i for i, v in enumerate(d)) == c, (c, d)
        assert all(-128 <= v <= 127 for v in d), (c, d)
        if len(d) == 8:
            assert -65 <= d[7] <= 65, (c, d)


def test_a_planted_fold_error_is_caught(harness):
    """The check has teeth: a fold that drops the top diagonal disagrees with
    the integer result on these inputs."""
    p = P32[-1]
    s = [0, 0, 0, 0, 0, 0, 1]
    got = int(harness([f"F32 {p} 5 " + " ".join(map(str, s))])[0])
    assert got == _fold_expected(p, 5, s) and got != (5 - 0) % p


def test_mersenne61_products_random(harness):
    """p = 2^61 - 1 takes the Mersenne path (four 32-bit products, a fold at
    bit 61, a rotation for R^-1): many random and near-2^32 operands."""
    p = 2 ** 61 - 1
    rng = random.Random(61)
    near = [2 ** 32 - 1, 2 ** 32, 2 ** 32 + 1, 2 ** 61 - 2, 2 ** 60, 2 ** 29, (2 ** 61 - 1) // 3]
    ops = near + [rng.randrange(p) for _ in range(60)]
    lines, want = [], []
    rinv = pow(2 ** 64, -1, p)
    for _ in range(3000):
        h, x, a, y = (rng.choice(ops) for _ in range(4))
        lines.append(f"U64 {p} {h} {x} {a} {y}")
        want.append((h * x - a * y) * rinv % p)
        lines.append(f"M64 {p} {h} {x}")
        want.append(h * x % p)
    got = harness(lines)
    bad = [(ln, g, w) for ln, g, w in zip(lines, got, want) if int(g) != w]
    assert not bad, bad[:3]
