"""The exact fixed-point accumulator of the march kernels' totals (FixSum and
fixed_total in flow_amd/csrc/kernels/reduce_common.h), compiled for the host
and checked against exact rational arithmetic: the 192-bit words equal the
sum of every value truncated toward zero to a multiple of 2^-64, the words do
not depend on the order of the additions, the converted double is within an
ulp of that sum, and a non-finite or too large value marks the total NaN."""
import math
import os
import random
import shutil
import subprocess
from fractions import Fraction

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "hostcheck", "fixsum_host.cpp")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def fixsum(tmp_path_factory):
    if not shutil.which(HIPCC) and not os.path.exists(HIPCC):
        pytest.skip("hipcc not found")
    exe = str(tmp_path_factory.mktemp("fixsum") / "fixsum_host")
    subprocess.run([HIPCC, "-O2", "-std=c++17", SRC, "-o", exe], check=True,
                   capture_output=True, timeout=600)

    def run(values):
        out = subprocess.run([exe], input="".join(f"{v.hex()}\n" for v in values),
                             capture_output=True, text=True, check=True, timeout=60)
        w0, w1, w2, bad, total = out.stdout.split()
        return int(w0), int(w1), int(w2), int(bad), float.fromhex(total)
    return run


def _exact(values):
    """Sum of every value truncated toward zero to a multiple of 2^-64, in
    units of 2^-64."""
    return sum(int(Fraction(v) * (1 << 64)) for v in values)


def _words(n):
    n &= (1 << 192) - 1
    return n & ((1 << 64) - 1), (n >> 64) & ((1 << 64) - 1), n >> 128


def _values(rng, count):
    out = []
    for _ in range(count):
        kind = rng.randrange(6)
        x = rng.uniform(-1.0, 1.0)
        if kind == 0:
            out.append(x * 2.0 ** rng.randrange(-80, 120))
        elif kind == 1:
            out.append(0.0 if rng.random() < 0.5 else -0.0)
        elif kind == 2:
            out.append(abs(x) * 1.0e8)          # a density total of a chunk
        elif kind == 3:
            out.append(x * 1.0e-30)             # below the 2^-64 unit
        else:
            out.append(x * 10.0 ** rng.randrange(-6, 12))
    return out


@pytest.mark.parametrize("seed", range(8))
def test_fixsum_words_are_the_exact_truncated_sum(fixsum, seed):
    rng = random.Random(seed)
    values = _values(rng, 400)
    w0, w1, w2, bad, total = fixsum(values)
    exact = _exact(values)
    assert bad == 0
    assert (w0, w1, w2) == _words(exact)
    ref = Fraction(exact, 1 << 64)
    if ref != 0:
        assert abs(Fraction(total) - ref) <= Fraction(math.ulp(float(ref)))
    else:
        assert total == 0.0


def test_fixsum_is_order_independent(fixsum):
    rng = random.Random(99)
    values = _values(rng, 300)
    first = fixsum(values)
    for _ in range(3):
        rng.shuffle(values)
        assert fixsum(values) == first


def test_fixsum_cancellation_and_negative_totals(fixsum):
    big = 2.0 ** 100
    assert fixsum([big, 1.0, -big])[:3] == _words(1 << 64)
    w = fixsum([-3.5, 1.25])
    assert w[:3] == _words(int(Fraction(-2.25) * (1 << 64)))
    assert w[4] == -2.25


@pytest.mark.parametrize("bad_value", [float("nan"), float("inf"), -float("inf"), 2.0 ** 127])
def test_fixsum_marks_out_of_range_values(fixsum, bad_value):
    w0, w1, w2, bad, total = fixsum([1.0, bad_value, 2.0])
    assert bad == 1
    assert math.isnan(total)
