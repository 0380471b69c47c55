"""The text format of save_second_order / save_third_order (src/io.jl:9-62), CPU side.

Pins: Julia's printed form of specific Float64 values (known answers of Julia's
print, e.g. 1.0e-5, 0.0001, 100000.0, 1.0e6), agreement of the library's formatter
(fc3_format_float, host only) with the oracle's independent restatement from Python's
shortest digits on random doubles, exact read-back of every printed value, and the
oracle's line order against a brute-force transliteration of the reference's six
nested loops on a tiny tensor.
"""
import itertools

import numpy as np
import pytest

from oracle import fc3_oracle as O

KNOWN = [
    (0.0001, "0.0001"), (1e-5, "1.0e-5"), (1.234e-5, "1.234e-5"), (100000.0, "100000.0"),
    (1e6, "1.0e6"), (123456.0, "123456.0"), (1234567.0, "1.234567e6"), (0.1, "0.1"),
    (-2.5, "-2.5"), (1.0, "1.0"), (0.0, "0.0"), (-0.0, "-0.0"), (100.0, "100.0"),
    (0.30000000000000004, "0.30000000000000004"), (5e-324, "5.0e-324"),
    (1.7976931348623157e308, "1.7976931348623157e308"), (1e16, "1.0e16"),
    (float("nan"), "NaN"), (float("inf"), "Inf"), (-float("inf"), "-Inf"),
    (-0.00805, "-0.00805"),
]


@pytest.mark.parametrize("x,s", KNOWN)
def test_oracle_prints_known_julia_strings(x, s):
    assert O.julia_float_str(x) == s


@pytest.mark.parametrize("x,s", KNOWN)
def test_library_prints_known_julia_strings(x, s):
    import fc3
    assert fc3.format_float(x) == s


def test_library_and_oracle_agree_and_read_back():
    import fc3
    rng = np.random.default_rng(7)
    xs = rng.standard_normal(20000) * 10.0 ** rng.integers(-40, 40, 20000)
    xs = np.concatenate([xs, rng.standard_normal(2000) * 1e-3, np.ldexp(1.0, np.arange(-1074, 1024, 37))])
    for x in xs:
        a = fc3.format_float(float(x))
        assert a == O.julia_float_str(float(x))
        assert float(a) == float(x)


def _brute_lines(v, N, D):
    """The reference's loops (io.jl:16-27, 46-61), transliterated for tiny tensors."""
    out = []
    if v.ndim == 2:
        for i, j, a, b in itertools.product(range(N), range(N), range(D), range(D)):
            x = v[D * i + a, D * j + b]
            if x != 0.0:
                out.append(f"{i + 1} {a + 1} {j + 1} {b + 1} {O.julia_float_str(float(x))}")
        return out
    for i, j, k, a, b, c in itertools.product(range(N), range(N), range(N), range(D), range(D),
                                              range(D)):
        x = v[D * i + a, D * j + b, D * k + c]
        if x != 0.0:
            out.append(f"{i + 1} {a + 1} {j + 1} {b + 1} {k + 1} {c + 1} "
                       f"{O.julia_float_str(float(x))}")
    return out


@pytest.mark.parametrize("D", [1, 2, 3])
@pytest.mark.parametrize("order", [2, 3])
def test_oracle_line_order_matches_reference_loops(D, order):
    N = 3
    rng = np.random.default_rng(10 * D + order)
    v = rng.standard_normal((N * D,) * order)
    v[np.abs(v) < 0.9] = 0.0
    assert O.save_txt_lines(v, N, D) == _brute_lines(v, N, D)
