"""The oracle's asr_satisfied / issymmetric (asr.jl:43-74, ifc.jl:15) pinned on the CPU.

Pins that do not restate the formula: planted entries whose effect on the row sums is
known in closed form (a +x entry makes exactly its own (slice, row, b) set fail, a -x
entry makes none fail, which fixes the one-sided `s > tol`), a tolerance below every
sum (every set fails: the count fixes the loop extents), and the invariant that a
tensor straight out of the reference's ASR! satisfies the rule.
"""
import numpy as np
import pytest

from fc3 import systems as S
from oracle import fc3_oracle as O

CALC = S.AnalyticalCalculator(1e-12, S.ARGON_LJ.r_cut)


def _planted(order, N, D, seed, n_plants=5):
    n = N * D
    rng = np.random.default_rng(seed)
    base = np.zeros((n,) * order)
    pos, sets = [], set()
    while len(pos) < n_plants:
        p = tuple(int(x) for x in rng.integers(0, n, order))
        key = (p[2] if order == 3 else 0, p[0], p[1] % D)
        if key not in sets:          # one plant per (slice, row, b) set
            sets.add(key)
            pos.append(p)
    expect = set()
    for t, p in enumerate(pos):
        sign = 1.0 if t % 2 == 0 else -1.0
        base[p] += sign * 0.5
        row, col = p[0], p[1]
        sl = p[2] if order == 3 else 0
        if sign > 0:
            expect.add((sl * n + row) * D + col % D)
    return base, pos, expect


@pytest.mark.parametrize("order", [2, 3])
@pytest.mark.parametrize("D", [1, 2, 3])
def test_planted_entries(order, D):
    N = 4
    t, pos, expect = _planted(order, N, D, seed=10 * order + D, n_plants=min(5, N * D * D))
    got = O.asr_violations(t, N, D, 1e-8)
    assert {k for k, _ in got} == expect
    for k, s in got:
        assert s == 0.5
    assert [k for k, _ in got] == sorted(k for k, _ in got)    # reference loop order


@pytest.mark.parametrize("order", [2, 3])
def test_every_set_fails_below_all_sums(order):
    N, D = 3, 3
    n = N * D
    t = np.zeros((n,) * order)
    got = O.asr_violations(t, N, D, -1.0)
    assert len(got) == (n if order == 3 else 1) * n * D


def test_reference_asr_output_satisfies():
    sysd = S.fcc(S.ARGON_A, (2, 2, 2), S.ARGON_MASS, 0.05, seed=3)
    psi = O.lj_third_order(sysd, S.ARGON_LJ, CALC)
    assert O.asr_violations(psi, sysd.n_atoms, 3, 1e-12) == []
    f2 = O.lj_second_order(sysd, S.ARGON_LJ, CALC)
    assert O.asr_violations(f2, sysd.n_atoms, 3, 1e-12) == []
    # the same tensor with one first-index row shifted by +1e-6 on a single column
    psi[5, 7, 11] += 1e-6
    got = O.asr_violations(psi, sysd.n_atoms, 3, 1e-12)
    assert [k for k, _ in got] == [(11 * psi.shape[0] + 5) * 3 + 7 % 3]


def test_issymmetric_count():
    rng = np.random.default_rng(4)
    a = rng.standard_normal((7, 7))
    a = a + a.T
    assert O.issymmetric_count(a) == 0
    a[1, 4] += 1.0
    a[6, 2] = np.nan
    a[2, 6] = np.nan                   # NaN != NaN: still asymmetric
    assert O.issymmetric_count(a) == 2
