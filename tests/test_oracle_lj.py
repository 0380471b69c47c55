"""Pins of the LJ part of the CPU oracle (oracle/fc3_oracle.c).

- known-answer values of the reference's own test suite (test/runtests.jl:109-136)
- IFC3 = d IFC2 / dr by central differences of the oracle's IFC2 (which the known
  answers pin), so a dropped term, a wrong sign or a transposed index in the third
  derivative fails
- invariants: ASR (asr.jl semantics), full index-permutation symmetry, linearity in
  epsilon, Psi[iii] = 0 on a centrosymmetric lattice
- the column oracle used at sizes too large for a dense CPU tensor equals the
  dense oracle entry for entry
"""
import math

import numpy as np
import pytest

from fc3 import systems as S
from oracle import fc3_oracle as O

RTOL_JULIA = math.sqrt(np.finfo(np.float64).eps)  # isapprox default rtol


def approx(a, b):
    # Julia's isapprox(a, b) with default rtol = sqrt(eps), atol = 0
    return abs(a - b) <= RTOL_JULIA * max(abs(a), abs(b))


@pytest.fixture(scope="module")
def fcc444():
    sys_ = S.fcc(S.ARGON_A, (4, 4, 4), S.ARGON_MASS)
    calc = S.AnalyticalCalculator(1e-8, S.ARGON_LJ.r_cut)
    return sys_, calc


def test_known_answers_ifc2(fcc444):
    sys_, calc = fcc444
    f2 = O.lj_second_order(sys_, S.ARGON_LJ, calc)
    # runtests.jl:119-125 (1-based indices)
    assert approx(f2[0, 0], 7.727196175)
    assert approx(f2[1, 1], 7.727196175)
    assert approx(f2[2, 2], 7.727196175)
    assert approx(f2[3, 0], -1.02002707)
    assert approx(f2[3, 1], -1.06589524)
    # asr_satisfied (asr.jl:43-60) and issymmetric
    N = sys_.n_atoms
    rows = f2.reshape(3 * N, 3, N, order="F").sum(axis=2)
    assert np.all(rows <= calc.tol)
    assert np.array_equal(f2, f2.T)


@pytest.fixture(scope="module")
def fcc444_ifc3(fcc444):
    sys_, calc = fcc444
    return O.lj_third_order(sys_, S.ARGON_LJ, calc)


def test_known_answers_ifc3(fcc444_ifc3):
    f3 = fcc444_ifc3
    # runtests.jl:128-134
    assert f3[0, 0, 0] == 0.0
    assert approx(f3[3, 0, 0], -3.167167895)
    assert approx(f3[3, 1, 0], -3.979773863)
    assert approx(f3[3, 3, 0], 3.167167895)
    assert approx(f3[4, 4, 0], 3.979773863)
    assert approx(f3[5, 5, 0], -0.406302984)
    assert approx(f3[3, 0, 1], -3.979773863)


def _ifc2(sys_, calc):
    return O.lj_second_order(sys_, S.ARGON_LJ, calc)


def test_ifc3_is_derivative_of_ifc2():
    rng = np.random.default_rng(7)
    sys_ = S.fcc(S.ARGON_A, (2, 2, 2), S.ARGON_MASS, 0.05, seed=3)
    calc = S.AnalyticalCalculator(0.0, S.ARGON_LJ.r_cut)
    psi = O.lj_third_order(sys_, S.ARGON_LJ, calc)
    n = sys_.n_dof
    h = 1e-5
    cols = rng.choice(n, size=12, replace=False)
    for kk in cols:
        a, c = divmod(int(kk), 3)
        p = sys_.positions.copy()
        p[a, c] += h
        fp = _ifc2(S.SuperCellSystem(p, sys_.masses, sys_.box_sizes), calc)
        p[a, c] -= 2 * h
        fm = _ifc2(S.SuperCellSystem(p, sys_.masses, sys_.box_sizes), calc)
        fd = (fp - fm) / (2 * h)
        err = np.abs(fd - psi[:, :, kk]).max()
        assert err < 2e-5 * max(1.0, np.abs(psi[:, :, kk]).max()), (kk, err)


def test_ifc2_is_hessian_of_energy():
    # IFC2 = d2U/dr dr with U the unshifted LJ pair sum (LennardJones.jl:28-31)
    sys_ = S.fcc(S.ARGON_A, (2, 2, 2), S.ARGON_MASS, 0.05, seed=5)
    pot = S.ARGON_LJ
    calc = S.AnalyticalCalculator(0.0, pot.r_cut)
    f2 = _ifc2(sys_, calc)
    L = sys_.box_sizes

    def energy(pos):
        E = 0.0
        N = pos.shape[0]
        for i in range(N):
            r = pos[i] - pos[i + 1:]
            r = np.where(r > L / 2, r - L, np.where(r < -L / 2, r + L, r))
            d2 = (r * r).sum(axis=1)
            d2 = d2[d2 < pot.r_cut ** 2]
            k = (pot.sigma ** 2 / d2) ** 3
            E += (4 * pot.epsilon * k * (k - 1)).sum()
        return E

    h = 1e-4
    p0 = sys_.positions
    for (x, y) in [(0, 0), (0, 4), (5, 17), (30, 31), (40, 2)]:
        a, c = divmod(x, 3)
        b, d = divmod(y, 3)
        vals = []
        for sx in (1, -1):
            for sy in (1, -1):
                p = p0.copy()
                p[a, c] += sx * h
                p[b, d] += sy * h
                vals.append(energy(p))
        fd = (vals[0] - vals[1] - vals[2] + vals[3]) / (4 * h * h)
        assert abs(fd - f2[x, y]) < 1e-4 * max(1.0, abs(f2[x, y])), (x, y, fd, f2[x, y])


def test_ifc3_invariants():
    sys_ = S.fcc(S.ARGON_A, (3, 3, 3), S.ARGON_MASS, 0.04, seed=11)
    calc = S.AnalyticalCalculator(1e-12, S.ARGON_LJ.r_cut)
    psi = O.lj_third_order(sys_, S.ARGON_LJ, calc)
    N = sys_.n_atoms
    n = 3 * N
    scale = np.abs(psi).max()
    # acoustic sum rule on the last index
    s = psi.reshape(n, n, 3, N, order="F").sum(axis=3)
    assert np.abs(s).max() < 1e-12 * scale * 100
    # full permutation symmetry (up to the rounding of the triple product)
    for perm in [(1, 0, 2), (0, 2, 1), (2, 1, 0), (1, 2, 0)]:
        assert np.abs(psi - psi.transpose(perm)).max() < 1e-13 * scale
    # linear in epsilon
    pot2 = S.LJ(S.ARGON_LJ.sigma, 2 * S.ARGON_LJ.epsilon, S.ARGON_LJ.r_cut)
    calc0 = S.AnalyticalCalculator(0.0, S.ARGON_LJ.r_cut)
    p1 = O.lj_third_order(sys_, S.ARGON_LJ, calc0)
    p2 = O.lj_third_order(sys_, pot2, calc0)
    assert np.abs(p2 - 2 * p1).max() < 1e-13 * scale


def test_ifc3_self_zero_on_perfect_lattice(fcc444, fcc444_ifc3):
    # FCC 4x4x4 with r_cut < L/2 is centrosymmetric about every atom
    psi = fcc444_ifc3
    N = fcc444[0].n_atoms
    for a in range(0, N, 7):
        blk = psi[3 * a:3 * a + 3, 3 * a:3 * a + 3, 3 * a:3 * a + 3]
        assert np.abs(blk).max() == 0.0


def test_minimum_image_quirk_fcc333():
    # r_cut = 8.5 > L/2 = 7.87 A: the reference keeps one image per pair
    # (third_order_analytical.jl:7 asserts r_cut < L only), giving 66 neighbours.
    sys_ = S.fcc(S.ARGON_A, (3, 3, 3), S.ARGON_MASS)
    calc = S.AnalyticalCalculator(1e-12, S.ARGON_LJ.r_cut)
    f2 = O.lj_second_order(sys_, S.ARGON_LJ, calc)
    N = sys_.n_atoms
    blocks = np.abs(f2).reshape(3, N, 3, N, order="F").max(axis=(0, 2))
    nbrs = (blocks[0] > 0).sum() - 1
    assert nbrs == 66


def test_cutoff_assertion():
    sys_ = S.fcc(S.ARGON_A, (1, 1, 1), S.ARGON_MASS)
    calc = S.AnalyticalCalculator(1e-12, S.ARGON_LJ.r_cut)
    with pytest.raises(AssertionError):
        O.lj_third_order(sys_, S.ARGON_LJ, calc)


def test_column_oracle_matches_dense():
    sys_ = S.fcc(S.ARGON_A, (2, 2, 2), S.ARGON_MASS, 0.05, seed=9)
    calc = S.AnalyticalCalculator(1e-12, S.ARGON_LJ.r_cut)
    psi = O.lj_third_order(sys_, S.ARGON_LJ, calc)
    n = sys_.n_dof
    rng = np.random.default_rng(1)
    jj = rng.integers(0, n, size=40)
    kk = np.where(rng.random(40) < 0.3, jj, rng.integers(0, n, size=40))
    cols = O.lj_third_order_columns(sys_, S.ARGON_LJ, calc, jj, kk)
    for c in range(len(jj)):
        assert np.array_equal(cols[:, c], psi[:, jj[c], kk[c]])


def test_mass_weight_oracle():
    sys_ = S.fcc(S.ARGON_A, (2, 2, 2), S.ARGON_MASS, 0.05, seed=9)
    calc = S.AnalyticalCalculator(1e-12, S.ARGON_LJ.r_cut)
    psi = O.lj_third_order(sys_, S.ARGON_LJ, calc)
    rng = np.random.default_rng(2)
    m = rng.uniform(10, 50, size=sys_.n_atoms)
    mw = O.mass_weight_third_order(psi.copy(order="F"), m)
    mm = np.repeat(m, 3)
    ref = psi / np.sqrt(mm[:, None, None] * mm[None, :, None] * mm[None, None, :])
    assert np.abs(mw - ref).max() <= 1e-15 * np.abs(ref).max()


def test_ipow_matches_correctly_rounded_pow():
    # The compensated power (Julia's Float64^Int is one) agrees with glibc's pow,
    # which is correctly rounded in practice, except on rare near-ties, and is
    # never more than one ulp away.
    L = O.lib()
    rng = np.random.default_rng(0)
    xs = rng.uniform(2.0, 9.0, size=2000)
    bad = 0
    for x in xs:
        for n in (6, 7, 8, 9, 12, 13, 14, 15):
            a, b = L.fco_ipow(float(x), n), float(x) ** n
            if a != b:
                bad += 1
                assert abs(a - b) <= np.spacing(b)
    assert bad <= 0.002 * len(xs) * 8


def test_coo_oracle_matches_dense():
    for cells, su, seed in [((2, 2, 2), 0.06, 12), ((3, 3, 3), 0.04, 13)]:
        sys_ = S.fcc(S.ARGON_A, cells, S.ARGON_MASS, su, seed=seed)
        calc = S.AnalyticalCalculator(1e-12, S.ARGON_LJ.r_cut)
        psi = O.lj_third_order(sys_, S.ARGON_LJ, calc)
        i, j, k, v = O.lj_third_order_coo(sys_, S.ARGON_LJ, calc)
        dense = np.zeros_like(psi)
        dense[i, j, k] = v
        assert len(v) == np.count_nonzero(psi)
        assert np.array_equal(dense, psi)


def test_threaded_oracle_is_bit_identical_to_one_thread():
    """The pair loop and ASR! are threaded over atoms (OpenMP, as the reference's @tasks
    loops); every entry is still written by one iteration and every ASR sum keeps its
    order, so the tensor must not depend on the thread count."""
    import hashlib
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = ("import sys, hashlib; sys.path[:0] = [%r, %r]\n"
            "from fc3 import systems as S\nfrom oracle import fc3_oracle as O\n"
            "s = S.fcc(S.ARGON_A, (3, 3, 3), S.ARGON_MASS, 0.05, seed=5)\n"
            "c = S.AnalyticalCalculator(1e-12, S.ARGON_LJ.r_cut)\n"
            "p = O.lj_third_order(s, S.ARGON_LJ, c)\n"
            "print(O.lib().fco_num_threads(), hashlib.sha1(p.tobytes(order='F')).hexdigest())\n"
            % (root, os.path.join(root, "forceconstants.jl_amd")))
    out = {}
    for t in ("1", "4"):
        env = dict(os.environ, OMP_NUM_THREADS=t)
        r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                           timeout=300, check=True)
        threads, digest = r.stdout.split()
        assert threads == t
        out[t] = digest
    assert out["1"] == out["4"]


@pytest.mark.parametrize("D", [1, 2, 3])
def test_oracle_mass_weight_dims(D):
    """mass_weight_third_order! for D = n / N_atoms in 1:3 (third_order_analytical.jl:
    185-214): unit masses leave Psi unchanged, equal masses m scale nonzeros by m^-1.5,
    zeros stay zero, and the atom of index ii is ii // D (a per-atom mass pattern)."""
    N = 5
    n = D * N
    rng = np.random.default_rng(D)
    psi = rng.standard_normal((n, n, n))
    psi[np.abs(psi) < 0.5] = 0.0
    psi = np.asfortranarray(psi)
    assert np.array_equal(O.mass_weight_third_order(psi.copy(order="F"), np.ones(N)), psi)
    got = O.mass_weight_third_order(psi.copy(order="F"), np.full(N, 4.0))
    assert np.allclose(got, psi / 8.0, rtol=1e-15, atol=0)
    m = np.array([1.0, 4.0, 9.0, 16.0, 25.0])
    got = O.mass_weight_third_order(psi.copy(order="F"), m)
    idx = np.arange(n) // D
    s = np.sqrt(m)[idx]
    assert np.allclose(got, psi / (s[:, None, None] * s[None, :, None] * s[None, None, :]),
                       rtol=1e-14, atol=0)
    with pytest.raises(AssertionError):
        O.mass_weight_third_order(psi.copy(order="F"), np.ones(N + 1))
