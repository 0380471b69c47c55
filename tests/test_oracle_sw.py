"""Pins of the Stillinger-Weber part of the CPU oracle.

- the reference's own AutoDiffCalculator blocks (FastDifferentiation output
  evaluated as text, tests/golden/make_sw_fixtures.py) at 20 seeded geometries
- independent sympy/mpmath derivatives (40 digits) with the per-block cos(theta0)
  constants (-0.333333333333 for the generated blocks, -1/3 for iik)
- FP64 central differences of the SW energy (IFC2 = Hessian, IFC3 = d IFC2/dr)
- invariants: ASR and permutation symmetry of the assembled tensors
"""
import os

import numpy as np
import pytest

from fc3 import systems as S
from oracle import fc3_oracle as O

FIX = np.load(os.path.join(os.path.dirname(__file__), "golden", "sw_blocks.npz"))
P = S.SWParams()


def test_fixture_sources_agree():
    # the two fixture sources (reference blocks, sympy) agree with each other
    m = ~np.isnan(FIX["ref3"])
    assert np.abs(FIX["ref3"][m] - FIX["sym3"][m]).max() < 1e-13
    assert np.abs(FIX["ref2"] - FIX["sym2"]).max() < 1e-13
    assert np.abs(FIX["refp3"] - FIX["symp3"]).max() < 1e-12


def test_oracle_three_body_blocks_match_reference_and_sympy():
    for t, r9 in enumerate(FIX["r9"]):
        blk = O.sw_blocks3(r9, P)
        for b in range(7):
            sym = FIX["sym3"][t, b]
            assert np.abs(blk[b] - sym).max() <= 1e-12 * max(1.0, np.abs(sym).max()), (t, b)
            if b != 2:   # iik has no straight-line reference form
                ref = FIX["ref3"][t, b]
                assert np.abs(blk[b] - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()), (t, b)
        h2 = O.sw_blocks2(r9, P)
        assert np.abs(h2 - FIX["ref2"][t]).max() <= 1e-12


def test_iik_constant_matters():
    # cos(theta0) differs between the iik block (-1/3) and the others
    # (-0.333333333333): using one value for all would move iik by ~1e-12 relative
    r9 = FIX["r9"][0]
    a = O.sw_blocks3(r9, P, c0_iik=O.SW_C0_IIK)[2]
    b = O.sw_blocks3(r9, P, c0_iik=O.SW_C0_FD)[2]
    assert np.abs(a - b).max() > 0
    assert np.abs(a - FIX["sym3"][0, 2]).max() < np.abs(b - FIX["sym3"][0, 2]).max()


def test_oracle_pair_blocks():
    for t, r3 in enumerate(FIX["r3"]):
        h2, h3 = O.sw_pair_blocks(r3, P)
        assert np.abs(h3 - FIX["refp3"][t]).max() <= 1e-12 * max(1.0, np.abs(FIX["refp3"][t]).max())
        assert np.abs(h2 - FIX["refp2"][t]).max() <= 1e-12 * max(1.0, np.abs(FIX["refp2"][t]).max())


@pytest.fixture(scope="module")
def small_si():
    sysd = S.diamond(S.SI_A, (2, 2, 2), S.SI_MASS, 0.03, seed=648)
    pot = S.StillingerWeberSilicon()
    calc = S.AutoDiffCalculator(0.0, pot.r_cut)
    return sysd, pot, calc


def test_sw_ifc2_is_hessian_of_energy(small_si):
    sysd, pot, calc = small_si
    f2 = O.sw_second_order(sysd, pot, calc)
    h = 1e-4
    p0 = sysd.positions
    rng = np.random.default_rng(3)
    for _ in range(8):
        x, y = rng.integers(0, sysd.n_dof, size=2)
        a, c = divmod(int(x), 3)
        b, d = divmod(int(y), 3)
        vals = []
        for sx in (1, -1):
            for sy in (1, -1):
                p = p0.copy()
                p[a, c] += sx * h
                p[b, d] += sy * h
                vals.append(O.sw_energy(S.SuperCellSystem(p, sysd.masses, sysd.box_sizes), P,
                                        calc.r_cut))
        fd = (vals[0] - vals[1] - vals[2] + vals[3]) / (4 * h * h)
        assert abs(fd - f2[x, y]) < 2e-5 * max(1.0, abs(f2[x, y])), (x, y, fd, f2[x, y])


def test_sw_ifc3_is_derivative_of_ifc2(small_si):
    sysd, pot, calc = small_si
    psi = O.sw_third_order(sysd, pot, calc)
    h = 1e-5
    rng = np.random.default_rng(4)
    for kk in rng.choice(sysd.n_dof, size=6, replace=False):
        a, c = divmod(int(kk), 3)
        p = sysd.positions.copy()
        p[a, c] += h
        fp = O.sw_second_order(S.SuperCellSystem(p, sysd.masses, sysd.box_sizes), pot, calc)
        p[a, c] -= 2 * h
        fm = O.sw_second_order(S.SuperCellSystem(p, sysd.masses, sysd.box_sizes), pot, calc)
        fd = (fp - fm) / (2 * h)
        scale = max(1.0, np.abs(psi[:, :, kk]).max())
        assert np.abs(fd - psi[:, :, kk]).max() < 1e-5 * scale, kk


def test_sw_ifc3_invariants(small_si):
    sysd, pot, calc = small_si
    psi = O.sw_third_order(sysd, pot, calc)
    N, n = sysd.n_atoms, sysd.n_dof
    scale = np.abs(psi).max()
    s = psi.reshape(n, n, 3, N, order="F").sum(axis=3)
    assert np.abs(s).max() < 1e-11 * scale
    for perm in [(1, 0, 2), (0, 2, 1), (2, 1, 0)]:
        # symmetric up to rounding and the two cos(theta0) constants
        assert np.abs(psi - psi.transpose(perm)).max() < 1e-10 * scale
    f2 = O.sw_second_order(sysd, pot, calc)
    assert np.abs(f2 - f2.T).max() < 1e-12 * np.abs(f2).max()
    r = f2.reshape(n, 3, N, order="F").sum(axis=2)
    assert np.abs(r).max() < 1e-11 * np.abs(f2).max()
