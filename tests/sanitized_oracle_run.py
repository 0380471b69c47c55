"""Runs every C entry point of the CPU oracle once on small systems.

tests/test_oracle_sanitized.py starts this script in a child process, with FC3_ORACLE_LIB
naming an AddressSanitizer + UndefinedBehaviorSanitizer build of oracle/fc3_oracle.c and
the ASan runtime preloaded. Any out-of-bounds access, use after free or undefined
arithmetic in the oracle ends the child with a sanitizer report. The values are checked
only lightly here; the other oracle tests pin them.
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, ".."))
sys.path.insert(0, os.path.join(HERE, "..", "forceconstants.jl_amd"))

from fc3 import systems as S  # noqa: E402
from oracle import fc3_oracle as O  # noqa: E402


def plant_overflow() -> None:
    """Negative control: mass weighting of N = 2, D = 3 touches 216 doubles of an array of
    8, so AddressSanitizer must stop the process with a heap-buffer-overflow report."""
    psi = np.zeros(8)
    O.lib().fco_mass_weight3_d(O._p(psi), 2, 3, O._p(np.ones(2)))
    print("planted overflow not detected", flush=True)


def main() -> None:
    assert os.environ.get("FC3_ORACLE_LIB"), "run by tests/test_oracle_sanitized.py"
    if "--plant" in sys.argv:
        plant_overflow()
        return
    lib = O.lib()
    ipow = lib.fco_ipow
    ipow.argtypes = [ctypes.c_double, ctypes.c_int]
    ipow.restype = ctypes.c_double
    assert ipow(1.5, 13) == 1.5 ** 13

    # LJ: FCC 2x2x2 argon, displaced, cutoff inside L/2
    lj_sys = S.fcc(S.ARGON_A, (2, 2, 2), S.ARGON_MASS, 0.05, seed=96)
    calc = S.AnalyticalCalculator(1e-12, 5.0)
    f2 = O.lj_second_order(lj_sys, S.ARGON_LJ, calc)
    f3 = O.lj_third_order(lj_sys, S.ARGON_LJ, calc)
    n = 3 * lj_sys.n_atoms
    assert f2.shape == (n, n) and f3.shape == (n, n, n)
    jj = np.array([0, 5, n - 1], np.int64)
    kk = np.array([n - 1, 7, 0], np.int64)
    cols = O.lj_third_order_columns(lj_sys, S.ARGON_LJ, calc, jj, kk)
    assert np.array_equal(cols, f3[:, jj, kk])
    coo = O.lj_third_order_coo(lj_sys, S.ARGON_LJ, calc)
    assert len(coo[0]) == np.count_nonzero(f3)
    # the COO writer's overflow return (-2) with a capacity far too small
    i32 = ctypes.POINTER(ctypes.c_int32)
    small = [np.empty(4, np.int32) for _ in range(3)]
    sv = np.empty(4, np.float64)
    pos, box = O._sysargs(lj_sys)
    rc = lib.fco_lj_ifc3_coo(O._p(pos), lj_sys.n_atoms, O._p(box), S.ARGON_LJ.sigma,
                             S.ARGON_LJ.epsilon, S.ARGON_LJ.r_cut, calc.r_cut, calc.tol, 4,
                             *[a.ctypes.data_as(i32) for a in small], O._p(sv))
    assert rc == -2
    for name in ("sigma", "epsilon"):
        d = O.lj_third_order_dparam(lj_sys, S.ARGON_LJ, calc, name)
        assert d.shape == f3.shape and np.all(np.isfinite(d))
    masses = np.linspace(20.0, 40.0, lj_sys.n_atoms)
    w = O.mass_weight_third_order(f3, masses)
    assert w.shape == f3.shape
    for D in (1, 2, 3):
        N = n // D
        psi = np.asfortranarray(np.random.default_rng(D).standard_normal((D * N,) * 3))
        lib.fco_mass_weight3_d(O._p(psi), N, D, O._p(np.linspace(1.0, 2.0, N)))
    # the IFC3 cutoff assertion (potential cutoff at least the box) returns before writing
    far = S.LJ(sigma=S.ARGON_LJ.sigma, epsilon=S.ARGON_LJ.epsilon, r_cut=11.0)
    try:
        O.lj_third_order(lj_sys, far, calc)
        raise SystemExit("potential cutoff beyond the box accepted")
    except AssertionError:
        pass

    # SW: diamond 2x2x2 silicon, displaced
    sw_sys = S.diamond(S.SI_A, (2, 2, 2), S.SI_MASS, 0.03, seed=192)
    pot = S.StillingerWeberSilicon()
    sw_calc = S.AutoDiffCalculator(1e-12, pot.r_cut)
    r9 = np.array([0.1, 0.2, 0.3, 2.3, 0.1, -0.2, -0.4, 2.2, 0.5])
    assert np.all(np.isfinite(O.sw_blocks3(r9, pot.params)))
    assert np.all(np.isfinite(O.sw_blocks2(r9, pot.params)))
    h2, h3 = O.sw_pair_blocks(np.array([2.2, 0.3, -0.1]), pot.params)
    assert np.all(np.isfinite(h2)) and np.all(np.isfinite(h3))
    assert np.isfinite(O.sw_energy(sw_sys, pot.params, pot.r_cut))
    g2 = O.sw_second_order(sw_sys, pot, sw_calc)
    g3 = O.sw_third_order(sw_sys, pot, sw_calc)
    assert g2.shape == (3 * sw_sys.n_atoms,) * 2 and np.all(np.isfinite(g3))
    for name in O.SW_PARAMS:
        d = O.sw_third_order_dparam(sw_sys, pot, sw_calc, name)
        assert d.shape == g3.shape and np.all(np.isfinite(d))
    print("oracle entry points ran clean", flush=True)


if __name__ == "__main__":
    main()
