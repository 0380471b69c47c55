"""Stillinger-Weber force constants on the GPU against the CPU oracle.

The oracle follows the reference's AutoDiffCalculator loops with hyper-dual
derivative blocks (pinned to the reference's generated blocks and to sympy in
test_oracle_sw.py); the GPU evaluates hand-derived closed forms.  Bar: 1e-12
absolute per entry (FP64), IFC2 and IFC3, with the reference's per-block
cos(theta0) constants; FP32 mass-weighted output equal to Float32 of the oracle's
mass-weighted values up to one FP32 rounding step.
"""
import numpy as np
import pytest

import fc3
from fc3 import systems as S
from oracle import fc3_oracle as O

pytestmark = pytest.mark.gpu

POT = S.StillingerWeberSilicon()


def _calc(tol=1e-12):
    return S.AutoDiffCalculator(tol, POT.r_cut)


@pytest.mark.parametrize("cells,su,seed", [((2, 2, 2), 0.0, 0), ((2, 2, 2), 0.03, 61),
                                           ((3, 3, 3), 0.03, 648)])
def test_sw_ifc2_matches_oracle(ctx, cells, su, seed):
    sysd = S.diamond(S.SI_A, cells, S.SI_MASS, su, seed=seed)
    g = fc3.second_order(sysd, POT, _calc(), ctx=ctx).values
    r = O.sw_second_order(sysd, POT, _calc())
    d = np.abs(g - r)
    print(f"sw ifc2 n={sysd.n_dof}: max {d.max():.3e} fro {np.sqrt((d**2).sum()):.3e}", flush=True)
    assert d.max() <= 1e-12


@pytest.mark.parametrize("cells,su,seed", [((2, 2, 2), 0.0, 0), ((2, 2, 2), 0.03, 62),
                                           ((3, 3, 3), 0.03, 648)])
def test_sw_ifc3_matches_oracle(ctx, cells, su, seed):
    sysd = S.diamond(S.SI_A, cells, S.SI_MASS, su, seed=seed)
    print(f"case diamond {cells} sigma_u={su} n={sysd.n_dof}", flush=True)
    g = fc3.third_order(sysd, POT, _calc(), ctx=ctx).values
    r = O.sw_third_order(sysd, POT, _calc())
    d = np.abs(g - r)
    print(f"sw ifc3 n={sysd.n_dof}: max {d.max():.3e} fro {np.sqrt((d**2).sum()):.3e} "
          f"nnz gpu {np.count_nonzero(g)} oracle {np.count_nonzero(r)}", flush=True)
    assert d.max() <= 1e-12


def test_sw_iik_constant_is_reproduced(ctx):
    # With one cos(theta0) for every block the result moves by ~1e-12: the GPU must
    # follow the reference's two constants, not the self-consistent potential.
    sysd = S.diamond(S.SI_A, (2, 2, 2), S.SI_MASS, 0.03, seed=63)
    g = fc3.third_order(sysd, POT, _calc(0.0), ctx=ctx).values
    ref = O.sw_third_order(sysd, POT, _calc(0.0))
    alt = O.sw_third_order(sysd, POT, _calc(0.0), c0_fd=O.SW_C0_IIK)
    assert np.abs(g - ref).max() < np.abs(g - alt).max()


def test_sw_fp32_mass_weighted(ctx):
    sysd = S.diamond(S.SI_A, (2, 2, 2), S.SI_MASS, 0.03, seed=64)
    rng = np.random.default_rng(5)
    m = rng.uniform(20.0, 40.0, size=sysd.n_atoms)
    n = sysd.n_dof
    d64 = fc3.DeviceArray((n, n, n), np.float64, ctx)
    d32 = fc3.DeviceArray((n, n, n), np.float32, ctx)
    fc3.third_order_device(sysd, POT, _calc(), psi64=d64, psi32=d32, masses=m, ctx=ctx)
    g64, g32 = d64.to_host(), d32.to_host()
    r = O.sw_third_order(sysd, POT, _calc())
    assert np.abs(g64 - r).max() <= 1e-12
    # the FP32 copy is Float32(g64 / sqrt(m m m)) exactly
    w = g64.copy(order="F")
    O.mass_weight_third_order(w, m)
    assert np.array_equal(g32, w.astype(np.float32))


def test_sw_cutoff_assertion(ctx):
    sysd = S.diamond(S.SI_A, (2, 2, 2), S.SI_MASS)
    with pytest.raises(AssertionError):
        fc3.third_order(sysd, POT, S.AutoDiffCalculator(1e-12, POT.r_cut + 0.1), ctx=ctx)
