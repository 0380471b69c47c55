"""Device dynamical matrix + normal modes (fc3_modes_lj) against the oracle's host path.

The oracle forms D = IFC2 / sqrt(m_I m_J) from its restatement of the reference's
second_order (dynmat.jl:5-33) and runs get_modes (dynmat.jl:123-131: numpy eigh, then the
3 eigenvalues of smallest magnitude set to zero).  Eigenvectors of a degenerate
eigenvalue are any orthonormal basis of its space, so the vectors are pinned by what
fixes them: the eigen-residual ||D V - V diag(w)||, orthonormality, and agreement up to
sign with the host vectors on eigenvalues separated from their neighbours.
"""
import numpy as np
import pytest

import fc3
from fc3 import systems as S
from oracle import fc3_oracle as O

pytestmark = pytest.mark.gpu

LJ = S.ARGON_LJ


@pytest.mark.parametrize("cells,seed,su", [((3, 3, 3), 61, 0.05), ((4, 4, 4), 62, 0.05),
                                           ((2, 2, 2), 63, 0.0)])
def test_modes_device_match_host(ctx, cells, seed, su):
    sysd = S.fcc(S.ARGON_A, cells, S.ARGON_MASS, su, seed=seed)
    calc = S.AnalyticalCalculator(1e-12, LJ.r_cut)
    n = sysd.n_dof
    D = O.dynamical_matrix(O.lj_second_order(sysd, LJ, calc), sysd.masses)
    w_ref, v_ref = O.get_modes(D)
    lam = np.linalg.eigvalsh(D)
    w, phi = fc3.modes_device(sysd, LJ, calc, fp64=True, ctx=ctx)
    V = phi.to_host()
    scale = float(np.abs(lam).max())
    dw = float(np.max(np.abs(w - w_ref)))
    res = float(np.linalg.norm(D @ V - V * lam) / np.linalg.norm(D))
    orth = float(np.linalg.norm(V.T @ V - np.eye(n)))
    print(f"{cells} n={n}: max|dw| {dw:.2e} (scale {scale:.2e}), residual {res:.2e}, "
          f"orthonormality {orth:.2e}", flush=True)
    assert dw <= 1e-10 * scale
    assert np.array_equal(w == 0.0, w_ref == 0.0) and int((w == 0.0).sum()) == 3
    assert res <= 1e-12 * n and orth <= 1e-12 * n
    # separated eigenvalues: the same vector up to sign
    gap = np.minimum(np.abs(np.diff(lam, prepend=-np.inf)), np.abs(np.diff(lam, append=np.inf)))
    sep = np.nonzero(gap > 1e-6 * scale)[0]
    if len(sep):
        dots = np.abs(np.sum(V[:, sep] * v_ref[:, sep], axis=0))
        print(f"{len(sep)} separated modes: min |<v_dev, v_host>| {dots.min():.12f}", flush=True)
        assert np.all(np.abs(dots - 1.0) <= 1e-8)
    # the FP32 output is the FP64 vectors rounded
    _, phi32 = fc3.modes_device(sysd, LJ, calc, ctx=ctx)
    assert np.max(np.abs(phi32.to_host() - V.astype(np.float32))) <= 1e-6


def test_modes_device_without_rigid_zeroing(ctx):
    sysd = S.fcc(S.ARGON_A, (2, 2, 2), S.ARGON_MASS, 0.04, seed=64)
    calc = S.AnalyticalCalculator(1e-12, LJ.r_cut)
    D = O.dynamical_matrix(O.lj_second_order(sysd, LJ, calc), sysd.masses)
    lam = np.linalg.eigvalsh(D)
    w, _ = fc3.modes_device(sysd, LJ, calc, num_rigid_translation=0, ctx=ctx)
    assert np.all(np.diff(w) >= 0)
    assert np.max(np.abs(w - lam)) <= 1e-10 * np.abs(lam).max()


@pytest.mark.parametrize("cells,su,seed", [((2, 2, 2), 0.03, 65), ((3, 3, 3), 0.03, 648)])
def test_modes_device_sw_match_host(ctx, cells, su, seed):
    pot = S.StillingerWeberSilicon()
    calc = S.AutoDiffCalculator(1e-12, pot.r_cut)
    sysd = S.diamond(S.SI_A, cells, S.SI_MASS, su, seed=seed)
    n = sysd.n_dof
    D = O.dynamical_matrix(O.sw_second_order(sysd, pot, calc), sysd.masses)
    w_ref, _ = O.get_modes(D)
    lam = np.linalg.eigvalsh(D)
    w, phi = fc3.modes_device(sysd, pot, calc, fp64=True, ctx=ctx)
    V = phi.to_host()
    scale = float(np.abs(lam).max())
    dw = float(np.max(np.abs(w - w_ref)))
    res = float(np.linalg.norm(D @ V - V * lam) / np.linalg.norm(D))
    print(f"SW diamond {cells} n={n}: max|dw| {dw:.2e} (scale {scale:.2e}), residual {res:.2e}",
          flush=True)
    assert dw <= 1e-10 * scale
    assert np.array_equal(w == 0.0, w_ref == 0.0)
    assert res <= 1e-12 * n
    assert float(np.linalg.norm(V.T @ V - np.eye(n))) <= 1e-12 * n
