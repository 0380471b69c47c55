"""Fused third_order + mcc3 without the dense tensor on the GPU: LJ
(fc3_third_order_mcc3_lj) and Stillinger-Weber (fc3_third_order_mcc3_sw).

Pins:
  * against the FP64 oracle: K3 = mcc3 of the oracle's mass-weighted Psi (the reference
    loop restated, then mass_weight_third_order!), contracted in FP64 from the same FP32
    Psi and phi, relative Frobenius error <= 1e-5 (the BF16x3 bar);
  * against the library's dense path (fc3_third_order_lj into psi32, then fc3_mcc3 in
    BF16x3) on the same inputs, <= 1e-5: the two differ only in how the mode-1 product
    is evaluated (over the nonzeros in FP32, or densely in BF16x3).
Sizes cover K padding (n = 324 -> 336) and none (n = 96, 768), with and without
masses, and a tolerance that zeroes entries (1e-8, as the reference's "LJ Values" test).
The SW cases hold the same two bars on silicon diamond 2x2x2 and 3x3x3 (192 and 648 DoF),
with the tolerance 1e-12 and 0.1 (the reference's SW test value, which zeroes entries).
"""
import numpy as np
import pytest

import fc3
from fc3 import systems as S
from fc3._lib import FC3Error
from oracle import fc3_oracle as O

pytestmark = pytest.mark.gpu

LJ = S.ARGON_LJ
SW = S.StillingerWeberSilicon()


def _rel(a, b):
    return float(np.linalg.norm((a.astype(np.float64) - b).ravel()) / np.linalg.norm(b.ravel()))


def _case(ctx, cells, seed, su, tol):
    sysd = S.fcc(S.ARGON_A, cells, S.ARGON_MASS, su, seed=seed)
    calc = S.AnalyticalCalculator(tol, LJ.r_cut)
    f2 = O.lj_second_order(sysd, LJ, calc)
    _, phi = O.get_modes(O.dynamical_matrix(f2, sysd.masses))
    phi32 = np.asfortranarray(phi.astype(np.float32))
    return sysd, calc, phi32, fc3.DeviceArray.from_host(phi32, ctx=ctx)


@pytest.mark.parametrize("cells,seed,su,tol,weighted", [
    ((3, 3, 3), 41, 0.05, 1e-12, True),
    ((2, 2, 2), 42, 0.08, 1e-8, False),
    ((2, 2, 2), 43, 0.03, 1e-12, True),
])
def test_fused_matches_fp64_oracle(ctx, cells, seed, su, tol, weighted):
    sysd, calc, phi32, dphi = _case(ctx, cells, seed, su, tol)
    masses = sysd.masses if weighted else None
    k3 = fc3.third_order_mcc3(sysd, LJ, calc, dphi, masses=masses, ctx=ctx).to_host()
    psi = O.lj_third_order(sysd, LJ, calc)
    if weighted:
        O.mass_weight_third_order(psi, sysd.masses)
    ref = O.mcc3_fp64(psi.astype(np.float32), phi32)
    err = _rel(k3, ref)
    print(f"{cells} seed {seed} tol {tol} weighted {weighted}: rel fro err vs FP64 {err:.3e}",
          flush=True)
    assert err <= 1e-5


@pytest.mark.parametrize("cells,seed", [((3, 3, 3), 46), ((4, 4, 4), 47), ((5, 5, 5), 48)])
def test_fused_direct_mode3_equals_split_path(ctx, cells, seed, monkeypatch):
    """Mode 3 of the fused path reads the FP32 intermediate directly (the bf16 hi/lo split
    inside the GEMM) when the workspace holds it; FC3_FUSED_SPLIT=1 forces the split pass
    into a bf16 scratch first.  Both multiply the same operands in the same order: K3 must
    agree bit for bit (324 and 1500 DoF have K padding, 768 none)."""
    sysd, calc, phi32, dphi = _case(ctx, cells, seed, 0.05, 1e-12)
    k_direct = fc3.third_order_mcc3(sysd, LJ, calc, dphi, masses=sysd.masses, ctx=ctx).to_host()
    monkeypatch.setenv("FC3_FUSED_SPLIT", "1")
    k_split = fc3.third_order_mcc3(sysd, LJ, calc, dphi, masses=sysd.masses, ctx=ctx).to_host()
    monkeypatch.delenv("FC3_FUSED_SPLIT")
    diff = int(np.count_nonzero(k_direct.view(np.uint32) != k_split.view(np.uint32)))
    print(f"fused {cells}: direct vs split mode 3: {diff} entries differ", flush=True)
    assert diff == 0
    dphi.free()


@pytest.mark.parametrize("cells,seed", [((3, 3, 3), 44), ((4, 4, 4), 45)])
def test_fused_matches_dense_path(ctx, cells, seed):
    sysd, calc, phi32, dphi = _case(ctx, cells, seed, 0.05, 1e-12)
    n = sysd.n_dof
    k3f = fc3.third_order_mcc3(sysd, LJ, calc, dphi, masses=sysd.masses, ctx=ctx).to_host()
    psi32 = fc3.DeviceArray((n, n, n), np.float32, ctx)
    fc3.third_order_device(sysd, LJ, calc, psi32=psi32, masses=sysd.masses, ctx=ctx)
    k3d = fc3.mcc3(psi32, dphi, precision="bf16x3").to_host()
    err = _rel(k3f, k3d.astype(np.float64))
    print(f"{cells} (n={n}): fused vs dense BF16x3 rel {err:.3e}", flush=True)
    assert err <= 1e-5


def test_fused_global_coefficients_bitwise(ctx, monkeypatch):
    """Neighbourhoods whose coefficients do not fit in LDS read them from global scratch
    (FC3_LJ_FUSED_GLOBAL=1 forces that kernel): the same values multiplied in the same
    order, so K3 must agree bit for bit."""
    sysd, calc, phi32, dphi = _case(ctx, (3, 3, 3), 49, 0.05, 1e-12)
    k_lds = fc3.third_order_mcc3(sysd, LJ, calc, dphi, masses=sysd.masses, ctx=ctx).to_host()
    monkeypatch.setenv("FC3_LJ_FUSED_GLOBAL", "1")
    k_glb = fc3.third_order_mcc3(sysd, LJ, calc, dphi, masses=sysd.masses, ctx=ctx).to_host()
    monkeypatch.delenv("FC3_LJ_FUSED_GLOBAL")
    diff = int(np.count_nonzero(k_lds.view(np.uint32) != k_glb.view(np.uint32)))
    print(f"LJ fused LDS vs global coefficients: {diff} entries differ", flush=True)
    assert diff == 0
    dphi.free()


def test_fused_long_cutoff(ctx):
    """r_cut = 12.5 A on FCC 4x4x4 with a = 4.4 A (768 DoF): 236 neighbours per atom, past
    the 199 whose coefficients fit in LDS, so mode 1 reads them from global scratch.
    Against the FP64 oracle on sampled slabs and against the dense route on the whole K3."""
    pot = S.LJ(sigma=3.4, epsilon=0.24037, r_cut=12.5)
    sysd = S.fcc(4.4, (4, 4, 4), S.ARGON_MASS, 0.05, seed=50)
    calc = S.AnalyticalCalculator(1e-12, pot.r_cut)
    n = sysd.n_dof
    f2 = O.lj_second_order(sysd, pot, calc)
    _, phi = O.get_modes(O.dynamical_matrix(f2, sysd.masses))
    phi32 = np.asfortranarray(phi.astype(np.float32))
    dphi = fc3.DeviceArray.from_host(phi32, ctx=ctx)
    k3 = fc3.third_order_mcc3(sysd, pot, calc, dphi, masses=sysd.masses, ctx=ctx).to_host()
    coo = O.lj_third_order_coo(sysd, pot, calc)
    v32 = O.coo_mass_weight(coo, sysd.masses).astype(np.float32).astype(np.float64)
    coo32 = (coo[0], coo[1], coo[2], v32)
    num = den = 0.0
    for l in (5, 200, 401, 767):
        ref = O.mcc3_fp64_slab(coo32, phi32.astype(np.float64), l)
        num += float(np.sum((k3[:, :, l].astype(np.float64) - ref) ** 2))
        den += float(np.sum(ref ** 2))
    err = np.sqrt(num / den)
    psi32 = fc3.DeviceArray((n, n, n), np.float32, ctx)
    fc3.third_order_device(sysd, pot, calc, psi32=psi32, masses=sysd.masses, ctx=ctx)
    k3d = fc3.mcc3(psi32, dphi, precision="bf16x3").to_host()
    psi32.free()
    errd = _rel(k3, k3d.astype(np.float64))
    print(f"LJ r_cut 12.5 (n={n}): rel fro err vs FP64 slabs {err:.3e}, vs dense {errd:.3e}",
          flush=True)
    assert err <= 1e-5 and errd <= 1e-5
    dphi.free()


def test_fused_rejects_unsupported_size(ctx):
    full = S.fcc(S.ARGON_A, (2, 2, 2), S.ARGON_MASS, 0.0, seed=1)
    sysd = S.SuperCellSystem(full.positions[:31], full.masses[:31], full.box_sizes)   # 3N = 93
    calc = S.AnalyticalCalculator(1e-12, LJ.r_cut)
    dphi = fc3.DeviceArray.from_host(np.eye(93, dtype=np.float32, order="F"), ctx=ctx)
    with pytest.raises(FC3Error):
        fc3.third_order_mcc3(sysd, LJ, calc, dphi, ctx=ctx)


def _sw_case(ctx, cells, seed, su, tol):
    sysd = S.diamond(S.SI_A, cells, S.SI_MASS, su, seed=seed)
    calc = S.AutoDiffCalculator(tol, SW.r_cut)
    f2 = O.sw_second_order(sysd, SW, S.AutoDiffCalculator(1e-12, SW.r_cut))
    _, phi = O.get_modes(O.dynamical_matrix(f2, sysd.masses))
    phi32 = np.asfortranarray(phi.astype(np.float32))
    return sysd, calc, phi32, fc3.DeviceArray.from_host(phi32, ctx=ctx)


@pytest.mark.parametrize("cells,seed,su,tol,masses", [
    ((2, 2, 2), 51, 0.03, 1e-12, "si"),
    ((2, 2, 2), 52, 0.03, 0.1, None),
    ((2, 2, 2), 53, 0.0, 1e-12, "random"),
    ((3, 3, 3), 54, 0.03, 1e-12, "si"),
    ((3, 3, 3), 55, 0.03, 0.1, "si"),
    ((1, 1, 1), 59, 0.03, 1e-12, "si"),      # 24 DoF: every atom's second shell wraps
    ((2, 2, 2), 60, 0.2, 1e-12, "random"),   # up to 14 neighbours, 484 blocks per slab
])
def test_fused_sw_matches_fp64_oracle(ctx, cells, seed, su, tol, masses):
    sysd, calc, phi32, dphi = _sw_case(ctx, cells, seed, su, tol)
    if masses == "si":
        m = sysd.masses
    elif masses == "random":
        m = np.random.default_rng(seed).uniform(20.0, 40.0, size=sysd.n_atoms)
    else:
        m = None
    k3 = fc3.third_order_mcc3(sysd, SW, calc, dphi, masses=m, ctx=ctx).to_host()
    psi = O.sw_third_order(sysd, SW, calc)
    if m is not None:
        O.mass_weight_third_order(psi, m)
    ref = O.mcc3_fp64(psi.astype(np.float32), phi32)
    err = _rel(k3, ref)
    print(f"SW {cells} seed {seed} tol {tol} masses {masses}: rel fro err vs FP64 {err:.3e}",
          flush=True)
    assert err <= 1e-5
    dphi.free()


@pytest.mark.parametrize("cells,seed,tol", [((2, 2, 2), 56, 1e-12), ((3, 3, 3), 57, 1e-12),
                                            ((3, 3, 3), 58, 0.1)])
def test_fused_sw_matches_dense_path(ctx, cells, seed, tol):
    sysd, calc, phi32, dphi = _sw_case(ctx, cells, seed, 0.03, tol)
    n = sysd.n_dof
    k3f = fc3.third_order_mcc3(sysd, SW, calc, dphi, masses=sysd.masses, ctx=ctx).to_host()
    psi32 = fc3.DeviceArray((n, n, n), np.float32, ctx)
    fc3.third_order_device(sysd, SW, calc, psi32=psi32, masses=sysd.masses, ctx=ctx)
    k3d = fc3.mcc3(psi32, dphi, precision="bf16x3").to_host()
    psi32.free()
    err = _rel(k3f, k3d.astype(np.float64))
    print(f"SW {cells} (n={n}) tol {tol}: fused vs dense BF16x3 rel {err:.3e}", flush=True)
    assert err <= 1e-5
    dphi.free()


def test_fused_sw_global_coefficients_bitwise(ctx, monkeypatch):
    """Slabs whose block coefficients do not fit in LDS read them from global scratch
    (FC3_SW_FUSED_GLOBAL=1 forces that kernel): the same values multiplied in the same
    order, so K3 must agree bit for bit."""
    sysd, calc, phi32, dphi = _sw_case(ctx, (3, 3, 3), 61, 0.03, 1e-12)
    k_lds = fc3.third_order_mcc3(sysd, SW, calc, dphi, masses=sysd.masses, ctx=ctx).to_host()
    monkeypatch.setenv("FC3_SW_FUSED_GLOBAL", "1")
    k_glb = fc3.third_order_mcc3(sysd, SW, calc, dphi, masses=sysd.masses, ctx=ctx).to_host()
    monkeypatch.delenv("FC3_SW_FUSED_GLOBAL")
    diff = int(np.count_nonzero(k_lds.view(np.uint32) != k_glb.view(np.uint32)))
    print(f"SW fused LDS vs global coefficients: {diff} entries differ", flush=True)
    assert diff == 0
    dphi.free()


@pytest.mark.parametrize("a,masses", [(4.8, "si"), (4.6, None)])
def test_fused_sw_compressed_diamond(ctx, a, masses):
    """Diamond compressed to a = 4.8 / 4.6 A: 16 / 20 neighbours and 727 / 973 blocks per
    slab atom, past what LDS holds (585), so mode 1 reads its coefficients from global
    scratch.  Against the FP64 oracle and the dense route."""
    sysd = S.diamond(a, (2, 2, 2), S.SI_MASS, 0.03, seed=62)
    calc = S.AutoDiffCalculator(1e-12, SW.r_cut)
    f2 = O.sw_second_order(sysd, SW, calc)
    _, phi = O.get_modes(O.dynamical_matrix(f2, sysd.masses))
    phi32 = np.asfortranarray(phi.astype(np.float32))
    dphi = fc3.DeviceArray.from_host(phi32, ctx=ctx)
    m = sysd.masses if masses == "si" else None
    k3 = fc3.third_order_mcc3(sysd, SW, calc, dphi, masses=m, ctx=ctx).to_host()
    psi = O.sw_third_order(sysd, SW, calc)
    if m is not None:
        O.mass_weight_third_order(psi, m)
    ref = O.mcc3_fp64(psi.astype(np.float32), phi32)
    err = _rel(k3, ref)
    n = sysd.n_dof
    psi32 = fc3.DeviceArray((n, n, n), np.float32, ctx)
    fc3.third_order_device(sysd, SW, calc, psi32=psi32, masses=m, ctx=ctx)
    k3d = fc3.mcc3(psi32, dphi, precision="bf16x3").to_host()
    psi32.free()
    errd = _rel(k3, k3d.astype(np.float64))
    print(f"SW compressed diamond a = {a}: rel fro err vs FP64 {err:.3e}, vs dense {errd:.3e}",
          flush=True)
    assert err <= 1e-5 and errd <= 1e-5
    dphi.free()


def test_fused_sw_rejects_bad_cutoff(ctx):
    sysd = S.diamond(S.SI_A, (2, 2, 2), S.SI_MASS)
    n = sysd.n_dof
    dphi = fc3.DeviceArray.from_host(np.eye(n, dtype=np.float32, order="F"), ctx=ctx)
    with pytest.raises(AssertionError):
        fc3.third_order_mcc3(sysd, SW, S.AutoDiffCalculator(1e-12, SW.r_cut + 0.1), dphi, ctx=ctx)
    dphi.free()
