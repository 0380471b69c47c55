"""Adjoints on the GPU: mcc3 backward and the LJ parameter backward of third_order.

Pins (none retypes the adjoint formulas):
  * mcc3 dL/dPsi: the adjoint identity <G, mcc3(dPsi_dir, phi)> = <dL/dPsi, dPsi_dir> for
    random directions, with the forward evaluated by the FP64 oracle contraction.
  * mcc3 dL/dphi: central differences in FP64 of L(phi) = <G, mcc3(Psi, phi)> along
    random directions (L is a cubic polynomial in phi, so the truncation error is
    O(h^2) relative; h = 1e-4).
  * LJ dL/dsigma: central differences (relative step 1e-6, as the C5 configuration asks)
    of F(sigma) = <dPsi, Psi(sigma)> with Psi from the FP64 oracle restatement of the
    reference loop; dL/depsilon: Psi is linear in epsilon, so dF/depsilon = F / epsilon
    exactly.
  * The whole chain L = sum w K3^2 (w ~ U(0,1) seeded) through mass weighting and mcc3,
    against FP64 central differences of the oracle chain.
Bars: FP64 paths 1e-6 relative (FD noise), FP32 paths 2e-5 relative.
"""
import numpy as np
import pytest

import fc3
from fc3 import systems as S
from oracle import fc3_oracle as O

pytestmark = pytest.mark.gpu

LJ = S.ARGON_LJ


def _rand(shape, seed, dtype=np.float64):
    return np.asfortranarray(np.random.default_rng(seed).standard_normal(shape).astype(dtype))


def _dot(a, b):
    return float(np.sum(np.asarray(a, dtype=np.float64) * np.asarray(b, dtype=np.float64)))


@pytest.mark.parametrize("n", [48, 96, 100])
def test_mcc3_backward_dpsi_adjoint_identity(ctx, n):
    psi = _rand((n, n, n), 1, np.float32)
    phi, _ = np.linalg.qr(np.random.default_rng(2).standard_normal((n, n)))
    phi = np.asfortranarray(phi.astype(np.float32))
    G = _rand((n, n, n), 3, np.float32)
    d = fc3.DeviceArray.from_host
    dpsi, dphi = fc3.mcc3_backward(d(psi, ctx=ctx), d(phi, ctx=ctx), d(G, ctx=ctx))
    gp = dpsi.to_host()
    for seed in (4, 5):
        X = _rand((n, n, n), seed)
        lhs = _dot(G, O.mcc3_fp64(X, phi))
        rhs = _dot(gp, X)
        print(f"n={n} <G, mcc3(X)> = {lhs:.9e}, <dPsi, X> = {rhs:.9e}", flush=True)
        assert abs(lhs - rhs) <= 2e-5 * np.sqrt(_dot(G, G) * _dot(O.mcc3_fp64(X, phi), O.mcc3_fp64(X, phi)))
    # dL/dPsi alone (the mcc3-kernel path) agrees with the path that also forms dL/dphi
    for prec in ("fp32", "bf16x3"):
        only, none = fc3.mcc3_backward(d(psi, ctx=ctx), d(phi, ctx=ctx), d(G, ctx=ctx),
                                       precision=prec, dphi=False)
        assert none is None
        rel = np.linalg.norm((only.to_host().astype(np.float64) - gp).ravel()) / np.linalg.norm(gp)
        print(f"n={n} dPsi via mcc3 kernels ({prec}) vs rocBLAS path: rel {rel:.2e}", flush=True)
        assert rel <= (2e-6 if prec == "fp32" else 1e-5)


@pytest.mark.parametrize("n,prec", [(48, "fp32"), (96, "fp32"), (20, "bf16x3"), (48, "bf16x3"),
                                    (96, "bf16x3"), (100, "bf16x3")])
def test_mcc3_backward_dphi_central_difference(ctx, n, prec):
    psi = _rand((n, n, n), 11, np.float32)
    phi = _rand((n, n), 12, np.float32) / np.sqrt(n)
    phi = np.asfortranarray(phi.astype(np.float32))
    G = _rand((n, n, n), 13, np.float32)
    d = fc3.DeviceArray.from_host
    _, dphi = fc3.mcc3_backward(d(psi, ctx=ctx), d(phi, ctx=ctx), d(G, ctx=ctx), dpsi=False,
                                precision=prec)
    gf = dphi.to_host().astype(np.float64)
    p64, f64, g64 = (a.astype(np.float64) for a in (psi, phi, G))
    h = 1e-4
    for seed in (14, 15, 16):
        D = _rand((n, n), seed)
        fd = (_dot(g64, O.mcc3_fp64(p64, f64 + h * D)) - _dot(g64, O.mcc3_fp64(p64, f64 - h * D))) / (2 * h)
        ad = _dot(gf, D)
        print(f"n={n} {prec} dir {seed}: FD {fd:.9e} adjoint {ad:.9e} rel {abs(fd - ad) / abs(fd):.2e}", flush=True)
        assert abs(fd - ad) <= 2e-5 * np.linalg.norm(gf) * np.linalg.norm(D)


@pytest.mark.parametrize("n", [20, 48, 100])
def test_mcc3_backward_bf16x3_matches_fp32(ctx, n):
    """The BF16x3 pullback (mode products and split-K contractions on the BF16x3 kernels)
    against the FP32 one, both outputs, relative Frobenius error <= 1e-5 (the BF16x3
    forward's bar)."""
    psi = _rand((n, n, n), 31, np.float32)
    phi, _ = np.linalg.qr(np.random.default_rng(32).standard_normal((n, n)))
    phi = np.asfortranarray(phi.astype(np.float32))
    G = _rand((n, n, n), 33, np.float32)
    d = fc3.DeviceArray.from_host
    args = (d(psi, ctx=ctx), d(phi, ctx=ctx), d(G, ctx=ctx))
    p32, f32 = (a.to_host().astype(np.float64) for a in fc3.mcc3_backward(*args, precision="fp32"))
    pbf, fbf = (a.to_host().astype(np.float64) for a in fc3.mcc3_backward(*args, precision="bf16x3"))
    rp = np.linalg.norm((pbf - p32).ravel()) / np.linalg.norm(p32.ravel())
    rf = np.linalg.norm((fbf - f32).ravel()) / np.linalg.norm(f32.ravel())
    print(f"n={n} bf16x3 vs fp32: dPsi rel {rp:.2e}, dphi rel {rf:.2e}", flush=True)
    assert rp <= 1e-5 and rf <= 1e-5


def _lj_sys(cells, su, seed):
    return S.fcc(S.ARGON_A, cells, S.ARGON_MASS, su, seed=seed)


@pytest.mark.parametrize("cells,su,seed,tol", [((2, 2, 2), 0.05, 21, 0.0), ((2, 2, 2), 0.05, 22, 1e-12),
                                               ((3, 3, 3), 0.03, 23, 1e-12)])
def test_lj_param_backward_fp64(ctx, cells, su, seed, tol):
    sysd = _lj_sys(cells, su, seed)
    calc = S.AnalyticalCalculator(tol, LJ.r_cut)
    n = sysd.n_dof
    dpsi = _rand((n, n, n), seed + 100)
    g = fc3.third_order_param_backward(sysd, LJ, calc, fc3.DeviceArray.from_host(dpsi, ctx=ctx))
    F = lambda pot: _dot(dpsi, O.lj_third_order(sysd, pot, calc))
    F0 = F(LJ)
    # epsilon: Psi is linear in epsilon
    print(f"{cells}: dF/deps {g['epsilon']:.12e} vs F/eps {F0 / LJ.epsilon:.12e}", flush=True)
    assert abs(g["epsilon"] - F0 / LJ.epsilon) <= 1e-10 * abs(F0 / LJ.epsilon)
    # sigma: central difference, relative step 1e-6
    hs = 1e-6 * LJ.sigma
    fd = (F(S.LJ(LJ.sigma + hs, LJ.epsilon, LJ.r_cut)) - F(S.LJ(LJ.sigma - hs, LJ.epsilon, LJ.r_cut))) / (2 * hs)
    print(f"{cells}: dF/dsigma adjoint {g['sigma']:.12e} FD {fd:.12e} rel {abs(fd - g['sigma']) / abs(fd):.2e}",
          flush=True)
    assert abs(fd - g["sigma"]) <= 1e-6 * abs(fd)


def test_lj_param_backward_fp32_mass_weighted(ctx):
    sysd = _lj_sys((2, 2, 2), 0.05, 31)
    calc = S.AnalyticalCalculator(1e-12, LJ.r_cut)
    n = sysd.n_dof
    m = np.random.default_rng(32).uniform(20.0, 40.0, size=sysd.n_atoms)
    d32 = _rand((n, n, n), 33, np.float32)
    g = fc3.third_order_param_backward(sysd, LJ, calc, fc3.DeviceArray.from_host(d32, ctx=ctx), masses=m)

    def F(pot):
        psi = O.lj_third_order(sysd, pot, calc)
        O.mass_weight_third_order(psi, m)
        return _dot(d32, psi)

    F0 = F(LJ)
    assert abs(g["epsilon"] - F0 / LJ.epsilon) <= 1e-10 * abs(F0 / LJ.epsilon)
    hs = 1e-6 * LJ.sigma
    fd = (F(S.LJ(LJ.sigma + hs, LJ.epsilon, LJ.r_cut)) - F(S.LJ(LJ.sigma - hs, LJ.epsilon, LJ.r_cut))) / (2 * hs)
    print(f"mass weighted: adjoint {g['sigma']:.12e} FD {fd:.12e}", flush=True)
    assert abs(fd - g["sigma"]) <= 1e-6 * abs(fd)


def test_lj_chain_sum_w_k3_squared(ctx):
    """L = sum w K3^2, K3 = mcc3(mass-weighted FP32 Psi(sigma, epsilon), phi), phi fixed:
    the GPU chain (forward, K3 cotangent 2 w K3, mcc3 backward, parameter backward)
    against FP64 central differences of the oracle chain."""
    sysd = _lj_sys((2, 2, 2), 0.05, 41)
    calc = S.AnalyticalCalculator(1e-12, LJ.r_cut)
    n = sysd.n_dof
    f2 = O.lj_second_order(sysd, LJ, calc)
    _, phi = O.get_modes(O.dynamical_matrix(f2, sysd.masses))
    phi32 = np.asfortranarray(phi.astype(np.float32))
    w = np.asfortranarray(np.random.default_rng(42).uniform(0.0, 1.0, size=(n, n, n)))
    # GPU forward
    psi32 = fc3.DeviceArray((n, n, n), np.float32, ctx)
    fc3.third_order_device(sysd, LJ, calc, psi32=psi32, masses=sysd.masses, ctx=ctx)
    dphi = fc3.DeviceArray.from_host(phi32, ctx=ctx)
    k3 = fc3.mcc3(psi32, dphi).to_host().astype(np.float64)
    G = fc3.DeviceArray.from_host(2.0 * w * k3, dtype=np.float32, ctx=ctx)
    dpsi, _ = fc3.mcc3_backward(psi32, dphi, G, dphi=False)
    g = fc3.third_order_param_backward(sysd, LJ, calc, dpsi, masses=sysd.masses)

    def L(pot):
        psi = O.lj_third_order(sysd, pot, calc)
        O.mass_weight_third_order(psi, sysd.masses)
        K = O.mcc3_fp64(psi, phi32)
        return float(np.sum(w * K * K))

    for name, plus, minus, h in (
            ("sigma", S.LJ(LJ.sigma * (1 + 1e-6), LJ.epsilon, LJ.r_cut),
             S.LJ(LJ.sigma * (1 - 1e-6), LJ.epsilon, LJ.r_cut), LJ.sigma * 1e-6),
            ("epsilon", S.LJ(LJ.sigma, LJ.epsilon * (1 + 1e-6), LJ.r_cut),
             S.LJ(LJ.sigma, LJ.epsilon * (1 - 1e-6), LJ.r_cut), LJ.epsilon * 1e-6)):
        fd = (L(plus) - L(minus)) / (2 * h)
        print(f"chain dL/d{name}: adjoint {g[name]:.9e} FD {fd:.9e} rel {abs(fd - g[name]) / abs(fd):.2e}",
              flush=True)
        assert abs(fd - g[name]) <= 2e-4 * abs(fd)


@pytest.mark.parametrize("shape,seed", [((96, 96, 96), 1500), ((7, 5, 3), (1 << 40) + 9),
                                        ((250, 250, 250), 77), ((251, 251, 250), 78)])
def test_weighted_square_loss_matches_oracle(ctx, shape, seed):
    """fc3_weighted_square_loss against the oracle's Philox restatement: the cotangent
    2 w K3 bit for bit, L to FP64 summation order (1e-13); a count that is not a
    multiple of 4 takes the scalar tail; counts past 4 x 3 grid strides (6.3M) take the
    four-step loop, and 251^2 x 250 leaves a scalar tail behind it."""
    k3 = np.asfortranarray(np.random.default_rng(5).standard_normal(shape).astype(np.float32))
    d = fc3.DeviceArray.from_host(k3, ctx=ctx)
    L, g = fc3.weighted_square_loss(d, seed)
    L0, g0 = O.weighted_square_loss(k3, seed)
    print(f"loss {shape}: device {L:.15e} oracle {L0:.15e}", flush=True)
    assert abs(L - L0) <= 1e-13 * abs(L0)
    assert np.array_equal(g.to_host(), g0)
    L1, none = fc3.weighted_square_loss(d, seed, grad=False)
    assert none is None and L1 == L
    g.free()
    d.free()


def test_lj_chain_device_weights_central_difference(ctx):
    """The differentiation workload's chain with the weights on the device (L = sum w K3^2,
    w from Philox keyed by the seed): forward, loss cotangent, mcc3 backward and the
    sigma / epsilon pullback, against FP64 central differences of the oracle chain with
    the same regenerated weights (FCC 2x2x2, 96 DoF)."""
    sysd = _lj_sys((2, 2, 2), 0.05, 43)
    calc = S.AnalyticalCalculator(1e-12, LJ.r_cut)
    n = sysd.n_dof
    seed = 20260
    f2 = O.lj_second_order(sysd, LJ, calc)
    _, phi = O.get_modes(O.dynamical_matrix(f2, sysd.masses))
    phi32 = np.asfortranarray(phi.astype(np.float32))
    psi32 = fc3.DeviceArray((n, n, n), np.float32, ctx)
    fc3.third_order_device(sysd, LJ, calc, psi32=psi32, masses=sysd.masses, ctx=ctx)
    dphi = fc3.DeviceArray.from_host(phi32, ctx=ctx)
    k3 = fc3.mcc3(psi32, dphi)
    _, G = fc3.weighted_square_loss(k3, seed)
    dpsi, _ = fc3.mcc3_backward(psi32, dphi, G, dphi=False)
    g = fc3.third_order_param_backward(sysd, LJ, calc, dpsi, masses=sysd.masses)
    w = O.philox_uniform(seed, n ** 3).reshape((n, n, n), order="F").astype(np.float64)

    def L(pot):
        psi = O.lj_third_order(sysd, pot, calc)
        O.mass_weight_third_order(psi, sysd.masses)
        K = O.mcc3_fp64(psi, phi32)
        return float(np.sum(w * K * K))

    for name, plus, minus, h in (
            ("sigma", S.LJ(LJ.sigma * (1 + 1e-6), LJ.epsilon, LJ.r_cut),
             S.LJ(LJ.sigma * (1 - 1e-6), LJ.epsilon, LJ.r_cut), LJ.sigma * 1e-6),
            ("epsilon", S.LJ(LJ.sigma, LJ.epsilon * (1 + 1e-6), LJ.r_cut),
             S.LJ(LJ.sigma, LJ.epsilon * (1 - 1e-6), LJ.r_cut), LJ.epsilon * 1e-6)):
        fd = (L(plus) - L(minus)) / (2 * h)
        print(f"device-weight chain dL/d{name}: adjoint {g[name]:.9e} FD {fd:.9e} "
              f"rel {abs(fd - g[name]) / abs(fd):.2e}", flush=True)
        assert abs(fd - g[name]) <= 2e-4 * abs(fd)
    for d in (psi32, dphi, k3, G, dpsi):
        d.free()


SW_POT = S.StillingerWeberSilicon()


def _sw_with(**kw):
    from dataclasses import replace
    return S.StillingerWeberSilicon(params=replace(SW_POT.params, **kw))


@pytest.mark.parametrize("tol,seed,weighted", [(0.0, 51, False), (1e-12, 52, False), (1e-12, 53, True)])
def test_sw_param_backward_central_difference(ctx, tol, seed, weighted):
    """dF/d(A, B, lambda, gamma) of F = <dPsi, Psi_SW> against FP64 central differences
    (relative step 1e-6) of the oracle's SW third_order (hyper-dual restatement of the
    reference's AutoDiffCalculator)."""
    sysd = S.diamond(S.SI_A, (2, 2, 2), S.SI_MASS, 0.03, seed=seed)
    calc = S.AutoDiffCalculator(tol, SW_POT.r_cut)
    n = sysd.n_dof
    masses = np.random.default_rng(seed).uniform(20.0, 40.0, size=sysd.n_atoms) if weighted else None
    if weighted:
        d = _rand((n, n, n), seed + 100, np.float32)
    else:
        d = _rand((n, n, n), seed + 100)
    g = fc3.third_order_param_backward(sysd, SW_POT, calc, fc3.DeviceArray.from_host(d, ctx=ctx),
                                       masses=masses)

    def F(pot):
        psi = O.sw_third_order(sysd, pot, calc)
        if weighted:
            O.mass_weight_third_order(psi, masses)
        return _dot(d, psi)

    prm = SW_POT.params
    for name, field in (("A", "A"), ("B", "B"), ("lambda", "lam"), ("gamma", "gamma")):
        x = getattr(prm, field)
        h = 1e-6 * x
        fd = (F(_sw_with(**{field: x + h})) - F(_sw_with(**{field: x - h}))) / (2 * h)
        print(f"SW tol={tol} weighted={weighted} dF/d{name}: adjoint {g[name]:.12e} FD {fd:.12e} "
              f"rel {abs(fd - g[name]) / abs(fd):.2e}", flush=True)
        assert abs(fd - g[name]) <= 1e-6 * abs(fd)


@pytest.mark.parametrize("cells,weighted", [((2, 2, 2), False), ((2, 2, 2), True),
                                            ((3, 3, 3), True)])
def test_lj_param_backward_against_dual_numbers(ctx, cells, weighted):
    """dL/dsigma, dL/depsilon for L = <G, Psi> against the oracle's forward-mode dual-number
    derivative d Psi / dp contracted with G (96 and 324 DoF): <= 1e-10 relative."""
    sysd = _lj_sys(cells, 0.05, 61)
    calc = S.AnalyticalCalculator(1e-12, LJ.r_cut)
    n = sysd.n_dof
    m = np.random.default_rng(62).uniform(20.0, 40.0, sysd.n_atoms) if weighted else None
    G = _rand((n, n, n), 63, np.float32 if weighted else np.float64)
    g = fc3.third_order_param_backward(sysd, LJ, calc, fc3.DeviceArray.from_host(G, ctx=ctx),
                                       masses=m)
    for name in ("sigma", "epsilon"):
        ref = _dot(G, O.lj_third_order_dparam(sysd, LJ, calc, name, masses=m))
        rel = abs(g[name] - ref) / abs(ref)
        print(f"LJ {cells} weighted={weighted} dL/d{name}: GPU {g[name]:.15e} dual {ref:.15e} "
              f"rel {rel:.2e}", flush=True)
        assert rel <= 1e-10


@pytest.mark.parametrize("tol,weighted", [(0.0, False), (1e-12, True), (0.1, False), (0.1, True)])
def test_sw_param_backward_against_dual_numbers(ctx, tol, weighted):
    """dL/d(A, B, lambda, gamma) against the oracle's 4-direction hyper-dual derivative of
    the SW tensor (192 DoF): <= 1e-10 relative.  tol = 0.1 zeroes 38% of the nonzero
    entries: the pullback's mask, read from the forward block table, must be the
    oracle's mask of the assembled values."""
    sysd = S.diamond(S.SI_A, (2, 2, 2), S.SI_MASS, 0.03, seed=64)
    calc = S.AutoDiffCalculator(tol, SW_POT.r_cut)
    n = sysd.n_dof
    m = np.random.default_rng(65).uniform(20.0, 40.0, sysd.n_atoms) if weighted else None
    G = _rand((n, n, n), 66, np.float32 if weighted else np.float64)
    g = fc3.third_order_param_backward(sysd, SW_POT, calc, fc3.DeviceArray.from_host(G, ctx=ctx),
                                       masses=m)
    for name in ("A", "B", "lambda", "gamma"):
        ref = _dot(G, O.sw_third_order_dparam(sysd, SW_POT, calc, name, masses=m))
        rel = abs(g[name] - ref) / abs(ref)
        print(f"SW tol={tol} weighted={weighted} dL/d{name}: GPU {g[name]:.15e} dual {ref:.15e} "
              f"rel {rel:.2e}", flush=True)
        assert rel <= 1e-10
