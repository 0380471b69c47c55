"""Run-to-run bitwise determinism of the product paths: the race detector.

No kernel of the library uses floating-point atomics. Every reduction has a fixed order:
- segmented sums in record order;
- ASR sums in neighbour order;
- block partials summed in block order.

So each entry point must return the same bits whenever it runs on the same input. Each
case runs twice. Between the runs:
- a call of another size leaves other data in the handle's workspace;
- the caller's output buffer is refilled with NaN.

Any of these makes the second run differ:
- a write race;
- a read of memory that another workgroup is still writing;
- a read of scratch or output memory that the call did not write first.

The reference's loops are deterministic by construction (owner-computes for LJ, serial SW
loops: third_order_analytical.jl:14-15, third_order_AD.jl:91,128,147). These tests hold
the library to the same standard.
"""
import numpy as np
import pytest

import fc3
from fc3 import systems as S
from oracle import fc3_oracle as O

pytestmark = pytest.mark.gpu

LJ = S.ARGON_LJ
SW = S.StillingerWeberSilicon()


def _nan(shape, dtype, ctx):
    return fc3.DeviceArray.from_host(np.full(shape, np.nan, dtype=dtype, order="F"), ctx=ctx)


def _same(a: np.ndarray, b: np.ndarray, what: str):
    assert a.dtype == b.dtype and a.shape == b.shape, what
    eq = (a == b) | (np.isnan(a) & np.isnan(b))
    bad = int(eq.size - np.count_nonzero(eq))
    print(f"{what}: {bad} of {a.size} entries differ between runs", flush=True)
    assert bad == 0, what
    assert not np.isnan(a).any(), f"{what}: NaN left in the output"


@pytest.fixture(scope="module")
def lj324(ctx):
    sysd = S.fcc(S.ARGON_A, (3, 3, 3), S.ARGON_MASS, 0.05, seed=324)
    calc = S.AnalyticalCalculator(1e-12, LJ.r_cut)
    f2 = O.lj_second_order(sysd, LJ, calc)
    _, phi = O.get_modes(O.dynamical_matrix(f2, sysd.masses))
    phi32 = np.asfortranarray(phi.astype(np.float32))
    return sysd, calc, fc3.DeviceArray.from_host(phi32, ctx=ctx)


def _other_work(ctx):
    """A call of another size between two runs (new data in the workspace)."""
    sysd = S.fcc(S.ARGON_A, (2, 2, 2), S.ARGON_MASS, 0.05, seed=7)
    n = sysd.n_dof
    calc = S.AnalyticalCalculator(1e-12, LJ.r_cut)
    psi = fc3.DeviceArray((n, n, n), np.float32, ctx)
    fc3.third_order_device(sysd, LJ, calc, psi32=psi, masses=sysd.masses, ctx=ctx)
    phi = fc3.DeviceArray.from_host(
        np.asfortranarray(np.random.default_rng(7).standard_normal((n, n)).astype(np.float32)),
        ctx=ctx)
    fc3.mcc3(psi, phi, precision="bf16x3").free()
    psi.free()
    phi.free()
    ctx.synchronize()


def test_lj_assembly_is_deterministic(ctx, lj324):
    sysd, calc, _ = lj324
    n = sysd.n_dof
    runs = []
    for r in range(2):
        p64 = _nan((n, n, n), np.float64, ctx)
        p32 = _nan((n, n, n), np.float32, ctx)
        fc3.third_order_device(sysd, LJ, calc, psi64=p64, psi32=p32, masses=sysd.masses, ctx=ctx)
        runs.append((p64.to_host(), p32.to_host()))
        p64.free()
        p32.free()
        _other_work(ctx)
    _same(runs[0][0], runs[1][0], "LJ FCC 3x3x3 FP64 Psi")
    _same(runs[0][1], runs[1][1], "LJ FCC 3x3x3 FP32 mass-weighted Psi")


@pytest.mark.parametrize("tol", [1e-12, 0.1])
def test_sw_assembly_is_deterministic(ctx, tol):
    sysd = S.diamond(S.SI_A, (2, 2, 2), S.SI_MASS, 0.03, seed=192)
    calc = S.AutoDiffCalculator(tol, SW.r_cut)
    n = sysd.n_dof
    runs = []
    for r in range(2):
        p64 = _nan((n, n, n), np.float64, ctx)
        p32 = _nan((n, n, n), np.float32, ctx)
        fc3.third_order_device(sysd, SW, calc, psi64=p64, psi32=p32, masses=sysd.masses, ctx=ctx)
        runs.append((p64.to_host(), p32.to_host()))
        p64.free()
        p32.free()
        _other_work(ctx)
    _same(runs[0][0], runs[1][0], f"SW diamond 2x2x2 tol {tol} FP64 Psi")
    _same(runs[0][1], runs[1][1], f"SW diamond 2x2x2 tol {tol} FP32 Psi")


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_mcc3_is_deterministic(ctx, lj324, precision):
    sysd, calc, phi = lj324
    n = sysd.n_dof
    psi = fc3.DeviceArray((n, n, n), np.float32, ctx)
    fc3.third_order_device(sysd, LJ, calc, psi32=psi, masses=sysd.masses, ctx=ctx)
    full, uniq = [], []
    for r in range(2):
        out = _nan((n, n, n), np.float32, ctx)
        fc3.mcc3(psi, phi, precision=precision, out=out)
        full.append(out.to_host())
        out.free()
        u = _nan((fc3.unique_count(n),), np.float32, ctx)
        fc3.mcc3_unique(psi, phi, precision=precision, out=u)
        uniq.append(u.to_host())
        u.free()
        _other_work(ctx)
    blocked = [fc3.mcc3(psi, phi, block_size=12, precision=precision) for _ in range(2)]
    psi.free()
    _same(full[0], full[1], f"mcc3 {precision} n={n}")
    _same(uniq[0], uniq[1], f"mcc3_unique {precision} n={n}")
    _same(blocked[0], blocked[1], f"mcc3 blocked (block 12) {precision} n={n}")


def test_fused_is_deterministic(ctx, lj324):
    sysd, calc, phi = lj324
    n = sysd.n_dof
    runs = []
    for r in range(2):
        out = _nan((n, n, n), np.float32, ctx)
        fc3.third_order_mcc3(sysd, LJ, calc, phi, masses=sysd.masses, out=out, ctx=ctx)
        runs.append(out.to_host())
        out.free()
        _other_work(ctx)
    _same(runs[0], runs[1], f"fused third_order + mcc3 n={n}")


@pytest.mark.parametrize("tol", [1e-12, 0.1])
def test_fused_sw_is_deterministic(ctx, tol):
    sysd = S.diamond(S.SI_A, (3, 3, 3), S.SI_MASS, 0.03, seed=648)
    calc = S.AutoDiffCalculator(tol, SW.r_cut)
    n = sysd.n_dof
    phi = fc3.DeviceArray.from_host(
        np.asfortranarray(np.random.default_rng(648).standard_normal((n, n)).astype(np.float32)),
        ctx=ctx)
    runs = []
    for r in range(2):
        out = _nan((n, n, n), np.float32, ctx)
        fc3.third_order_mcc3(sysd, SW, calc, phi, masses=sysd.masses, out=out, ctx=ctx)
        runs.append(out.to_host())
        out.free()
        _other_work(ctx)
    phi.free()
    _same(runs[0], runs[1], f"fused SW third_order + mcc3 n={n} tol {tol}")


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_adjoints_are_deterministic(ctx, lj324, precision):
    sysd, calc, phi = lj324
    n = sysd.n_dof
    psi = fc3.DeviceArray((n, n, n), np.float32, ctx)
    fc3.third_order_device(sysd, LJ, calc, psi32=psi, masses=sysd.masses, ctx=ctx)
    k3 = fc3.mcc3(psi, phi, precision=precision)
    res = []
    for r in range(2):
        loss, dk3 = fc3.weighted_square_loss(k3, seed=324)
        dpsi, dphi = fc3.mcc3_backward(psi, phi, dk3, precision=precision)
        g = fc3.third_order_param_backward(sysd, LJ, calc, dpsi, masses=sysd.masses, ctx=ctx)
        res.append((loss, dk3.to_host(), dpsi.to_host(), dphi.to_host(), g))
        for a in (dk3, dpsi, dphi):
            a.free()
        _other_work(ctx)
    k3.free()
    psi.free()
    assert res[0][0] == res[1][0], "weighted square loss"
    _same(res[0][1], res[1][1], "dL/dK3")
    _same(res[0][2], res[1][2], f"dL/dPsi {precision}")
    _same(res[0][3], res[1][3], f"dL/dphi {precision}")
    assert res[0][4] == res[1][4], f"LJ parameter gradients {res[0][4]} {res[1][4]}"


def test_sw_param_backward_is_deterministic(ctx):
    sysd = S.diamond(S.SI_A, (2, 2, 2), S.SI_MASS, 0.03, seed=193)
    calc = S.AutoDiffCalculator(0.1, SW.r_cut)
    n = sysd.n_dof
    g = np.random.default_rng(193).standard_normal((n, n, n)).astype(np.float32)
    dpsi = fc3.DeviceArray.from_host(np.asfortranarray(g), ctx=ctx)
    res = []
    for r in range(2):
        res.append(fc3.third_order_param_backward(sysd, SW, calc, dpsi, masses=sysd.masses,
                                                  ctx=ctx))
        _other_work(ctx)
    dpsi.free()
    print(f"SW parameter gradients {res[0]}", flush=True)
    assert res[0] == res[1]
