"""LJ force constants on the GPU against the CPU oracle (through the C ABI).

Bar: max |GPU - oracle| <= 1e-12 per entry (FP64, "1e-12 abs"), and the Frobenius
norm of the difference <= 1e-12 where the reference's own test uses that norm
(runtests.jl:155,160, FCC 4x4x4).  The FP32 mass-weighted output must equal
Float32(oracle / sqrt(m_i m_j m_k)) exactly.
"""
import ctypes

import numpy as np
import pytest

import fc3
from fc3 import systems as S
from oracle import fc3_oracle as O

pytestmark = pytest.mark.gpu

POT = S.ARGON_LJ


def _calc(tol):
    return S.AnalyticalCalculator(tol, POT.r_cut)


def _cmp(gpu, ref, label):
    d = np.abs(gpu - ref)
    mx = float(d.max())
    fro = float(np.sqrt((d.astype(np.float64) ** 2).sum()))
    eq = float((gpu == ref).mean())
    print(f"{label}: max|d|={mx:.3e} fro={fro:.3e} bit-equal={eq:.6f}", flush=True)
    return mx, fro


def test_ifc2_matches_oracle(ctx):
    for sysd, tol in [(S.fcc(S.ARGON_A, (4, 4, 4), S.ARGON_MASS), 1e-8),
                      (S.fcc(S.ARGON_A, (4, 4, 4), S.ARGON_MASS, 0.05, seed=21), 1e-12),
                      (S.fcc(S.ARGON_A, (3, 3, 3), S.ARGON_MASS, 0.05, seed=22), 1e-12)]:
        g = fc3.second_order(sysd, POT, _calc(tol), ctx=ctx).values
        r = O.lj_second_order(sysd, POT, _calc(tol))
        mx, fro = _cmp(g, r, f"ifc2 n={sysd.n_dof}")
        assert mx <= 1e-12 and fro <= 1e-12


def test_ifc3_known_answers(ctx):
    # runtests.jl:109-136 on the GPU result
    sysd = S.fcc(S.ARGON_A, (4, 4, 4), S.ARGON_MASS)
    f3 = fc3.third_order(sysd, POT, _calc(1e-8), ctx=ctx)
    rt = np.sqrt(np.finfo(np.float64).eps)
    for idx, val in [((3, 0, 0), -3.167167895), ((3, 1, 0), -3.979773863),
                     ((3, 3, 0), 3.167167895), ((4, 4, 0), 3.979773863),
                     ((5, 5, 0), -0.406302984), ((3, 0, 1), -3.979773863)]:
        assert abs(f3[idx] - val) <= rt * max(abs(val), abs(f3[idx])), (idx, f3[idx])
    assert f3[0, 0, 0] == 0.0
    f2 = fc3.second_order(sysd, POT, _calc(1e-8), ctx=ctx)
    for idx, val in [((0, 0), 7.727196175), ((1, 1), 7.727196175), ((2, 2), 7.727196175),
                     ((3, 0), -1.02002707), ((3, 1), -1.06589524)]:
        assert abs(f2[idx] - val) <= rt * abs(val)


@pytest.mark.parametrize("cells,sigma_u,seed", [((4, 4, 4), 0.0, 0), ((4, 4, 4), 0.05, 31),
                                                ((3, 3, 3), 0.05, 32), ((2, 2, 2), 0.08, 33)])
def test_ifc3_matches_oracle(ctx, cells, sigma_u, seed):
    sysd = S.fcc(S.ARGON_A, cells, S.ARGON_MASS, sigma_u, seed=seed)
    calc = _calc(1e-12)
    print(f"case cells={cells} sigma_u={sigma_u} n={sysd.n_dof}", flush=True)
    g = fc3.third_order(sysd, POT, calc, ctx=ctx).values
    r = O.lj_third_order(sysd, POT, calc)
    mx, fro = _cmp(g, r, f"ifc3 n={sysd.n_dof}")
    assert mx <= 1e-12
    assert fro <= 1e-12
    del g, r


def test_ifc3_fp32_mass_weighted_matches_oracle(ctx):
    sysd = S.fcc(S.ARGON_A, (3, 3, 3), S.ARGON_MASS, 0.05, seed=41)
    rng = np.random.default_rng(4)
    masses = rng.uniform(20.0, 60.0, size=sysd.n_atoms)
    calc = _calc(1e-12)
    n = sysd.n_dof
    d64 = fc3.DeviceArray((n, n, n), np.float64, ctx)
    d32 = fc3.DeviceArray((n, n, n), np.float32, ctx)
    fc3.third_order_device(sysd, POT, calc, psi64=d64, psi32=d32, masses=masses, ctx=ctx)
    g64, g32 = d64.to_host(), d32.to_host()
    r = O.lj_third_order(sysd, POT, calc)
    assert np.abs(g64 - r).max() <= 1e-12
    O.mass_weight_third_order(r, masses)
    r32 = r.astype(np.float32)
    assert np.array_equal(g32, r32), float(np.abs(g32 - r32).max())
    # the in-place mass weighting entry point (mass_weight_third_order!)
    fc3.mass_weight_third_order_(d64, masses, ctx=ctx)
    mw = d64.to_host()
    assert np.abs(mw - r).max() <= 1e-15 * np.abs(r).max()


def test_ifc3_odd_dof_count(ctx):
    # 31 atoms -> n = 93: exercises the scalar (non-16-byte) store path
    base = S.fcc(S.ARGON_A, (2, 2, 2), S.ARGON_MASS, 0.05, seed=51)
    sysd = S.SuperCellSystem(base.positions[:31], base.masses[:31], base.box_sizes)
    calc = _calc(1e-12)
    g = fc3.third_order(sysd, POT, calc, ctx=ctx).values
    r = O.lj_third_order(sysd, POT, calc)
    assert np.abs(g - r).max() <= 1e-12


def test_cutoff_assertion(ctx):
    sysd = S.fcc(S.ARGON_A, (1, 1, 1), S.ARGON_MASS)
    with pytest.raises(AssertionError, match="Cutoff larger than L/2"):
        fc3.third_order(sysd, POT, _calc(1e-12), ctx=ctx)


def test_ifc3_fcc666_sampled_columns(ctx):
    """FCC 6x6x6 (2592 DoF, sigma_u = 0.05 A): dense FP64 (139 GB) and FP32 mass-weighted
    (70 GB) tensors on the device; sampled columns against the column oracle."""
    sysd = S.fcc(S.ARGON_A, (6, 6, 6), S.ARGON_MASS, 0.05, seed=2592)
    calc = _calc(1e-12)
    n = sysd.n_dof
    free, _ = ctx.memory()
    need = n ** 3 * 12
    if free < need + (4 << 30):
        pytest.skip(f"needs {need / 1e9:.0f} GB of device memory, {free / 1e9:.0f} GB free")
    d64 = fc3.DeviceArray((n, n, n), np.float64, ctx)
    d32 = fc3.DeviceArray((n, n, n), np.float32, ctx)
    fc3.third_order_device(sysd, POT, calc, psi64=d64, psi32=d32, masses=sysd.masses, ctx=ctx)
    ctx.synchronize()
    rng = np.random.default_rng(6)
    # self columns, neighbour columns and random columns
    J = rng.integers(0, sysd.n_atoms, size=6)
    jj = list(3 * J + rng.integers(0, 3, size=6))
    kk = list(3 * J + rng.integers(0, 3, size=6))
    jj += list(3 * J + 1)
    kk += list(3 * ((J + 1) % sysd.n_atoms))
    jj += list(rng.integers(0, n, size=6))
    kk += list(rng.integers(0, n, size=6))
    ref = O.lj_third_order_columns(sysd, POT, calc, np.array(jj), np.array(kk))
    mm = np.repeat(sysd.masses, 3)
    col = np.empty(n, dtype=np.float64)
    col32 = np.empty(n, dtype=np.float32)
    lib = ctx.lib
    for c, (j, k) in enumerate(zip(jj, kk)):
        off = (int(j) + n * int(k)) * n
        fc3._lib.check(lib.fc3_memcpy_d2h(ctx.h, col.ctypes.data_as(ctypes.c_void_p),
                                          ctypes.c_void_p(d64.ptr.value + 8 * off), 8 * n))
        fc3._lib.check(lib.fc3_memcpy_d2h(ctx.h, col32.ctypes.data_as(ctypes.c_void_p),
                                          ctypes.c_void_p(d32.ptr.value + 4 * off), 4 * n))
        assert np.abs(col - ref[:, c]).max() <= 1e-12, (j, k)
        w = np.where(ref[:, c] != 0, ref[:, c] / np.sqrt((mm * mm[j]) * mm[k]), ref[:, c])
        assert np.array_equal(col32, w.astype(np.float32)), (j, k)
    d64.free()
    d32.free()
