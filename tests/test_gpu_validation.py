"""Caller buffers are validated before any library call, and mass weighting takes
D = n / N_atoms in 1:3 (src/analytical/third_order_analytical.jl:185-214).

Each rejected call must raise before a kernel runs: the buffers hold a sentinel that
is checked to be unchanged afterwards.
"""
import numpy as np
import pytest

import fc3
from fc3 import systems as S
from oracle import fc3_oracle as O

pytestmark = pytest.mark.gpu

CALC = S.AnalyticalCalculator(1e-12, S.ARGON_LJ.r_cut)


def _sentinel(shape, dtype, ctx):
    host = np.full(shape, 7.0, dtype=dtype, order="F")
    return fc3.DeviceArray.from_host(host, ctx=ctx), host


def test_third_order_device_rejects_wrong_buffers(ctx):
    sysd = S.fcc(S.ARGON_A, (2, 2, 2), S.ARGON_MASS, 0.05, seed=3)
    n = sysd.n_dof
    small, h_small = _sentinel((n, n, n), np.float32, ctx)
    with pytest.raises(AssertionError, match="psi64"):
        fc3.third_order_device(sysd, S.ARGON_LJ, CALC, psi64=small, ctx=ctx)   # FP32 as FP64
    short, h_short = _sentinel((n, n, n - 1), np.float32, ctx)
    with pytest.raises(AssertionError, match="psi32"):
        fc3.third_order_device(sysd, S.ARGON_LJ, CALC, psi32=short, ctx=ctx)
    with pytest.raises(AssertionError, match="masses"):
        fc3.third_order_device(sysd, S.ARGON_LJ, CALC, psi32=small,
                               masses=sysd.masses[:-1], ctx=ctx)
    with pytest.raises(AssertionError):
        fc3.third_order_device(sysd, S.ARGON_LJ, CALC, ctx=ctx)
    ctx.synchronize()
    assert np.array_equal(small.to_host(), h_small)
    assert np.array_equal(short.to_host(), h_short)


def test_other_entry_points_reject_wrong_buffers(ctx):
    sysd = S.fcc(S.ARGON_A, (2, 2, 2), S.ARGON_MASS, 0.05, seed=4)
    n = sysd.n_dof
    f32_2d, h2 = _sentinel((n, n), np.float32, ctx)
    f32_3d, h3 = _sentinel((n, n, n), np.float32, ctx)
    with pytest.raises(AssertionError, match="out"):
        fc3.second_order_device(sysd, S.ARGON_LJ, CALC, out=f32_2d, ctx=ctx)
    with pytest.raises(AssertionError, match="Psi"):
        fc3.mass_weight_third_order_(f32_3d, sysd.masses, ctx=ctx)       # FP32 tensor
    with pytest.raises(AssertionError, match="D"):
        fc3.mass_weight_third_order_(f32_3d, np.ones(7), ctx=ctx)        # 96 / 7 not whole
    f64_3d, h64 = _sentinel((n, n, n), np.float64, ctx)
    with pytest.raises(AssertionError, match="Psi"):
        fc3.mcc3_(f64_3d, f32_2d)
    with pytest.raises(AssertionError, match="dK3"):
        fc3.mcc3_backward(f32_3d, f32_2d, f32_2d)
    with pytest.raises(AssertionError, match="out"):
        fc3.mcc3(f32_3d, f32_2d, out=f64_3d)
    with pytest.raises(AssertionError, match="masses"):
        fc3.third_order_param_backward(sysd, S.ARGON_LJ, CALC, f32_3d, masses=np.ones(3))
    gone, _ = _sentinel((n, n), np.float32, ctx)
    gone.free()
    with pytest.raises(AssertionError, match="live"):
        fc3.mcc3(f32_3d, gone)
    ctx.synchronize()
    assert np.array_equal(f32_2d.to_host(), h2)
    assert np.array_equal(f32_3d.to_host(), h3)
    assert np.array_equal(f64_3d.to_host(), h64)


@pytest.mark.parametrize("D", [1, 2, 3])
def test_mass_weight_dims_match_oracle(ctx, D):
    """D = n / N_atoms: atom of index ii is ii // D.  Bit-exact against the oracle's
    restatement, and against the definition evaluated with numpy broadcasting."""
    N = 9
    n = D * N
    rng = np.random.default_rng(60 + D)
    psi = rng.standard_normal((n, n, n))
    psi[np.abs(psi) < 0.7] = 0.0
    psi = np.asfortranarray(psi)
    m = rng.uniform(1.0, 60.0, N)
    ref = O.mass_weight_third_order(psi.copy(order="F"), m)
    mr = np.repeat(m, D)
    direct = np.where(psi != 0.0, psi / np.sqrt((mr[:, None, None] * mr[None, :, None]) *
                                                mr[None, None, :]), 0.0)
    assert np.array_equal(ref, direct)
    d = fc3.DeviceArray.from_host(psi, ctx=ctx)
    fc3.mass_weight_third_order_(d, m, ctx=ctx)
    assert np.array_equal(d.to_host(), ref)
    dfc = fc3.DenseForceConstants(psi.copy(order="F"), "eV / Angstrom^3", 0.0)
    fc3.mass_weight_third_order_(dfc, m, ctx=ctx)
    assert np.array_equal(dfc.values, ref)
