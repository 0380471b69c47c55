"""Symmetric-unique K3 (fc3_mcc3_unique; the unique-part TODO of src/mcc.jl:116-117).

The packed entries K3[a, b, l], a <= b <= l, must equal the full mcc3's entries bit for
bit: in the BF16X3 path only the mode-1 / mode-2 GEMM tiles that hold such entries run,
and a tile's result does not depend on the other tiles of its launch.  Shapes cover the
permuted mode-1 layout (n % 16 == 0), the plain one with K padding (324, 1500), several
slab batches (768, 1500) and the pack-only fallbacks (FP32, n % 4 != 0).  At 96 DoF the
expanded tensor (every permutation of (a, b, l) takes the sorted entry) is checked against
the oracle's FP64 contraction: K3 of a symmetric Psi is symmetric up to rounding.
"""
import ctypes

import numpy as np
import pytest

import fc3
from fc3 import systems as S
from oracle import fc3_oracle as O

pytestmark = pytest.mark.gpu

CALC = S.AnalyticalCalculator(1e-12, S.ARGON_LJ.r_cut)


def _workload(ctx, cells, seed):
    sysd = S.fcc(S.ARGON_A, cells, S.ARGON_MASS, 0.05, seed=seed)
    n = sysd.n_dof
    _, phi = fc3.get_modes(fc3.dynamical_matrix(sysd, S.ARGON_LJ, CALC, ctx=ctx))
    phi32 = np.asfortranarray(phi.astype(np.float32))
    psi32 = fc3.DeviceArray((n, n, n), np.float32, ctx)
    fc3.third_order_device(sysd, S.ARGON_LJ, CALC, psi32=psi32, masses=sysd.masses, ctx=ctx)
    return sysd, psi32, fc3.DeviceArray.from_host(phi32, ctx=ctx), phi32


def _compare_to_full(ctx, psi32, dphi, n, prec):
    full = fc3.mcc3(psi32, dphi, precision=prec)
    uq = fc3.mcc3_unique(psi32, dphi, precision=prec)
    ctx.synchronize()
    u = uq.to_host()
    uq.free()
    assert u.shape == (fc3.unique_count(n),)
    slab = np.empty((n, n), dtype=np.float32, order="F")
    bad = 0
    for l in range(n):
        fc3._lib.check(ctx.lib.fc3_memcpy_d2h(ctx.h, slab.ctypes.data_as(ctypes.c_void_p),
                                              ctypes.c_void_p(full.ptr.value + 4 * n * n * l),
                                              4 * n * n))
        tri = slab[: l + 1, : l + 1].T[np.tril_indices(l + 1)]   # (a, b), b outer, a <= b
        lo = int(fc3.unique_index(0, 0, l))
        got = u[lo: lo + tri.size]
        bad += int(np.count_nonzero(got.view(np.uint32) != tri.view(np.uint32)))
    full.free()
    return u, bad


@pytest.mark.parametrize("cells,prec", [((2, 2, 2), "bf16x3"), ((2, 2, 2), "fp32"),
                                        ((3, 3, 3), "bf16x3"), ((3, 3, 3), "fp32"),
                                        ((4, 4, 4), "bf16x3"), ((5, 5, 5), "bf16x3")])
def test_unique_entries_bitwise_full(ctx, cells, prec):
    sysd, psi32, dphi, phi32 = _workload(ctx, cells, seed=40 + cells[0])
    n = sysd.n_dof
    print(f"mcc3_unique {prec} n={n}: comparing with the full K3", flush=True)
    u, bad = _compare_to_full(ctx, psi32, dphi, n, prec)
    print(f"mcc3_unique {prec} n={n}: {bad} entries differ", flush=True)
    assert bad == 0
    if n == 96:
        coo = O.lj_third_order_coo(sysd, S.ARGON_LJ, CALC)
        v32 = O.coo_mass_weight(coo, sysd.masses).astype(np.float32).astype(np.float64)
        psi = np.zeros((n, n, n))
        psi[coo[0], coo[1], coo[2]] = v32
        ref = O.mcc3_fp64(np.asfortranarray(psi), phi32.astype(np.float64))
        k = fc3.k3_from_unique(u, n).astype(np.float64)
        rel = np.linalg.norm((k - ref).ravel()) / np.linalg.norm(ref.ravel())
        assert rel <= (1e-5 if prec == "bf16x3" else 2e-6)
    psi32.free()
    dphi.free()


def test_unique_other_shapes(ctx):
    """n % 4 != 0 (the FP32 kernels for either precision): full K3 packed."""
    n = 30
    rng = np.random.default_rng(12)
    a = rng.standard_normal((n, n, n)).astype(np.float32)
    sym = (a + a.transpose(1, 0, 2) + a.transpose(2, 1, 0) + a.transpose(0, 2, 1) +
           a.transpose(1, 2, 0) + a.transpose(2, 0, 1)) / 6
    psi32 = fc3.DeviceArray.from_host(np.asfortranarray(sym), ctx=ctx)
    phi = np.asfortranarray(np.linalg.qr(rng.standard_normal((n, n)))[0].astype(np.float32))
    dphi = fc3.DeviceArray.from_host(phi, ctx=ctx)
    for prec in ("bf16x3", "fp32"):
        _, bad = _compare_to_full(ctx, psi32, dphi, n, prec)
        assert bad == 0
    psi32.free()
    dphi.free()


def test_unique_rejects_wrong_buffers(ctx):
    n = 16
    psi = fc3.DeviceArray((n, n, n), np.float32, ctx, zero=True)
    phi = fc3.DeviceArray((n, n), np.float32, ctx, zero=True)
    short = fc3.DeviceArray((fc3.unique_count(n) - 1,), np.float32, ctx)
    with pytest.raises(AssertionError, match="out"):
        fc3.mcc3_unique(psi, phi, out=short)
    for d in (psi, phi, short):
        d.free()
