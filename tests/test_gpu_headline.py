"""The headline configuration at full size: LJ argon on FCC 6x6x6 (864 atoms,
2592 DoF), seeded Gaussian displacements sigma_u = 0.05 A (PCG64 seed 2592).

Psi (FP64 blocks, mass weighted, rounded to FP32) is assembled on the GPU, phi comes
from the host FP64 eigendecomposition of the dynamical matrix (src/dynmat.jl:123-131),
and mcc3 (src/mcc.jl:29-51) runs in both forms.  K3 is compared with the FP64
contraction of the oracle's COO tensor (the same FP32 values) on 18 output slabs
spread over the whole spectrum; the three rigid-translation modes are left out, since
their slabs vanish by the acoustic sum rule and would say nothing about the error.

Bars, relative Frobenius error over the sampled slabs,
sqrt(sum_l |K3_l - ref_l|^2 / sum_l |ref_l|^2):  FP32 <= 2e-6,  BF16X3 <= 1e-5.
The rigid slabs themselves are checked to be negligible against the others.
"""
import ctypes

import numpy as np
import pytest

import fc3
from fc3 import systems as S
from oracle import fc3_oracle as O

pytestmark = pytest.mark.gpu

TOL = {"fp32": 2e-6, "bf16x3": 1e-5}
CALC = S.AnalyticalCalculator(1e-12, S.ARGON_LJ.r_cut)


@pytest.fixture(scope="module")
def headline(ctx):
    sysd = S.fcc(S.ARGON_A, (6, 6, 6), S.ARGON_MASS, 0.05, seed=2592)
    n = sysd.n_dof
    assert n == 2592
    w2, phi = fc3.get_modes(fc3.dynamical_matrix(sysd, S.ARGON_LJ, CALC, ctx=ctx))
    rigid = np.argsort(np.abs(w2), kind="stable")[:3]
    phi32 = np.asfortranarray(phi.astype(np.float32))
    dphi = fc3.DeviceArray.from_host(phi32, ctx=ctx)
    psi32 = fc3.DeviceArray((n, n, n), np.float32, ctx)
    fc3.third_order_device(sysd, S.ARGON_LJ, CALC, psi32=psi32, masses=sysd.masses, ctx=ctx)
    coo = O.lj_third_order_coo(sysd, S.ARGON_LJ, CALC)
    v32 = O.coo_mass_weight(coo, sysd.masses).astype(np.float32).astype(np.float64)
    coo32 = (coo[0], coo[1], coo[2], v32)
    phi64 = phi32.astype(np.float64)
    cand = np.unique(np.linspace(0, n - 1, 24).round().astype(int))
    slabs = [int(l) for l in cand if l not in set(rigid.tolist())][:18]
    assert len(slabs) >= 16
    ref = {l: O.mcc3_fp64_slab(coo32, phi64, l) for l in slabs}
    ref_rigid = {int(l): O.mcc3_fp64_slab(coo32, phi64, int(l)) for l in rigid}
    print(f"headline: n={n}, rigid modes {sorted(rigid.tolist())}, slabs {slabs}", flush=True)
    k3 = fc3.DeviceArray((n, n, n), np.float32, ctx)
    yield dict(n=n, psi32=psi32, dphi=dphi, k3=k3, slabs=slabs, ref=ref, ref_rigid=ref_rigid)
    for d in (psi32, dphi, k3):
        d.free()


def _slab(ctx, k3, n, l):
    buf = np.empty((n, n), dtype=np.float32, order="F")
    fc3._lib.check(ctx.lib.fc3_memcpy_d2h(ctx.h, buf.ctypes.data_as(ctypes.c_void_p),
                                          ctypes.c_void_p(k3.ptr.value + 4 * n * n * l), 4 * n * n))
    return buf.astype(np.float64)


@pytest.mark.parametrize("prec", ["bf16x3", "fp32"])
def test_mcc3_headline_fcc666(ctx, headline, prec):
    h = headline
    n = h["n"]
    fc3.mcc3(h["psi32"], h["dphi"], precision=prec, out=h["k3"])
    num = den = 0.0
    worst = 0.0
    for l in h["slabs"]:
        got = _slab(ctx, h["k3"], n, l)
        ref = h["ref"][l]
        e2 = float(np.sum((got - ref) ** 2))
        r2 = float(np.sum(ref * ref))
        num += e2
        den += r2
        worst = max(worst, (e2 / r2) ** 0.5)
    err = (num / den) ** 0.5
    print(f"headline {prec}: rel fro err {err:.3e} over {len(h['slabs'])} slabs "
          f"(worst slab {worst:.3e})", flush=True)
    assert err <= TOL[prec]
    # rigid-translation slabs: K3 vanishes there up to rounding, on the GPU as in FP64
    scale = (den / len(h["slabs"])) ** 0.5
    for l, ref in h["ref_rigid"].items():
        got = _slab(ctx, h["k3"], n, l)
        assert np.linalg.norm(ref) <= 1e-6 * scale
        assert np.linalg.norm(got) <= 1e-4 * scale
