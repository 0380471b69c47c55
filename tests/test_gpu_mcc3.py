"""mcc3 on the GPU against an FP64 host contraction (src/mcc.jl:15-17 definition).

Bars (relative Frobenius error vs FP64): FP32 <= 2e-6 (plain FP32 accumulation
over n-term sums), BF16X3 <= 1e-5.  Exact invariants: phi = I gives K3 = Psi and a
permutation phi permutes Psi (FP32).  Sizes cover both BF16X3 kernels (n % 16 == 0
takes the pre-split path, with K padded to a multiple of 16 when n % 16 != 0), single and partial tiles, and, at n = 1536, several
mode-3 row blocks against a closed form (Psi a sum of rank-1 terms).
"""
import ctypes

import numpy as np
import pytest

import fc3
from fc3 import systems as S
from oracle import fc3_oracle as O

pytestmark = pytest.mark.gpu

TOL = {"fp32": 2e-6, "bf16x3": 1e-5}


def rel_fro(a, b):
    return float(np.linalg.norm((a.astype(np.float64) - b).ravel()) / np.linalg.norm(b.ravel()))


def random_case(n, seed):
    rng = np.random.default_rng(seed)
    psi = rng.standard_normal((n, n, n))
    # symmetrise like a force-constant tensor
    psi = (psi + psi.transpose(1, 0, 2) + psi.transpose(0, 2, 1) + psi.transpose(2, 1, 0) +
           psi.transpose(1, 2, 0) + psi.transpose(2, 0, 1)) / 6
    phi, _ = np.linalg.qr(rng.standard_normal((n, n)))
    return np.asfortranarray(psi.astype(np.float32)), np.asfortranarray(phi.astype(np.float32))


@pytest.mark.parametrize("n", [16, 20, 48, 96, 200, 99, 256, 272, 288, 324])
@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_mcc3_random(ctx, n, prec):
    psi, phi = random_case(n, 100 + n)
    ref = O.mcc3_fp64(psi, phi)
    dpsi = fc3.DeviceArray.from_host(psi, ctx=ctx)
    dphi = fc3.DeviceArray.from_host(phi, ctx=ctx)
    k3 = fc3.mcc3(dpsi, dphi, precision=prec).to_host()
    err = rel_fro(k3, ref)
    print(f"mcc3 n={n} {prec}: rel fro err {err:.3e}", flush=True)
    assert err <= TOL[prec]


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
@pytest.mark.parametrize("n", [96, 100, 112])
def test_mcc3_inplace_and_blocked(ctx, prec, n):
    psi, phi = random_case(n, 7)
    ref = O.mcc3_fp64(psi, phi)
    dphi = fc3.DeviceArray.from_host(phi, ctx=ctx)
    dpsi = fc3.DeviceArray.from_host(psi, ctx=ctx)
    blocked = fc3.mcc3(dpsi, dphi, 4, precision=prec)       # host result, block_size = 4
    assert isinstance(blocked, np.ndarray)
    assert rel_fro(blocked, ref) <= TOL[prec]
    fc3.mcc3_(dpsi, dphi, precision=prec)                   # mcc3!: overwrites Psi
    assert rel_fro(dpsi.to_host(), ref) <= TOL[prec]


def test_mcc3_block_size_assertion(ctx):
    n = 96
    psi, phi = random_case(n, 8)
    dpsi = fc3.DeviceArray.from_host(psi, ctx=ctx)
    dphi = fc3.DeviceArray.from_host(phi, ctx=ctx)
    with pytest.raises(AssertionError):
        fc3.mcc3(dpsi, dphi, 50)
    with pytest.raises(AssertionError):
        fc3.mcc3(dpsi, dphi, 0)


def test_mcc3_identity_and_permutation_exact(ctx):
    n = 132
    psi, _ = random_case(n, 9)
    dpsi = fc3.DeviceArray.from_host(psi, ctx=ctx)
    eye = fc3.DeviceArray.from_host(np.eye(n, dtype=np.float32), ctx=ctx)
    assert np.array_equal(fc3.mcc3(dpsi, eye, precision="fp32").to_host(), psi)
    perm = np.random.default_rng(3).permutation(n)
    P = np.zeros((n, n), dtype=np.float32)
    P[perm, np.arange(n)] = 1.0   # phi[:, c] = e_{perm[c]}
    k3 = fc3.mcc3(dpsi, fc3.DeviceArray.from_host(P, ctx=ctx), precision="fp32").to_host()
    assert np.array_equal(k3, psi[np.ix_(perm, perm, perm)])
    # BF16X3 with phi = I reproduces psi to the split's representation error
    k3b = fc3.mcc3(dpsi, eye, precision="bf16x3").to_host()
    assert rel_fro(k3b, psi.astype(np.float64)) <= 1e-5


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_mcc3_lj_fcc222_modes(ctx, prec):
    """The reference test's setting (runtests.jl:62-107): 96 modes of FCC 2x2x2 argon,
    F3 ./= sqrt(m^3), phi from the dynamical matrix; the reference's criterion is
    norm(K3 - K3_ref) <= 1e-6 (its TEP.jld2 data is not available, so K3_ref is the
    FP64 contraction of the oracle's tensors)."""
    sysd = S.fcc(S.ARGON_A, (2, 2, 2), S.ARGON_MASS)
    calc = S.AnalyticalCalculator(1e-12, S.ARGON_LJ.r_cut)
    f3 = O.lj_third_order(sysd, S.ARGON_LJ, calc) / np.sqrt(S.ARGON_MASS ** 3)
    f2 = O.lj_second_order(sysd, S.ARGON_LJ, calc)
    _, phi = O.get_modes(O.dynamical_matrix(f2, sysd.masses))
    ref = O.mcc3_fp64(f3.astype(np.float32), phi.astype(np.float32))
    dpsi = fc3.DeviceArray.from_host(f3, dtype=np.float32, ctx=ctx)
    dphi = fc3.DeviceArray.from_host(phi, dtype=np.float32, ctx=ctx)
    k3 = fc3.mcc3(dpsi, dphi, precision=prec).to_host()
    err = np.linalg.norm((k3.astype(np.float64) - ref).ravel())
    print(f"fcc222 {prec}: |K3 - K3_64| = {err:.3e}, rel {rel_fro(k3, ref):.3e}", flush=True)
    assert err <= 1e-6 or rel_fro(k3, ref) <= TOL[prec]


@pytest.mark.parametrize("inplace", [False, True])
def test_mcc3_bf16x3_rank3_large(ctx, inplace):
    """n = 1536: Psi = sum_t a_t (x) b_t (x) c_t, so K3 = sum_t (phi^T a_t) (x) (phi^T b_t)
    (x) (phi^T c_t) in closed form; Psi is uploaded slab by slab.  The bar is the
    whole-tensor one, 1e-5 relative Frobenius, estimated on 14 output slabs (the two
    ends and 12 drawn at random): sqrt(sum_l |K3_l - ref_l|^2 / sum_l |ref_l|^2)."""
    n, T = 1536, 3
    rng = np.random.default_rng(11)
    a, b, c = (rng.standard_normal((T, n)) for _ in range(3))
    phi, _ = np.linalg.qr(rng.standard_normal((n, n)))
    phi32 = phi.astype(np.float32)
    psi = fc3.DeviceArray((n, n, n), np.float32, ctx)
    slab_bytes = 4 * n * n
    for k in range(n):
        slab = np.asfortranarray(np.einsum("ti,tj,t->ij", a, b, c[:, k]).astype(np.float32))
        fc3._lib.check(ctx.lib.fc3_memcpy_h2d(ctx.h, ctypes.c_void_p(psi.ptr.value + slab_bytes * k),
                                              slab.ctypes.data_as(ctypes.c_void_p), slab_bytes))
    dphi = fc3.DeviceArray.from_host(phi32, ctx=ctx)
    if inplace:
        fc3.mcc3_(psi, dphi, precision="bf16x3")
        out = psi
    else:
        out = fc3.mcc3(psi, dphi, precision="bf16x3")
    p64 = phi32.astype(np.float64)
    A, B, C = a @ p64, b @ p64, c @ p64   # (T, n): phi^T a_t as rows
    slabs = sorted({0, n - 1, *np.random.default_rng(12).choice(n, 12, replace=False).tolist()})
    num = den = 0.0
    for l in slabs:
        got = np.empty((n, n), dtype=np.float32, order="F")
        fc3._lib.check(ctx.lib.fc3_memcpy_d2h(ctx.h, got.ctypes.data_as(ctypes.c_void_p),
                                              ctypes.c_void_p(out.ptr.value + slab_bytes * l),
                                              slab_bytes))
        ref = np.einsum("ti,tj,t->ij", A, B, C[:, l])
        e2 = float(np.sum((got.astype(np.float64) - ref) ** 2))
        r2 = float(np.sum(ref * ref))
        num += e2
        den += r2
        print(f"rank3 n={n} inplace={inplace} slab {l}: rel err {np.sqrt(e2 / r2):.3e}", flush=True)
    err = np.sqrt(num / den)
    print(f"rank3 n={n} inplace={inplace}: rel fro err over {len(slabs)} slabs {err:.3e}", flush=True)
    assert err <= 1e-5


def _sparse(psi):
    i, j, k = np.nonzero(psi)
    return i, j, k, psi[i, j, k].astype(np.float64)


def test_config_fcc444_fp32_block_size_256(ctx):
    """The plain small configuration: LJ argon FCC 4x4x4 (768 DoF), FP64 IFC2/IFC3 with
    tol 1e-12, mass weighting, phi from the host FP64 eigendecomposition, plain FP32
    mcc3 with block_size = 256 (host result, README.md:76), against the FP64 contraction
    of the same FP32 Psi on sampled slabs."""
    sysd = S.fcc(S.ARGON_A, (4, 4, 4), S.ARGON_MASS)
    calc = S.AnalyticalCalculator(1e-12, S.ARGON_LJ.r_cut)
    n = sysd.n_dof
    _, phi = fc3.get_modes(fc3.dynamical_matrix(sysd, S.ARGON_LJ, calc, ctx=ctx))
    phi32 = np.asfortranarray(phi.astype(np.float32))
    psi32 = fc3.DeviceArray((n, n, n), np.float32, ctx)
    fc3.third_order_device(sysd, S.ARGON_LJ, calc, psi32=psi32, masses=sysd.masses, ctx=ctx)
    k3 = fc3.mcc3(psi32, fc3.DeviceArray.from_host(phi32, ctx=ctx), 256, precision="fp32")
    sp = _sparse(psi32.to_host())
    num = den = 0.0
    for l in (0, 1, 2, 3, n // 2, n - 1):
        ref = O.mcc3_fp64_slab(sp, phi32.astype(np.float64), l)
        num += float(np.sum((k3[:, :, l].astype(np.float64) - ref) ** 2))
        den += float(np.sum(ref * ref))
    err = np.sqrt(num / den)
    print(f"FCC 4x4x4 fp32 block 256: rel fro err {err:.3e}", flush=True)
    assert err <= TOL["fp32"]


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_config_sw_diamond333(ctx, prec):
    """The three-body configuration: SW silicon on diamond 3x3x3 (648 DoF), sigma_u =
    0.03 A; IFC2/IFC3 on the GPU (their parity with the oracle is in test_gpu_sw.py),
    mass-weighted FP32 Psi, phi from the host eigendecomposition; mcc3 in both forms
    against the FP64 contraction of the same FP32 Psi."""
    pot = S.StillingerWeberSilicon()
    calc = S.AutoDiffCalculator(1e-12, pot.r_cut)
    sysd = S.diamond(S.SI_A, (3, 3, 3), S.SI_MASS, 0.03, seed=648)
    n = sysd.n_dof
    _, phi = fc3.get_modes(fc3.dynamical_matrix(sysd, pot, calc, ctx=ctx))
    phi32 = np.asfortranarray(phi.astype(np.float32))
    psi32 = fc3.DeviceArray((n, n, n), np.float32, ctx)
    fc3.third_order_device(sysd, pot, calc, psi32=psi32, masses=sysd.masses, ctx=ctx)
    k3 = fc3.mcc3(psi32, fc3.DeviceArray.from_host(phi32, ctx=ctx), precision=prec).to_host()
    ref = O.mcc3_fp64(psi32.to_host(), phi32)
    err = rel_fro(k3, ref)
    print(f"SW diamond 3x3x3 {prec}: rel fro err {err:.3e}", flush=True)
    assert err <= TOL[prec]


@pytest.mark.parametrize("n,block_size,prec", [(1536, 256, "bf16x3"), (1500, 250, "bf16x3"),
                                               (1500, 250, "fp32")])
def test_mcc3_blocked_bounded_memory_bitwise(n, block_size, prec):
    """mcc3(Psi, phi, block_size) (src/mcc.jl:89-120) exists to bound device memory.  At
    n = 1500/1536 (K3 = 13.5/14.5 GB) the call may grow device memory by at most two blocks
    of output slabs plus 3 GiB of scratch, well below the whole K3, and its host result
    must equal the unblocked device mcc3 bit for bit (every slab runs the same per-element
    arithmetic).  Several blocks run (6 at both sizes), and n = 1500 (n % 16 != 0) takes
    the separate split-operand scratch.  Psi is a rank-3 tensor uploaded slab by slab."""
    other = fc3.Context(0)   # its close() frees the arrays below
    try:
        rng = np.random.default_rng(n + block_size)
        a, b, c = (rng.standard_normal((3, n)) for _ in range(3))
        phi32 = np.linalg.qr(rng.standard_normal((n, n)))[0].astype(np.float32)
        psi = fc3.DeviceArray((n, n, n), np.float32, other)
        slab_bytes = 4 * n * n
        for k in range(n):
            s = np.asfortranarray(np.einsum("ti,tj,t->ij", a, b, c[:, k]).astype(np.float32))
            fc3._lib.check(other.lib.fc3_memcpy_h2d(other.h, ctypes.c_void_p(psi.ptr.value + slab_bytes * k),
                                                    s.ctypes.data_as(ctypes.c_void_p), slab_bytes))
        dphi = fc3.DeviceArray.from_host(phi32, ctx=other)
        other.synchronize()
        free0 = other.memory()[0]
        host = fc3.mcc3(psi, dphi, block_size, precision=prec)
        growth = free0 - other.memory()[0]
        bmem = max(1, (4 << 30) // slab_bytes)
        bs = min(n, block_size if block_size >= bmem else (bmem // block_size) * block_size)
        print(f"blocked n={n} block_size={block_size} {prec}: {(n + bs - 1) // bs} blocks of {bs} "
              f"slabs, device growth {growth / 2**30:.2f} GiB (K3 {4 * n**3 / 2**30:.2f} GiB)",
              flush=True)
        assert (n + bs - 1) // bs >= 2
        # two output blocks, the split mode-1 operand of a block when n % 16 != 0 (its K is
        # padded, so it cannot live in the output block), and 3 GiB of bounded scratch
        kp = -(-n // 16) * 16
        split = bs * 4 * n * kp if (prec == "bf16x3" and kp != n) else 0
        assert growth <= 2 * bs * slab_bytes + split + (3 << 30)
        assert growth < 0.7 * 4 * n ** 3
        dev = fc3.mcc3(psi, dphi, precision=prec).to_host()
        assert np.array_equal(host, dev)
        # and the result itself, on a few slabs, against the rank-3 closed form
        p64 = phi32.astype(np.float64)
        A, B, C = a @ p64, b @ p64, c @ p64
        num = den = 0.0
        for l in (0, bs - 1, bs, n - 1):
            ref = np.einsum("ti,tj,t->ij", A, B, C[:, l])
            num += float(np.sum((host[:, :, l].astype(np.float64) - ref) ** 2))
            den += float(np.sum(ref * ref))
        assert (num / den) ** 0.5 <= TOL[prec]
    finally:
        other.close()


@pytest.mark.parametrize("cells", [(4, 4, 4), (5, 5, 5)])
def test_mcc3_zero_block_skipping(ctx, cells, monkeypatch):
    """BF16X3 mode 3 multiplies only the 256-row x 16-K blocks of Psi that hold a nonzero
    (an LJ tensor is mostly zero blocks: 9% of them are nonzero at 768 DoF, 3% at 1500).
    The skipped products are exactly zero: with n % 16 != 0 (1500; U in scratch, dense
    mode 1) and in the blocked form, K3 must equal, bit for bit, the result with every
    block multiplied (FC3_MCC3_DENSE=1).  For n % 16 == 0 (768) mode 1 runs over each
    column group's gathered rows i (those where Psi's (i, j) pattern is not empty), whose
    16-wide K blocks group the sum over i differently from the dense form: there the two
    must agree to FP32 summation-order rounding (relative Frobenius <= 4e-6, against about
    8e-6 of BF16X3 error; a wrong row or block would differ at order 1), and both must
    agree with the FP64 contraction."""
    sysd = S.fcc(S.ARGON_A, cells, S.ARGON_MASS, 0.05, seed=3)
    calc = S.AnalyticalCalculator(1e-12, S.ARGON_LJ.r_cut)
    n = sysd.n_dof
    _, phi = fc3.get_modes(fc3.dynamical_matrix(sysd, S.ARGON_LJ, calc, ctx=ctx))
    dphi = fc3.DeviceArray.from_host(phi, dtype=np.float32, ctx=ctx)
    psi = fc3.DeviceArray((n, n, n), np.float32, ctx)
    fc3.third_order_device(sysd, S.ARGON_LJ, calc, psi32=psi, masses=sysd.masses, ctx=ctx)
    monkeypatch.setenv("FC3_MCC3_DENSE", "1")
    dense = fc3.mcc3(psi, dphi, precision="bf16x3").to_host()
    monkeypatch.delenv("FC3_MCC3_DENSE")
    sparse = fc3.mcc3(psi, dphi, precision="bf16x3").to_host()
    if n % 16 == 0:
        diff = float(np.linalg.norm((sparse.astype(np.float64) - dense).ravel()) /
                     np.linalg.norm(dense.astype(np.float64).ravel()))
        print(f"gathered mode 1 vs dense FCC {cells}: rel fro diff {diff:.2e}", flush=True)
        assert diff <= 4e-6
    else:
        assert np.array_equal(sparse, dense)
    blocked = fc3.mcc3(psi, dphi, n // 4, precision="bf16x3")
    assert np.array_equal(blocked, dense)
    # and the value against the FP64 contraction on a few slabs
    coo = O.lj_third_order_coo(sysd, S.ARGON_LJ, calc)
    v32 = O.coo_mass_weight(coo, sysd.masses).astype(np.float32).astype(np.float64)
    p64 = phi.astype(np.float32).astype(np.float64)
    num = den = 0.0
    for l in (n // 3, n // 2, n - 1):
        ref = O.mcc3_fp64_slab((coo[0], coo[1], coo[2], v32), p64, l)
        num += float(np.sum((sparse[:, :, l].astype(np.float64) - ref) ** 2))
        den += float(np.sum(ref * ref))
    print(f"zero-block skipping FCC {cells}: rel err {(num / den) ** 0.5:.2e}",
          flush=True)
    assert (num / den) ** 0.5 <= TOL["bf16x3"]
    psi.free()
    dphi.free()
