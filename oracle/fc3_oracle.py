"""Python face of the CPU oracle.  TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import this
module; the product never does.  Loops live in fc3_oracle.c (a restatement of the
reference's CPU algorithm, built by oracle/Makefile into oracle/liboracle.so); this
file adds the dense FP64 linear algebra the reference does on the host:

  dynamical_matrix   src/dynmat.jl:5-33
  get_modes          src/dynmat.jl:123-131 (LAPACK eigen of the Hermitian matrix,
                     the num_rigid smallest |omega^2| set to zero)
  mcc3_fp64          src/mcc.jl:15-17, K3[n,m,l] = sum Psi[i,j,k] phi[i,n] phi[j,m] phi[k,l]
                     evaluated in FP64 as three mode products

Parity pins (tests/test_oracle_*.py): the 12 known-answer values of runtests.jl,
an independent closed-form LJ, sympy/mpmath SW derivatives, FP64 finite
differences of the SW energy, and ASR / permutation-symmetry invariants.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# FC3_ORACLE_LIB names another build of fc3_oracle.c (the sanitizer test's ASan/UBSan one)
_LIB_PATH = os.environ.get("FC3_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)

# cos(theta0) constants of the reference's SW derivative blocks: the
# FastDifferentiation output bakes -0.333333333333 (SW/H3_exec_ijk.jl:39 and the
# other generated blocks); the runtime-built iik block uses params.cos_theta0 = -1/3.
SW_C0_FD = -0.333333333333
SW_C0_IIK = -1.0 / 3.0


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.fco_lj_ifc2.argtypes = [_dp, ctypes.c_int, _dp, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp]
        L.fco_lj_ifc3.argtypes = L.fco_lj_ifc2.argtypes
        L.fco_lj_ifc3_columns.argtypes = [_dp, ctypes.c_int, _dp, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_int, _i64p, _i64p, _dp]
        L.fco_mass_weight3.argtypes = [_dp, ctypes.c_int, _dp]
        L.fco_mass_weight3.restype = None
        L.fco_mass_weight3_d.argtypes = [_dp, ctypes.c_int, ctypes.c_int, _dp]
        L.fco_mass_weight3_d.restype = None
        L.fco_sw_blocks3.argtypes = [_dp, _dp, ctypes.c_double, ctypes.c_double, _dp]
        L.fco_sw_blocks3.restype = None
        L.fco_sw_blocks2.argtypes = [_dp, _dp, ctypes.c_double, _dp]
        L.fco_sw_blocks2.restype = None
        L.fco_sw_pair_blocks.argtypes = [_dp, _dp, _dp, _dp]
        L.fco_sw_pair_blocks.restype = None
        L.fco_sw_energy.argtypes = [_dp, ctypes.c_int, _dp, _dp, ctypes.c_double]
        L.fco_sw_energy.restype = ctypes.c_double
        L.fco_sw_ifc3.argtypes = [_dp, ctypes.c_int, _dp, _dp, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_double, _dp]
        L.fco_sw_ifc2.argtypes = [_dp, ctypes.c_int, _dp, _dp, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, _dp]
        _i32p = ctypes.POINTER(ctypes.c_int32)
        L.fco_lj_ifc3_coo.argtypes = [_dp, ctypes.c_int, _dp, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_long, _i32p, _i32p, _i32p, _dp]
        L.fco_lj_ifc3_coo.restype = ctypes.c_long
        L.fco_lj_ifc3_dparam.argtypes = [_dp, ctypes.c_int, _dp, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_int, _dp]
        L.fco_sw_ifc3_dparam.argtypes = [_dp, ctypes.c_int, _dp, _dp, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp]
        L.fco_ipow.argtypes = [ctypes.c_double, ctypes.c_int]
        L.fco_ipow.restype = ctypes.c_double
        L.fco_num_threads.argtypes = []
        L.fco_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def _sysargs(sys):
    pos = np.ascontiguousarray(sys.positions, dtype=np.float64)
    box = np.ascontiguousarray(sys.box_sizes, dtype=np.float64)
    return pos, box


def lj_second_order(sys, pot, calc) -> np.ndarray:
    """second_order(sys, LJ, AnalyticalCalculator) -> n x n (column-major values)."""
    pos, box = _sysargs(sys)
    n = 3 * sys.n_atoms
    out = np.zeros((n, n), dtype=np.float64, order="F")
    rc = lib().fco_lj_ifc2(_p(pos), sys.n_atoms, _p(box), pot.sigma, pot.epsilon, pot.r_cut,
                           calc.r_cut, calc.tol, _p(out))
    if rc:
        raise AssertionError("Cutoff larger than L/2")
    return out


def lj_third_order(sys, pot, calc) -> np.ndarray:
    """third_order(sys, LJ, AnalyticalCalculator) -> n x n x n (Fortran order)."""
    pos, box = _sysargs(sys)
    n = 3 * sys.n_atoms
    out = np.zeros((n, n, n), dtype=np.float64, order="F")
    rc = lib().fco_lj_ifc3(_p(pos), sys.n_atoms, _p(box), pot.sigma, pot.epsilon, pot.r_cut,
                           calc.r_cut, calc.tol, _p(out))
    if rc:
        raise AssertionError("Cutoff larger than L/2")
    return out


def lj_third_order_columns(sys, pot, calc, jj, kk) -> np.ndarray:
    """Columns Psi[:, jj[c], kk[c]] of third_order(sys, LJ, ...) without the dense tensor."""
    pos, box = _sysargs(sys)
    jj = np.ascontiguousarray(jj, dtype=np.int64)
    kk = np.ascontiguousarray(kk, dtype=np.int64)
    n = 3 * sys.n_atoms
    out = np.zeros((n, len(jj)), dtype=np.float64, order="F")
    rc = lib().fco_lj_ifc3_columns(_p(pos), sys.n_atoms, _p(box), pot.sigma, pot.epsilon,
                                   pot.r_cut, calc.r_cut, calc.tol, len(jj),
                                   jj.ctypes.data_as(_i64p), kk.ctypes.data_as(_i64p), _p(out))
    if rc:
        raise AssertionError("Cutoff larger than L/2")
    return out


def lj_third_order_coo(sys, pot, calc):
    """Nonzero entries (i, j, k, value) of third_order(sys, LJ, AnalyticalCalculator)
    without forming the dense tensor (for sizes a dense host tensor cannot hold)."""
    pos, box = _sysargs(sys)
    N = sys.n_atoms
    cap = N * 80 * 81 + N * 27
    while True:
        oi = np.empty(cap, np.int32)
        oj = np.empty(cap, np.int32)
        ok = np.empty(cap, np.int32)
        ov = np.empty(cap, np.float64)
        _i32p = ctypes.POINTER(ctypes.c_int32)
        m = lib().fco_lj_ifc3_coo(_p(pos), N, _p(box), pot.sigma, pot.epsilon, pot.r_cut,
                                  calc.r_cut, calc.tol, cap, oi.ctypes.data_as(_i32p),
                                  oj.ctypes.data_as(_i32p), ok.ctypes.data_as(_i32p), _p(ov))
        if m == -1:
            raise AssertionError("Cutoff larger than L/2")
        if m == -2:
            cap *= 2
            continue
        return oi[:m].copy(), oj[:m].copy(), ok[:m].copy(), ov[:m].copy()


def coo_mass_weight(coo, masses):
    """mass_weight_third_order! applied to a COO tensor (returns new values)."""
    i, j, k, v = coo
    m = np.asarray(masses, dtype=np.float64)
    return v / np.sqrt((m[i // 3] * m[j // 3]) * m[k // 3])


def mass_weight_third_order(psi: np.ndarray, masses) -> np.ndarray:
    """mass_weight_third_order! on a Fortran-ordered FP64 array (in place), with
    D = size(psi, 1) / length(masses) in 1:3 as the reference asserts."""
    assert psi.flags.f_contiguous and psi.dtype == np.float64
    m = np.ascontiguousarray(masses, dtype=np.float64)
    D, rem = divmod(psi.shape[0], len(m))
    if rem or D not in (1, 2, 3):
        raise AssertionError("D ∈ [1, 2, 3]")
    lib().fco_mass_weight3_d(_p(psi), len(m), D, _p(m))
    return psi


def asr_violations(ifc: np.ndarray, N_atoms: int, D: int, tol: float):
    """asr_satisfied (asr.jl:43-74) as a list of failures: for each slice ifc[:, :, kk]
    (the matrix itself for order 2), row D(i-1)+a and component b (0-based here), the
    sum s = sum(ifc[row, b::D]) fails where s > tol (signed).  Returned in the
    reference's loop order (slice, atom, a, b) as ((slice * n + row) * D + b, s); the
    rule is satisfied when the list is empty.  Pure-Python loop over the sums: small
    tensors only."""
    v = np.asarray(ifc, dtype=np.float64)
    n = v.shape[0]
    assert n == N_atoms * D and all(x == n for x in v.shape)
    slices = [v] if v.ndim == 2 else [v[:, :, kk] for kk in range(n)]
    out = []
    for sl, m in enumerate(slices):
        for row in range(n):
            for b in range(D):
                acc = 0.0
                for x in m[row, b::D]:
                    acc += float(x)
                if acc > tol:
                    out.append(((sl * n + row) * D + b, acc))
    return out


def issymmetric_count(ifc2: np.ndarray) -> int:
    """Pairs r < c with ifc2[r, c] != ifc2[c, r] (issymmetric, ifc.jl:15; NaN counts)."""
    a = np.asarray(ifc2)
    iu = np.triu_indices(a.shape[0], 1)
    return int(np.count_nonzero(~(a[iu] == a.T[iu])))


def sw_blocks3(r9, params, c0_fd=SW_C0_FD, c0_iik=SW_C0_IIK) -> np.ndarray:
    """7 x (3,3,3) blocks iij, ijj, iik, ikk, jjk, jkk, ijk at r_arr = [r_i; r_j; r_k]."""
    r9 = np.ascontiguousarray(r9, dtype=np.float64)
    out = np.zeros(7 * 27, dtype=np.float64)
    lib().fco_sw_blocks3(_p(r9), _p(params.as_array()), c0_fd, c0_iik, _p(out))
    return out.reshape(7, 3, 3, 3, order="C").transpose(0, 3, 2, 1).copy()  # [blk,a,b,c]


def sw_blocks2(r9, params, c0_fd=SW_C0_FD) -> np.ndarray:
    r9 = np.ascontiguousarray(r9, dtype=np.float64)
    out = np.zeros(27, dtype=np.float64)
    lib().fco_sw_blocks2(_p(r9), _p(params.as_array()), c0_fd, _p(out))
    return out.reshape(3, 3, 3, order="C").transpose(0, 2, 1).copy()  # [blk,a,b]


def sw_pair_blocks(r3, params):
    r3 = np.ascontiguousarray(r3, dtype=np.float64)
    h2 = np.zeros(9)
    h3 = np.zeros(27)
    lib().fco_sw_pair_blocks(_p(r3), _p(params.as_array()), _p(h2), _p(h3))
    return h2.reshape(3, 3, order="F"), h3.reshape(3, 3, 3, order="F")


def sw_energy(sys, params, r_cut) -> float:
    pos, box = _sysargs(sys)
    return lib().fco_sw_energy(_p(pos), sys.n_atoms, _p(box), _p(params.as_array()), r_cut)


def sw_third_order(sys, pot, calc, c0_fd=SW_C0_FD, c0_iik=SW_C0_IIK) -> np.ndarray:
    pos, box = _sysargs(sys)
    n = 3 * sys.n_atoms
    out = np.zeros((n, n, n), dtype=np.float64, order="F")
    rc = lib().fco_sw_ifc3(_p(pos), sys.n_atoms, _p(box), _p(pot.params.as_array()),
                           calc.r_cut, calc.tol, c0_fd, c0_iik, _p(out))
    if rc:
        raise AssertionError("For AutoDiff SW silicon force constant cutoff must be less than potential cutoff")
    return out


LJ_PARAMS = {"sigma": 0, "epsilon": 1}
SW_PARAMS = {"epsilon": 0, "sigma": 1, "a": 2, "lambda": 3, "gamma": 4, "A": 6, "B": 7}


def lj_third_order_dparam(sys, pot, calc, name: str, masses=None) -> np.ndarray:
    """d Psi / d p (p = "sigma" or "epsilon") of third_order(sys, LJ, calc) by forward-mode
    dual numbers through the same loops (fco_lj_ifc3_dparam), entries that the tolerance
    pass zeroes set to zero; mass weighted as mass_weight_third_order! when masses given."""
    pos, box = _sysargs(sys)
    n = 3 * sys.n_atoms
    out = np.zeros((n, n, n), dtype=np.float64, order="F")
    rc = lib().fco_lj_ifc3_dparam(_p(pos), sys.n_atoms, _p(box), pot.sigma, pot.epsilon,
                                  pot.r_cut, calc.r_cut, LJ_PARAMS[name], _p(out))
    if rc:
        raise AssertionError("Cutoff larger than L/2")
    out[lj_third_order(sys, pot, calc) == 0.0] = 0.0
    if masses is not None:
        mass_weight_third_order(out, masses)
    return out


def sw_third_order_dparam(sys, pot, calc, name: str, masses=None, c0_fd=SW_C0_FD,
                          c0_iik=SW_C0_IIK) -> np.ndarray:
    """d Psi / d p of third_order(sys, StillingerWeberSilicon, AutoDiffCalculator) for a
    parameter name of SW_PARAMS, by 4-direction hyper-dual numbers (the coordinates'
    three directions and the parameter's) through the same loops (fco_sw_ifc3_dparam);
    tolerance mask and mass weighting as lj_third_order_dparam."""
    pos, box = _sysargs(sys)
    n = 3 * sys.n_atoms
    out = np.zeros((n, n, n), dtype=np.float64, order="F")
    rc = lib().fco_sw_ifc3_dparam(_p(pos), sys.n_atoms, _p(box), _p(pot.params.as_array()),
                                  calc.r_cut, c0_fd, c0_iik, SW_PARAMS[name], _p(out))
    if rc:
        raise AssertionError("For AutoDiff SW silicon force constant cutoff must be less than potential cutoff")
    out[sw_third_order(sys, pot, calc, c0_fd, c0_iik) == 0.0] = 0.0
    if masses is not None:
        mass_weight_third_order(out, masses)
    return out


def sw_second_order(sys, pot, calc, c0_fd=SW_C0_FD) -> np.ndarray:
    pos, box = _sysargs(sys)
    n = 3 * sys.n_atoms
    out = np.zeros((n, n), dtype=np.float64, order="F")
    rc = lib().fco_sw_ifc2(_p(pos), sys.n_atoms, _p(box), _p(pot.params.as_array()),
                           calc.r_cut, calc.tol, c0_fd, _p(out))
    if rc:
        raise AssertionError("For SW silicon force constant cutoff must be less than potential cutoff")
    return out


def dynamical_matrix(ifc2: np.ndarray, masses) -> np.ndarray:
    m = np.repeat(np.asarray(masses, dtype=np.float64), 3)
    return ifc2 / np.sqrt(np.outer(m, m))


def get_modes(dynmat: np.ndarray, num_rigid_translation: int = 3):
    w2, phi = np.linalg.eigh(dynmat)
    idx = np.argsort(np.abs(w2), kind="stable")
    w2 = w2.copy()
    w2[idx[:num_rigid_translation]] = 0.0
    return w2, np.asfortranarray(phi)


def mcc3_fp64(psi: np.ndarray, phi: np.ndarray) -> np.ndarray:
    """K3[n,m,l] = sum_ijk Psi[i,j,k] phi[i,n] phi[j,m] phi[k,l] in FP64 (Fortran order)."""
    P = np.asarray(psi, dtype=np.float64)
    F = np.asarray(phi, dtype=np.float64)
    n = P.shape[0]
    # mode 1: K1[n,j,k] = sum_i phi[i,n] Psi[i,j,k]
    K1 = (F.T @ P.reshape(n, n * n, order="F")).reshape(n, n, n, order="F")
    # mode 2: K2[n,m,k] = sum_j K1[n,j,k] phi[j,m]
    K2 = np.einsum("njk,jm->nmk", K1, F, optimize=True)
    # mode 3: K3[n,m,l] = sum_k K2[n,m,k] phi[k,l]
    K3 = (K2.reshape(n * n, n, order="F") @ F).reshape(n, n, n, order="F")
    return np.asfortranarray(K3)


def mcc3_fp64_slab(psi_sparse, phi: np.ndarray, l: int) -> np.ndarray:
    """K3[:,:,l] in FP64 from a sparse Psi given as (rows, cols, slabs, values) COO,
    for tensors too large to hold densely on the host."""
    i, j, k, v = psi_sparse
    F = np.asarray(phi, dtype=np.float64)
    n = F.shape[0]
    w = v * F[k, l]
    import scipy.sparse as sp
    M = sp.coo_matrix((w, (i, j)), shape=(n, n)).tocsr()
    return F.T @ (M @ F)


# ---- text output (src/io.jl:9-62) ------------------------------------------------------

def julia_float_str(x: float) -> str:
    """print(::Float64) as Julia does it (Base.Ryu.writeshortest with print's defaults),
    restated from Python's shortest round-trip digits (repr): digits d_1..d_m, pt =
    (decimal exponent of d_1) + 1; plain notation for -4 < pt <= 6 (for pt >= m only
    when x is that integer up to 0.05), '.0' after integers; else 'd_1.d_2..d_m' + 'e' +
    str(pt - 1), with 'd_1.0' when m = 1."""
    import math
    from decimal import Decimal
    if x != x:
        return "NaN"
    neg = math.copysign(1.0, x) < 0
    s = "-" if neg else ""
    if math.isinf(x):
        return s + "Inf"
    if x == 0.0:
        return s + "0.0"
    _, digs, exp = Decimal(repr(abs(x))).as_tuple()
    digs = list(digs)
    while len(digs) > 1 and digs[-1] == 0:
        digs.pop()
        exp += 1
    m = len(digs)
    d = "".join(map(str, digs))
    pt = exp + m
    plain = -4 < pt <= 6
    if plain and pt >= m:
        y = 10.0 ** (pt - m)
        if abs(math.fmod(x + 0.05, y) % y - 0.05) > 0.05:
            plain = False
    if plain:
        if pt <= 0:
            return s + "0." + "0" * (-pt) + d
        if pt < m:
            return s + d[:pt] + "." + d[pt:]
        return s + d + "0" * (pt - m) + ".0"
    return s + d[0] + "." + (d[1:] or "0") + "e" + str(pt - 1)


def save_txt_lines(values: np.ndarray, N_atoms: int, D: int):
    """The data lines save_second_order / save_third_order (fmt = :txt) print for a dense
    tensor: one 'i a j b [k c] value' line (1-based) per nonzero, in the order of the
    reference's loops, atoms i, j(, k) outermost and components a, b(, c) innermost
    (io.jl:16-27, 46-61)."""
    v = np.asarray(values)
    order = v.ndim
    idx = np.nonzero(v)
    atoms = [ix // D for ix in idx]
    comps = [ix % D for ix in idx]
    key = np.zeros(len(idx[0]), dtype=np.int64)
    for a in atoms:
        key = key * N_atoms + a
    for c in comps:
        key = key * D + c
    perm = np.argsort(key, kind="stable")
    lines = []
    for t in perm:
        cols = []
        for q in range(order):
            cols += [str(int(atoms[q][t]) + 1), str(int(comps[q][t]) + 1)]
        lines.append(" ".join(cols) + " " + julia_float_str(float(v[tuple(ix[t] for ix in idx)])))
    return lines


# ---- weighted square loss of the differentiation workload --------------------------
# L = sum_e w_e K3_e^2 with w_e ~ U[0, 1) from a counter-based generator, so that device
# and host produce the same weights without storing n^3 of them.  The generator is
# Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2,
# 3", SC11): 10 rounds of two 32x32->64-bit multiplies with the key bumped by the Weyl
# constants between rounds.  Element e takes word e % 4 of the block for counter
# (e / 4 low word, e / 4 high word, 0, 0) under key (seed low word, seed high word), and
# w_e = (word >> 8) 2^-24.  The product path is fc3_weighted_square_loss (adjoint.hip).
_PHILOX_M0, _PHILOX_M1 = 0xD2511F53, 0xCD9E8D57
_PHILOX_W0, _PHILOX_W1 = 0x9E3779B9, 0xBB67AE85


def philox4x32_10(ctr, key):
    """Philox4x32-10 on arrays: ctr (..., 4) and key (..., 2) of uint32 -> (..., 4) uint32."""
    c = [np.asarray(ctr[..., i], dtype=np.uint64) for i in range(4)]
    k0 = np.asarray(key[..., 0], dtype=np.uint64)
    k1 = np.asarray(key[..., 1], dtype=np.uint64)
    mask = np.uint64(0xFFFFFFFF)
    for r in range(10):
        if r > 0:
            k0 = (k0 + np.uint64(_PHILOX_W0)) & mask
            k1 = (k1 + np.uint64(_PHILOX_W1)) & mask
        p0 = np.uint64(_PHILOX_M0) * c[0]
        p1 = np.uint64(_PHILOX_M1) * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & mask
        hi1, lo1 = p1 >> np.uint64(32), p1 & mask
        c = [hi1 ^ c[1] ^ k0, lo1, hi0 ^ c[3] ^ k1, lo0]
    return np.stack(c, axis=-1).astype(np.uint32)


def philox_uniform(seed: int, count: int, start: int = 0) -> np.ndarray:
    """w_e for e in [start, start + count) as float32 in [0, 1) (see above)."""
    e = np.arange(start, start + count, dtype=np.uint64)
    q = e >> np.uint64(2)
    ctr = np.zeros((count, 4), dtype=np.uint32)
    ctr[:, 0] = (q & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    ctr[:, 1] = (q >> np.uint64(32)).astype(np.uint32)
    key = np.empty((count, 2), dtype=np.uint32)
    key[:, 0] = np.uint32(seed & 0xFFFFFFFF)
    key[:, 1] = np.uint32((seed >> 32) & 0xFFFFFFFF)
    words = philox4x32_10(ctr, key)
    w = words[np.arange(count), (e & np.uint64(3)).astype(np.int64)]
    return ((w >> np.uint32(8)).astype(np.float64) * 2.0 ** -24).astype(np.float32)


def weighted_square_loss(k3: np.ndarray, seed: int):
    """(L, dL/dK3) for L = sum_e w_e K3_e^2 over K3 in column-major order (FP64 sum of
    w K3^2 in FP64; the cotangent 2 w K3 rounded to FP32 as the device forms it)."""
    flat = np.asarray(k3, dtype=np.float32).ravel(order="F")
    w = philox_uniform(seed, flat.size)
    L = float(np.sum(w.astype(np.float64) * flat.astype(np.float64) * flat.astype(np.float64)))
    g = (np.float32(2.0) * w) * flat
    return L, g.reshape(k3.shape, order="F")
