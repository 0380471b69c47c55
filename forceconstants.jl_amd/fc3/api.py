"""Host-side mirror of ForceConstants.jl's third-order API over the HIP library.

Same names, argument meaning and error behaviour as the reference (Julia `!` is
spelled with a trailing underscore):

  second_order(sys, pot, calc)              src/ifc.jl:3-18
  third_order(sys, pot, calc)               src/ifc.jl:20-35
  mass_weight_third_order_(ifc3, masses)    src/analytical/third_order_analytical.jl:185-214
  dynamical_matrix(sys, pot, calc)          src/dynmat.jl:5-33
  get_modes(dynmat, num_rigid_translation)  src/dynmat.jl:123-131
  mcc3(psi, phi) / mcc3(psi, phi, block_size) / mcc3_(psi, phi)   src/mcc.jl:29,60,89

GPU tensors are `DeviceArray`s (a device pointer with Fortran-order shape), the
counterpart of CuArray in the reference's examples (README.md:72-76).  All compute
goes through libfc3.so; nothing here computes force constants on the CPU.
"""
from __future__ import annotations

import ctypes
import weakref
from dataclasses import dataclass
from typing import Optional, Tuple

import numpy as np

from . import _lib
from ._lib import FC3_PREC_BF16X3, FC3_PREC_FP32, check
from .systems import (AnalyticalCalculator, AutoDiffCalculator, LJ, StillingerWeberSilicon,
                      SuperCellSystem)

_dp = ctypes.POINTER(ctypes.c_double)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


class Context:
    """One fc3 handle: a device, a HIP stream and a workspace."""

    _default: Optional["Context"] = None

    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        check(self.lib.fc3_create(ctypes.byref(h), int(device)))
        self.h = h
        self.device = device
        self._arrays = weakref.WeakSet()   # live DeviceArrays allocated on this handle

    @classmethod
    def default(cls) -> "Context":
        if cls._default is None:
            cls._default = Context(0)
        return cls._default

    def close(self):
        """Free the device arrays still alive on this handle, then the handle (its stream
        and workspace).  Arrays freed here read as freed afterwards (ptr None)."""
        if getattr(self, "h", None):
            for a in list(self._arrays):
                a.free()
            self.lib.fc3_destroy(self.h)
            self.h = None

    def synchronize(self):
        check(self.lib.fc3_synchronize(self.h))

    def release_workspace(self):
        """Return the handle's workspace to the device (it regrows on demand)."""
        check(self.lib.fc3_release_workspace(self.h))

    def timer_start(self):
        check(self.lib.fc3_timer_start(self.h))

    def timer_stop(self) -> float:
        ms = ctypes.c_float()
        check(self.lib.fc3_timer_stop(self.h, ctypes.byref(ms)))
        return float(ms.value)

    def memory(self) -> Tuple[int, int]:
        f, t = ctypes.c_size_t(), ctypes.c_size_t()
        check(self.lib.fc3_device_memory(self.h, ctypes.byref(f), ctypes.byref(t)))
        return int(f.value), int(t.value)


class DeviceArray:
    """A device allocation with a Fortran-order shape and a dtype."""

    def __init__(self, shape, dtype, ctx: Optional[Context] = None, zero: bool = False):
        self.ctx = ctx or Context.default()
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        if not self.ctx.h:
            raise AssertionError("DeviceArray on a closed Context")
        p = ctypes.c_void_p()
        check(self.ctx.lib.fc3_device_alloc(self.ctx.h, ctypes.byref(p), self.nbytes))
        self.ptr = p
        self.ctx._arrays.add(self)
        if zero:
            check(self.ctx.lib.fc3_memset_zero(self.ctx.h, self.ptr, self.nbytes))

    @classmethod
    def from_host(cls, a: np.ndarray, dtype=None, ctx: Optional[Context] = None) -> "DeviceArray":
        a = np.asfortranarray(a, dtype=dtype or a.dtype)
        d = cls(a.shape, a.dtype, ctx)
        check(d.ctx.lib.fc3_memcpy_h2d(d.ctx.h, d.ptr, a.ctypes.data_as(ctypes.c_void_p), d.nbytes))
        return d

    def to_host(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype, order="F")
        check(self.ctx.lib.fc3_memcpy_d2h(self.ctx.h, out.ctypes.data_as(ctypes.c_void_p), self.ptr,
                                          self.nbytes))
        return out

    def free(self):
        """Release the allocation (after the handle's queued work).  Idempotent; an array
        whose Context was closed was already freed by close()."""
        if self.ptr and self.ctx.h:
            self.ctx.lib.fc3_device_free(self.ctx.h, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


@dataclass
class DenseForceConstants:
    """types.jl:76-90: values, units, tol.  `values` is a host array (Fortran order)."""

    values: np.ndarray
    units: str
    tol: float

    def __getitem__(self, idx):
        return self.values[idx]

    @property
    def shape(self):
        return self.values.shape


def _check_buf(arr, shape, dtype, name: str, ctx: Optional[Context] = None) -> DeviceArray:
    """AssertionError unless `arr` is a live DeviceArray of exactly this shape and dtype
    (on ctx's device when ctx is given).  The kernels write through the pointer as given,
    so a smaller or differently typed buffer would be written out of bounds; every
    caller-supplied buffer passes through here before any library call."""
    if not isinstance(arr, DeviceArray) or not arr.ptr:
        raise AssertionError(f"{name} must be a live DeviceArray")
    shape = tuple(int(x) for x in shape)
    if arr.shape != shape or arr.dtype != np.dtype(dtype):
        raise AssertionError(f"{name} must be {shape} {np.dtype(dtype).name}, "
                             f"got {arr.shape} {arr.dtype.name}")
    if ctx is not None and arr.ctx.device != ctx.device:
        raise AssertionError(f"{name} lives on device {arr.ctx.device}, the call on {ctx.device}")
    return arr


def _check_masses(masses, n_atoms: int, name: str = "masses") -> np.ndarray:
    """Host masses as n_atoms contiguous doubles (the C side reads exactly that many)."""
    m = np.ascontiguousarray(masses, dtype=np.float64)
    if m.shape != (int(n_atoms),):
        raise AssertionError(f"{name} must hold one value per atom ({n_atoms}), got shape {m.shape}")
    return m


def _units(pot, order: int) -> str:
    e, l = ("kcal/mol", "Angstrom") if isinstance(pot, LJ) else ("eV", "Angstrom")
    return f"{e} / {l}^{order}"


def _sys_args(sys: SuperCellSystem):
    pos = np.ascontiguousarray(sys.positions, dtype=np.float64)
    box = np.ascontiguousarray(sys.box_sizes, dtype=np.float64)
    return pos, box


def second_order_device(sys: SuperCellSystem, pot, calc, out: Optional[DeviceArray] = None,
                        ctx: Optional[Context] = None) -> DeviceArray:
    ctx = ctx or Context.default()
    n = sys.n_dof
    out = _check_buf(out, (n, n), np.float64, "out", ctx) if out is not None else \
        DeviceArray((n, n), np.float64, ctx)
    pos, box = _sys_args(sys)
    if isinstance(pot, LJ) and isinstance(calc, AnalyticalCalculator):
        st = ctx.lib.fc3_second_order_lj(ctx.h, _ptr(pos), sys.n_atoms, _ptr(box), pot.sigma,
                                         pot.epsilon, pot.r_cut, calc.r_cut, calc.tol, out.ptr)
    elif isinstance(pot, StillingerWeberSilicon) and isinstance(calc, AutoDiffCalculator):
        prm = pot.params.as_abi_array()
        st = ctx.lib.fc3_second_order_sw(ctx.h, _ptr(pos), sys.n_atoms, _ptr(box), _ptr(prm),
                                         calc.r_cut, calc.tol, out.ptr)
    else:
        raise NotImplementedError(f"second_order for {type(pot).__name__} with {type(calc).__name__}")
    check(st)
    return out


def _as_device_f64(ifc, order: int, ctx: Context):
    """(DeviceArray, temporary?) for a DeviceArray, DenseForceConstants or host array."""
    if isinstance(ifc, DeviceArray):
        n = ifc.shape[0]
        return _check_buf(ifc, (n,) * order, np.float64, "ifc", ctx), False
    vals = ifc.values if isinstance(ifc, DenseForceConstants) else ifc
    vals = np.asarray(vals)
    if vals.ndim != order or len(set(vals.shape)) != 1:
        raise AssertionError(f"ifc must be n^{order}, got shape {vals.shape}")
    return DeviceArray.from_host(np.asfortranarray(vals, dtype=np.float64), ctx=ctx), True


def asr_satisfied(ifc, N_atoms: int, D: int, tol: float, verbose: bool = False,
                  ctx: Optional[Context] = None) -> bool:
    """asr_satisfied(ifc, N_atoms, D, tol; verbose) of asr.jl:43-74 for second- (n x n)
    or third-order (n x n x n) force constants, on the GPU (fc3_asr_satisfied): false
    where some sum(ifc[D(i-1)+a, b:D:end(, kk)]) > tol.  verbose prints the first failing
    set in the reference's words and the number of failing sets."""
    ctx = ctx or Context.default()
    order = len(ifc.shape)
    if order not in (2, 3):
        raise AssertionError("asr_satisfied takes second- or third-order force constants")
    if ifc.shape[0] != N_atoms * D:
        raise AssertionError("Incorrect dimensions")
    d, tmp = _as_device_f64(ifc, order, ctx)
    cnt, first, fsum = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_double()
    try:
        check(ctx.lib.fc3_asr_satisfied(ctx.h, d.ptr, order, int(N_atoms), int(D), float(tol),
                                        ctypes.byref(cnt), ctypes.byref(first),
                                        ctypes.byref(fsum)))
    finally:
        if tmp:
            d.free()
    if verbose and cnt.value:
        n = N_atoms * D
        key = first.value
        b, rs = key % D, key // D
        row, sl = rs % n, rs // n
        if order == 3:
            print(f"=======\n On k = {sl // D + 1}, γ = {sl % D + 1} \n=======")
        print(f"Set i = {row // D + 1}, α = {row % D + 1}, β = {b + 1} does not satisfy ASR. "
              f"Sum was: {format_float(fsum.value)}")
        print(f"({cnt.value} sets do not satisfy ASR)")
    return cnt.value == 0


def issymmetric(ifc2, ctx: Optional[Context] = None) -> bool:
    """issymmetric(Phi) (ifc.jl:15) on the GPU: Phi[r, c] == Phi[c, r] for all r, c."""
    ctx = ctx or Context.default()
    d, tmp = _as_device_f64(ifc2, 2, ctx)
    cnt = ctypes.c_longlong()
    try:
        check(ctx.lib.fc3_issymmetric2(ctx.h, d.ptr, d.shape[0], ctypes.byref(cnt)))
    finally:
        if tmp:
            d.free()
    return cnt.value == 0


def second_order(sys: SuperCellSystem, pot, calc, n_threads: Optional[int] = None,
                 check_asr: bool = False, check_symmetry: bool = False,
                 ctx: Optional[Context] = None) -> DenseForceConstants:
    """second_order(sys, pot, calc; n_threads, check_asr, check_symmetry) ->
    DenseForceConstants{2} (host values), ifc.jl:3-18.  n_threads is accepted for the
    reference's signature; the work runs on the GPU."""
    ctx = ctx or Context.default()
    d = second_order_device(sys, pot, calc, ctx=ctx)
    try:
        if check_asr and not asr_satisfied(d, sys.n_atoms, 3, calc.tol, ctx=ctx):
            raise AssertionError("ASR not satisfied for second order force constants, "
                                 "please report this")
        if check_symmetry and not issymmetric(d, ctx=ctx):
            raise AssertionError("IFC were not symmetric, please report this")
        vals = d.to_host()
    finally:
        d.free()
    return DenseForceConstants(vals, _units(pot, 2), calc.tol)


def third_order_device(sys: SuperCellSystem, pot, calc, psi64: Optional[DeviceArray] = None,
                       psi32: Optional[DeviceArray] = None, masses=None,
                       ctx: Optional[Context] = None) -> None:
    """Fill psi64 (FP64 third_order result) and/or psi32 (FP32, mass weighted when
    `masses` is given) on the device."""
    ctx = ctx or Context.default()
    pos, box = _sys_args(sys)
    n = sys.n_dof
    if psi64 is None and psi32 is None:
        raise AssertionError("third_order_device: psi64 and psi32 are both None")
    if psi64 is not None:
        _check_buf(psi64, (n, n, n), np.float64, "psi64", ctx)
    if psi32 is not None:
        _check_buf(psi32, (n, n, n), np.float32, "psi32", ctx)
    m = None if masses is None else _check_masses(masses, sys.n_atoms)
    mp = _ptr(m) if m is not None else None
    p64 = psi64.ptr if psi64 is not None else None
    p32 = psi32.ptr if psi32 is not None else None
    if isinstance(pot, LJ) and isinstance(calc, AnalyticalCalculator):
        st = ctx.lib.fc3_third_order_lj(ctx.h, _ptr(pos), sys.n_atoms, _ptr(box), pot.sigma,
                                        pot.epsilon, pot.r_cut, calc.r_cut, calc.tol, mp, p64, p32)
    elif isinstance(pot, StillingerWeberSilicon) and isinstance(calc, AutoDiffCalculator):
        prm = pot.params.as_abi_array()
        st = ctx.lib.fc3_third_order_sw(ctx.h, _ptr(pos), sys.n_atoms, _ptr(box), _ptr(prm),
                                        calc.r_cut, calc.tol, mp, p64, p32)
    else:
        raise NotImplementedError(f"third_order for {type(pot).__name__} with {type(calc).__name__}")
    check(st)


def third_order_mcc3(sys: SuperCellSystem, pot, calc, phi: DeviceArray, masses=None,
                     out: Optional[DeviceArray] = None, ctx: Optional[Context] = None
                     ) -> DeviceArray:
    """K3 = mcc3(mass-weighted FP32 third_order(sys, pot, calc), phi) without forming Psi:
    the mode-1 product over Psi's nonzeros (LJ: straight from the pair blocks,
    fc3_third_order_mcc3_lj; SW: from the assembly's reduced block table,
    fc3_third_order_mcc3_sw), modes 2 and 3 in BF16x3.  LJ with the analytical calculator
    or SW with the autodiff one; 3N % 4 == 0."""
    ctx = ctx or Context.default()
    is_lj = isinstance(pot, LJ) and isinstance(calc, AnalyticalCalculator)
    is_sw = isinstance(pot, StillingerWeberSilicon) and isinstance(calc, AutoDiffCalculator)
    if not (is_lj or is_sw):
        raise NotImplementedError(f"third_order_mcc3 for {type(pot).__name__} with {type(calc).__name__}")
    n = sys.n_dof
    _check_buf(phi, (n, n), np.float32, "third_order_mcc3: phi", ctx)
    if out is not None:
        _check_buf(out, (n, n, n), np.float32, "third_order_mcc3: out", ctx)
    k3 = out if out is not None else DeviceArray((n, n, n), np.float32, ctx)
    pos, box = _sys_args(sys)
    m = None if masses is None else _check_masses(masses, sys.n_atoms)
    mp = _ptr(m) if m is not None else None
    if is_lj:
        check(ctx.lib.fc3_third_order_mcc3_lj(ctx.h, _ptr(pos), sys.n_atoms, _ptr(box), pot.sigma,
                                              pot.epsilon, pot.r_cut, calc.r_cut, calc.tol, mp,
                                              phi.ptr, k3.ptr))
    else:
        prm = pot.params.as_abi_array()
        check(ctx.lib.fc3_third_order_mcc3_sw(ctx.h, _ptr(pos), sys.n_atoms, _ptr(box), _ptr(prm),
                                              calc.r_cut, calc.tol, mp, phi.ptr, k3.ptr))
    return k3


def third_order(sys: SuperCellSystem, pot, calc, n_threads: Optional[int] = None,
                check_asr: bool = False, ctx: Optional[Context] = None) -> DenseForceConstants:
    """third_order(sys, pot, calc; n_threads, check_asr) -> DenseForceConstants{3} (host
    values), ifc.jl:20-35.  n_threads is accepted for the reference's signature."""
    ctx = ctx or Context.default()
    n = sys.n_dof
    d = DeviceArray((n, n, n), np.float64, ctx)
    try:
        third_order_device(sys, pot, calc, psi64=d, ctx=ctx)
        if check_asr and not asr_satisfied(d, sys.n_atoms, 3, calc.tol, ctx=ctx):
            raise AssertionError("ASR not satisfied for third order force constants, "
                                 "please report this")
        vals = d.to_host()
    finally:
        d.free()
    return DenseForceConstants(vals, _units(pot, 3), calc.tol)


def mass_weight_third_order_(ifc3, masses, ctx: Optional[Context] = None):
    """mass_weight_third_order!(Psi, masses): divides nonzero entries by
    sqrt(m_i m_j m_k) in place.  Accepts a DeviceArray (FP64) or a
    DenseForceConstants whose host values are weighted on the GPU."""
    ctx = ctx or Context.default()
    m = np.ascontiguousarray(masses, dtype=np.float64).reshape(-1)
    n = ifc3.shape[0]
    # D = Int(N_modes / length(masses)); @assert D ∈ [1,2,3] (third_order_analytical.jl:186-190)
    dim, rem = divmod(n, max(len(m), 1))
    if len(m) == 0 or rem or dim not in (1, 2, 3):
        raise AssertionError("D ∈ [1, 2, 3]")
    if isinstance(ifc3, DeviceArray):
        _check_buf(ifc3, (n, n, n), np.float64, "mass_weight_third_order!: Psi", ctx)
        check(ctx.lib.fc3_mass_weight_third_order(ctx.h, ifc3.ptr, len(m), dim, _ptr(m)))
        return ifc3
    if tuple(ifc3.shape) != (n, n, n):
        raise AssertionError("mass_weight_third_order!: Psi must be n x n x n")
    d = DeviceArray.from_host(np.asfortranarray(ifc3.values, dtype=np.float64), ctx=ctx)
    check(ctx.lib.fc3_mass_weight_third_order(ctx.h, d.ptr, len(m), dim, _ptr(m)))
    ifc3.values = d.to_host()
    d.free()
    ifc3.units = ifc3.units + " / amu"
    return ifc3


def dynamical_matrix(sys: SuperCellSystem, pot, calc, ctx: Optional[Context] = None
                     ) -> DenseForceConstants:
    """dynmat.jl:5-33: IFC2 on the GPU, divided by sqrt(m_i m_j) on the host."""
    ifc2 = second_order(sys, pot, calc, ctx=ctx)
    m = np.repeat(sys.masses, 3)
    vals = ifc2.values / np.sqrt(np.outer(m, m))
    return DenseForceConstants(np.asfortranarray(vals), ifc2.units + " / amu", 0.0)


def modes_device(sys: SuperCellSystem, pot, calc, num_rigid_translation: int = 3,
                 fp64: bool = False, out: Optional[DeviceArray] = None,
                 ctx: Optional[Context] = None):
    """dynamical_matrix + get_modes on the GPU (fc3_modes_lj / fc3_modes_sw: LJ with the
    analytical calculator, SW with the autodiff one): (freqs_sq host array, device
    eigenvectors FP32, or FP64 if fp64), written into `out` when given."""
    ctx = ctx or Context.default()
    lj = isinstance(pot, LJ) and isinstance(calc, AnalyticalCalculator)
    sw = isinstance(pot, StillingerWeberSilicon) and isinstance(calc, AutoDiffCalculator)
    if not (lj or sw):
        raise NotImplementedError(f"modes_device for {type(pot).__name__} with {type(calc).__name__}")
    n = sys.n_dof
    pos, box = _sys_args(sys)
    m = _check_masses(sys.masses, sys.n_atoms)
    w = np.empty(n, dtype=np.float64)
    dt = np.float64 if fp64 else np.float32
    if out is not None:
        _check_buf(out, (n, n), dt, "modes_device: out", ctx)
    phi = out if out is not None else DeviceArray((n, n), dt, ctx)
    p64, p32 = (phi.ptr, None) if fp64 else (None, phi.ptr)
    if lj:
        check(ctx.lib.fc3_modes_lj(ctx.h, _ptr(pos), sys.n_atoms, _ptr(box), pot.sigma,
                                   pot.epsilon, pot.r_cut, calc.r_cut, calc.tol, _ptr(m),
                                   num_rigid_translation, _ptr(w), p64, p32))
    else:
        prm = pot.params.as_abi_array()
        check(ctx.lib.fc3_modes_sw(ctx.h, _ptr(pos), sys.n_atoms, _ptr(box), _ptr(prm), calc.r_cut,
                                   calc.tol, _ptr(m), num_rigid_translation, _ptr(w), p64, p32))
    return w, phi


def get_modes(dynmat, num_rigid_translation: int = 3):
    """dynmat.jl:123-131: FP64 LAPACK eigendecomposition of the Hermitian matrix;
    the num_rigid_translation eigenvalues of smallest |w^2| are set to zero."""
    D = dynmat.values if isinstance(dynmat, DenseForceConstants) else np.asarray(dynmat)
    w2, phi = np.linalg.eigh(D)
    idx = np.argsort(np.abs(w2), kind="stable")
    w2 = w2.copy()
    w2[idx[:num_rigid_translation]] = 0.0
    return w2, np.asfortranarray(phi)


_PREC = {"fp32": FC3_PREC_FP32, "bf16x3": FC3_PREC_BF16X3}


def _prec(precision) -> int:
    if isinstance(precision, str):
        return _PREC[precision.lower()]
    return int(precision)


def mcc3(psi: DeviceArray, phi: DeviceArray, block_size: Optional[int] = None,
         precision="fp32", out: Optional[DeviceArray] = None):
    """mcc3(Psi, phi) -> device K3;  mcc3(Psi, phi, block_size) -> host K3."""
    ctx = psi.ctx
    n = phi.shape[0]
    if psi.dtype != np.float32 or phi.dtype != np.float32:
        raise TypeError("mcc3 takes Float32 device arrays")
    _check_buf(psi, (n, n, n), np.float32, "mcc3: Psi")
    _check_buf(phi, (n, n), np.float32, "mcc3: phi", ctx)
    p = _prec(precision)
    if out is not None:
        _check_buf(out, (n, n, n), np.float32, "mcc3: out", ctx)
    if block_size is None:
        out = out or DeviceArray((n, n, n), np.float32, ctx)
        check(ctx.lib.fc3_mcc3(ctx.h, psi.ptr, phi.ptr, n, p, out.ptr))
        return out
    host = np.empty((n, n, n), dtype=np.float32, order="F")
    check(ctx.lib.fc3_mcc3_blocked(ctx.h, psi.ptr, phi.ptr, n, int(block_size), p,
                                   host.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
    return host


def unique_count(n: int) -> int:
    """Entries K3[a, b, l] with a <= b <= l: n (n + 1) (n + 2) / 6."""
    return n * (n + 1) * (n + 2) // 6


def unique_index(a, b, l):
    """Position of K3[a, b, l] (a <= b <= l, 0-based) in the packed unique vector:
    l (l + 1) (l + 2) / 6 + b (b + 1) / 2 + a."""
    a, b, l = (np.asarray(x, dtype=np.int64) for x in (a, b, l))
    return l * (l + 1) * (l + 2) // 6 + b * (b + 1) // 2 + a


def mcc3_unique(psi: DeviceArray, phi: DeviceArray, precision="fp32",
                out: Optional[DeviceArray] = None) -> DeviceArray:
    """The symmetric-unique part of mcc3(Psi, phi) (the unique-part TODO of mcc.jl:116-117):
    a device vector of n (n + 1) (n + 2) / 6 floats holding K3[a, b, l] for a <= b <= l at
    unique_index(a, b, l), each bitwise the entry mcc3 computes (fc3_mcc3_unique)."""
    ctx = psi.ctx
    n = phi.shape[0]
    _check_buf(psi, (n, n, n), np.float32, "mcc3_unique: Psi")
    _check_buf(phi, (n, n), np.float32, "mcc3_unique: phi", ctx)
    if out is not None:
        _check_buf(out, (unique_count(n),), np.float32, "mcc3_unique: out", ctx)
    out = out or DeviceArray((unique_count(n),), np.float32, ctx)
    check(ctx.lib.fc3_mcc3_unique(ctx.h, psi.ptr, phi.ptr, n, _prec(precision), out.ptr))
    return out


def k3_from_unique(values: np.ndarray, n: int) -> np.ndarray:
    """The full symmetric K3 (host, Fortran order) from the packed unique entries: every
    permutation of (a, b, l) takes the entry of its sorted indices ("rotate to enforce
    symmetry", mcc.jl:116-117)."""
    v = np.asarray(values).reshape(-1)
    if v.shape != (unique_count(n),):
        raise AssertionError(f"values must hold {unique_count(n)} entries")
    idx = np.indices((n, n, n), dtype=np.int64)
    s = np.sort(idx.reshape(3, -1), axis=0)
    full = v[unique_index(s[0], s[1], s[2])]
    return np.asfortranarray(full.reshape(n, n, n))


def mcc3_(psi: DeviceArray, phi: DeviceArray, precision="fp32") -> DeviceArray:
    """mcc3!(Psi, phi): K3 overwrites Psi; returns the same DeviceArray."""
    ctx = psi.ctx
    n = phi.shape[0]
    _check_buf(psi, (n, n, n), np.float32, "mcc3!: Psi")
    _check_buf(phi, (n, n), np.float32, "mcc3!: phi", ctx)
    check(ctx.lib.fc3_mcc3_inplace(ctx.h, psi.ptr, phi.ptr, n, _prec(precision)))
    return psi


# ---- output formats (src/io.jl:9-80) ---------------------------------------------------

def format_float(x: float) -> str:
    """x as the text files print it: Julia's print of a Float64."""
    L = _lib.load()
    buf = ctypes.create_string_buffer(40)
    check(L.fc3_format_float(float(x), buf, len(buf)))
    return buf.value.decode()


def _save(fc, N_atoms, D, outpath, filename, fmt, order, units, ctx):
    import os
    if fmt == "JLD2":
        raise NotImplementedError("fmt = :JLD2 needs an HDF5 writer, which this build does not "
                                  "have; use fmt = 'txt'")
    if fmt != "txt":
        raise ValueError(f"Invalid format, {fmt}")
    n = int(N_atoms) * int(D)
    if tuple(fc.shape) != (n,) * order:     # io.jl:10 / :40
        raise AssertionError("Incorrect dimensions")
    path = os.path.join(outpath, filename + ".txt")
    if isinstance(fc, DeviceArray):
        ctx = ctx or fc.ctx
        _check_buf(fc, (n,) * order, np.float64, "force constants", ctx)
        d, owned = fc, False
    else:
        ctx = ctx or Context.default()
        units = fc.units if units is None else units
        d, owned = DeviceArray.from_host(np.asfortranarray(fc.values, dtype=np.float64), ctx=ctx), True
    fn = ctx.lib.fc3_save_third_order_txt if order == 3 else ctx.lib.fc3_save_second_order_txt
    try:
        check(fn(ctx.h, d.ptr, int(N_atoms), int(D), (units or "").encode(), path.encode()))
    finally:
        if owned:
            d.free()
    return path


def save_second_order(Phi, N_atoms: int, D: int, outpath: str, filename: str = "second_order",
                      fmt: str = "JLD2", units: Optional[str] = None,
                      ctx: Optional[Context] = None) -> str:
    """save_second_order(Phi, N_atoms, D, outpath; filename, fmt) (src/io.jl:9-37).
    Phi: DenseForceConstants (host values) or a DeviceArray (FP64, n x n).  fmt "txt"
    writes `outpath/filename.txt`; the nonzeros are found on the GPU."""
    return _save(Phi, N_atoms, D, outpath, filename, fmt, 2, units, ctx)


def save_third_order(Psi, N_atoms: int, D: int, outpath: str, filename: str = "third_order",
                     fmt: str = "JLD2", units: Optional[str] = None,
                     ctx: Optional[Context] = None) -> str:
    """save_third_order(Psi, N_atoms, D, outpath; filename, fmt) (src/io.jl:39-69).
    Psi: DenseForceConstants (host values) or a DeviceArray (FP64, n x n x n)."""
    return _save(Psi, N_atoms, D, outpath, filename, fmt, 3, units, ctx)


# ---- adjoints (reverse mode; the targets of the Julia rrules) -------------------------

def mcc3_backward(psi: DeviceArray, phi: DeviceArray, dk3: DeviceArray, precision="fp32",
                  dpsi: bool = True, dphi: bool = True):
    """Pullback of K3 = mcc3(Psi, phi) for the cotangent dk3 = dL/dK3 (src/mcc.jl:15-17):
    returns (dL/dPsi or None, dL/dphi or None) as device arrays."""
    ctx = psi.ctx
    n = phi.shape[0]
    _check_buf(psi, (n, n, n), np.float32, "mcc3 backward: Psi")
    _check_buf(phi, (n, n), np.float32, "mcc3 backward: phi", ctx)
    _check_buf(dk3, (n, n, n), np.float32, "mcc3 backward: dK3", ctx)
    gp = DeviceArray((n, n, n), np.float32, ctx) if dpsi else None
    gf = DeviceArray((n, n), np.float32, ctx) if dphi else None
    check(ctx.lib.fc3_mcc3_backward(ctx.h, psi.ptr, phi.ptr, dk3.ptr, n, _prec(precision),
                                    gp.ptr if gp else None, gf.ptr if gf else None))
    return gp, gf


def weighted_square_loss(k3: DeviceArray, seed: int, grad: bool = True):
    """L = sum_e w_e K3_e^2 with w_e ~ U[0, 1) from Philox4x32-10 keyed by `seed` (the
    differentiation workload's scalar; the weights are regenerated, never stored).
    Returns (L, dL/dK3 as a device array, or None without grad)."""
    ctx = k3.ctx
    if not isinstance(k3, DeviceArray) or k3.dtype != np.float32:
        raise TypeError("k3 must be a float32 DeviceArray")
    if not 0 <= int(seed) < 2 ** 64:
        raise ValueError("seed must be in [0, 2^64)")
    g = DeviceArray(k3.shape, np.float32, ctx) if grad else None
    L = ctypes.c_double()
    count = int(np.prod(k3.shape, dtype=np.int64))
    check(ctx.lib.fc3_weighted_square_loss(ctx.h, k3.ptr, count, int(seed),
                                           g.ptr if g else None, ctypes.byref(L)))
    return float(L.value), g


def mcc3_rrule(psi: DeviceArray, phi: DeviceArray, precision="fp32"):
    """(K3, pullback) as a ChainRulesCore rrule returns them; pullback(dk3) ->
    (dL/dPsi, dL/dphi)."""
    k3 = mcc3(psi, phi, precision=precision)

    def pullback(dk3: DeviceArray):
        return mcc3_backward(psi, phi, dk3, precision=precision)

    return k3, pullback


def third_order_param_backward(sys: SuperCellSystem, pot, calc, dpsi: DeviceArray,
                               masses=None, ctx: Optional[Context] = None):
    """dL/d(potential parameters) for the cotangent dpsi of third_order's output.

    dpsi FP32: cotangent of the FP32 tensor third_order_device(..., psi32=...) writes
    (mass weighted when `masses` is given); dpsi FP64: of the unweighted FP64 tensor.
    Returns {"sigma", "epsilon"} for LJ and {"A", "B", "lambda", "gamma"} for SW."""
    ctx = ctx or dpsi.ctx
    pos, L = _sys_args(sys)
    n = sys.n_dof
    if not isinstance(dpsi, DeviceArray) or dpsi.dtype not in (np.float32, np.float64):
        raise TypeError("dpsi must be a float32 or float64 DeviceArray")
    _check_buf(dpsi, (n, n, n), dpsi.dtype, "dpsi", ctx)
    m = None if masses is None else _check_masses(masses, sys.n_atoms)
    if dpsi.dtype == np.float32:
        d32, d64 = dpsi.ptr, None
    elif dpsi.dtype == np.float64:
        if m is not None:
            raise AssertionError("an FP64 cotangent refers to the unweighted tensor: masses must be None")
        d32, d64 = None, dpsi.ptr
    else:
        raise TypeError("dpsi must be float32 or float64")
    if isinstance(pot, LJ):
        g = np.zeros(2, dtype=np.float64)
        check(ctx.lib.fc3_third_order_lj_param_backward(
            ctx.h, _ptr(pos), sys.n_atoms, _ptr(L), pot.sigma, pot.epsilon, pot.r_cut, calc.r_cut,
            calc.tol, _ptr(m) if m is not None else None, d32, d64, _ptr(g)))
        return {"sigma": float(g[0]), "epsilon": float(g[1])}
    if isinstance(pot, StillingerWeberSilicon):
        prm = np.ascontiguousarray(pot.params.as_abi_array(), dtype=np.float64)
        g = np.zeros(4, dtype=np.float64)
        check(ctx.lib.fc3_third_order_sw_param_backward(
            ctx.h, _ptr(pos), sys.n_atoms, _ptr(L), _ptr(prm), calc.r_cut, calc.tol,
            _ptr(m) if m is not None else None, d32, d64, _ptr(g)))
        return {"A": float(g[0]), "B": float(g[1]), "lambda": float(g[2]), "gamma": float(g[3])}
    raise NotImplementedError(f"parameter adjoint for {type(pot).__name__}")


# ---- batched snapshots ------------------------------------------------------------------

def batch_third_order_mcc3(snapshots, pot, calc, precision="bf16x3", autotune: bool = True,
                           graph: bool = True, device_modes: bool = False,
                           ctx: Optional[Context] = None):
    """For each (sys, phi, k3) in `snapshots`: Psi = third_order(sys, pot, calc) mass
    weighted to FP32 with sys.masses, then k3 <- mcc3(Psi, phi), all in one graph launch.
    phi, k3: device FP32 arrays (n x n, n x n x n).  Returns once the work is enqueued.

    device_modes: phi is an output: get_modes(dynamical_matrix(sys, pot, calc)) of each
    snapshot is computed on the device first (rocSOLVER, batched over same-size
    snapshots); returns the list of eigenvalue arrays (3 rigid modes zeroed)."""
    if not isinstance(pot, LJ):
        raise NotImplementedError("batched assembly implemented for LJ")
    snaps = list(snapshots)
    if not snaps:
        raise AssertionError("empty batch")
    ctx = ctx or snaps[0][1].ctx
    arr = (_lib.Snapshot * len(snaps))()
    keep, freqs = [], []
    for i, (sys, phi, k3) in enumerate(snaps):
        n = sys.n_dof
        _check_buf(phi, (n, n), np.float32, f"snapshot {i}: phi", ctx)
        _check_buf(k3, (n, n, n), np.float32, f"snapshot {i}: k3", ctx)
        pos, L = _sys_args(sys)
        m = _check_masses(sys.masses, sys.n_atoms)
        keep += [pos, m]
        arr[i].positions = _ptr(pos)
        arr[i].n_atoms = sys.n_atoms
        for c in range(3):
            arr[i].box[c] = float(L[c])
        arr[i].masses = _ptr(m)
        arr[i].phi = phi.ptr
        arr[i].k3 = k3.ptr
        if device_modes:
            w = np.zeros(n)
            freqs.append(w)
            arr[i].freqs_sq = _ptr(w)
    flags = ((_lib.FC3_BATCH_AUTOTUNE if autotune else 0) |
             (0 if graph else _lib.FC3_BATCH_NO_GRAPH) |
             (_lib.FC3_BATCH_DEVICE_MODES if device_modes else 0))
    check(ctx.lib.fc3_batch_third_order_mcc3_lj(ctx.h, arr, len(snaps), pot.sigma, pot.epsilon,
                                                pot.r_cut, calc.r_cut, calc.tol, _prec(precision),
                                                flags))
    del keep
    return freqs if device_modes else None


def batch_graph_nodes(ctx: Optional[Context] = None) -> int:
    ctx = ctx or Context.default()
    v = ctypes.c_size_t()
    check(ctx.lib.fc3_batch_graph_nodes(ctx.h, ctypes.byref(v)))
    return int(v.value)


def batch_graph_lanes(ctx: Optional[Context] = None) -> int:
    """Parallel branches of the last captured batch graph."""
    ctx = ctx or Context.default()
    v = ctypes.c_int()
    check(ctx.lib.fc3_batch_graph_lanes(ctx.h, ctypes.byref(v)))
    return int(v.value)


def autotune_query(n: int, precision="bf16x3", ctx: Optional[Context] = None):
    """(variant, tiles, ms) chosen for size n: variant 2 = FP32-operand kernels, 3 =
    pre-split BF16x3 kernels."""
    ctx = ctx or Context.default()
    v, t, ms = ctypes.c_int(), ctypes.c_long(), ctypes.c_float()
    check(ctx.lib.fc3_autotune_query(ctx.h, n, _prec(precision), ctypes.byref(v), ctypes.byref(t),
                                     ctypes.byref(ms)))
    return int(v.value), int(t.value), float(ms.value)


def autotune_save(path: str, ctx: Optional[Context] = None):
    ctx = ctx or Context.default()
    check(ctx.lib.fc3_autotune_save(ctx.h, path.encode()))


def autotune_load(path: str, ctx: Optional[Context] = None):
    ctx = ctx or Context.default()
    check(ctx.lib.fc3_autotune_load(ctx.h, path.encode()))
