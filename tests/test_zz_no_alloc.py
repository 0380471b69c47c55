"""fc3_third_order_lj_noalloc, the allocation-free variant of fc3_third_order_lj, called
through the C ABI: its FP64 and mass-weighted FP32 outputs equal the allocating entry
point's bit for bit on two calls in a row (the second with a smaller system), the calls
grow device memory by nothing, and a system whose scratch exceeds the static array is
refused with FC3_ERR_CAPACITY (its outputs are allocated at full size all the same)."""
import ctypes

import numpy as np
import pytest

import fc3
from fc3 import _lib
from fc3 import systems as S

pytestmark = pytest.mark.gpu

LJ = S.ARGON_LJ
CALC = S.AnalyticalCalculator(1e-12, LJ.r_cut)
_dp = ctypes.POINTER(ctypes.c_double)


def _noalloc(ctx, sysd, calc, d64, d32, masses=True):
    pos = np.ascontiguousarray(sysd.positions, dtype=np.float64)
    box = np.ascontiguousarray(sysd.box_sizes, dtype=np.float64)
    m = np.ascontiguousarray(sysd.masses, dtype=np.float64)
    return ctx.lib.fc3_third_order_lj_noalloc(
        ctx.h, pos.ctypes.data_as(_dp), sysd.n_atoms, box.ctypes.data_as(_dp), LJ.sigma,
        LJ.epsilon, LJ.r_cut, calc.r_cut, calc.tol, m.ctypes.data_as(_dp) if masses else None,
        d64.ptr if d64 is not None else None, d32.ptr if d32 is not None else None)


def test_noalloc_matches_allocating_entry_point():
    ctx = fc3.Context(0)
    try:
        for cells, seed in (((3, 3, 3), 5), ((2, 2, 2), 6)):
            sysd = S.fcc(S.ARGON_A, cells, S.ARGON_MASS, 0.05, seed=seed)
            n = sysd.n_dof
            print(f"noalloc: FCC {cells} (n = {n})", flush=True)
            ref64 = fc3.DeviceArray((n, n, n), np.float64, ctx)
            ref32 = fc3.DeviceArray((n, n, n), np.float32, ctx)
            fc3.third_order_device(sysd, LJ, CALC, psi64=ref64, psi32=ref32, masses=sysd.masses,
                                   ctx=ctx)
            out64 = fc3.DeviceArray((n, n, n), np.float64, ctx)
            out32 = fc3.DeviceArray((n, n, n), np.float32, ctx)
            ctx.synchronize()
            free0 = ctx.memory()[0]
            _lib.check(_noalloc(ctx, sysd, CALC, out64, out32))
            ctx.synchronize()
            assert ctx.memory()[0] == free0
            assert np.array_equal(out64.to_host(), ref64.to_host())
            assert np.array_equal(out32.to_host(), ref32.to_host())
            # unweighted FP32 only, on the same scratch
            fc3.third_order_device(sysd, LJ, CALC, psi32=ref32, ctx=ctx)
            _lib.check(_noalloc(ctx, sysd, CALC, None, out32, masses=False))
            ctx.synchronize()
            assert np.array_equal(out32.to_host(), ref32.to_host())
            for d in (ref64, ref32, out64, out32):
                d.free()
    finally:
        ctx.close()


def test_noalloc_refuses_systems_beyond_its_scratch():
    """FCC 5x5x5 (500 atoms) with a neighbour cutoff that takes in every other atom needs
    500 * 499 * 296 bytes > 64 MiB of lists and blocks."""
    ctx = fc3.Context(0)
    try:
        sysd = S.fcc(S.ARGON_A, (5, 5, 5), S.ARGON_MASS, 0.05, seed=7)
        calc = S.AnalyticalCalculator(1e-12, 26.0)
        n = sysd.n_dof
        d64 = fc3.DeviceArray((n, n, n), np.float64, ctx)
        d32 = fc3.DeviceArray((n, n, n), np.float32, ctx)
        st = _noalloc(ctx, sysd, calc, d64, d32)
        assert st == _lib.FC3_ERR_CAPACITY
        assert b"static scratch" in ctx.lib.fc3_last_error()
        ctx.synchronize()
    finally:
        ctx.close()
