"""save_second_order / save_third_order with fmt = :txt (src/io.jl:9-62) on the GPU.

The nonzeros are found and ordered on the device, so the file's data lines must equal,
byte for byte, the oracle's restatement of the reference's loops (oracle.save_txt_lines)
for the same tensor, and read back to the exact values.  The header carries the atom
count, round(sum, sigdigits = 7) and the units; the sum's last digits depend on the
summation order (the reference's is Julia's pairwise sum), so it is compared to the
oracle's sum within that rounding.  The JLD2 format is not built (no HDF5 writer).
"""
import numpy as np
import pytest

import fc3
from fc3 import systems as S
from oracle import fc3_oracle as O

pytestmark = pytest.mark.gpu

CALC = S.AnalyticalCalculator(1e-12, S.ARGON_LJ.r_cut)


def _check_file(path, values, N, D, symbol, units):
    text = open(path, encoding="utf-8").read().split("\n")
    assert text[-1] == ""
    header, lines = text[0], text[1:-1]
    want = f"#{N} atoms, Sum {symbol} = "
    assert header.startswith(want) and header.endswith(f", Units: {units}")
    s = float(header[len(want):header.index(", Units:")])
    ref_sum = float(np.sum(values))
    assert abs(s - ref_sum) <= 1e-6 * max(abs(ref_sum), 1e-300) + 1e-9 * float(np.abs(values).sum())
    assert lines == O.save_txt_lines(values, N, D)
    # read back: positions and exact values
    back = np.zeros_like(values)
    for ln in lines:
        f = ln.split()
        ids = [int(t) - 1 for t in f[:-1]]
        atoms, comps = ids[0::2], ids[1::2]
        back[tuple(D * a + c for a, c in zip(atoms, comps))] = float(f[-1])
    assert np.array_equal(back, values)
    return len(lines)


@pytest.mark.parametrize("cells", [(2, 2, 2), (3, 3, 3)])
def test_save_third_order_txt_lj(ctx, tmp_path, cells):
    sysd = S.fcc(S.ARGON_A, cells, S.ARGON_MASS, 0.05, seed=17)
    n, N = sysd.n_dof, sysd.n_atoms
    d64 = fc3.DeviceArray((n, n, n), np.float64, ctx)
    fc3.third_order_device(sysd, S.ARGON_LJ, CALC, psi64=d64, ctx=ctx)
    host = d64.to_host()
    assert np.array_equal(host, O.lj_third_order(sysd, S.ARGON_LJ, CALC))
    path = fc3.save_third_order(d64, N, 3, str(tmp_path), fmt="txt", units="kcal Å^-3 mol^-1")
    assert path.endswith("third_order.txt")
    print(f"save_third_order txt FCC {cells}: written, checking against the oracle", flush=True)
    nl = _check_file(path, host, N, 3, "Ψ", "kcal Å^-3 mol^-1")
    print(f"save_third_order txt FCC {cells}: {nl} nonzeros", flush=True)
    d64.free()


def test_save_from_host_force_constants(ctx, tmp_path):
    sysd = S.fcc(S.ARGON_A, (2, 2, 2), S.ARGON_MASS, 0.03, seed=5)
    ifc3 = fc3.third_order(sysd, S.ARGON_LJ, CALC, ctx=ctx)
    p3 = fc3.save_third_order(ifc3, sysd.n_atoms, 3, str(tmp_path), filename="psi", fmt="txt",
                              ctx=ctx)
    _check_file(p3, ifc3.values, sysd.n_atoms, 3, "Ψ", ifc3.units)
    ifc2 = fc3.second_order(sysd, S.ARGON_LJ, CALC, ctx=ctx)
    p2 = fc3.save_second_order(ifc2, sysd.n_atoms, 3, str(tmp_path), fmt="txt", ctx=ctx)
    assert p2.endswith("second_order.txt")
    _check_file(p2, ifc2.values, sysd.n_atoms, 3, "Φ", ifc2.units)


@pytest.mark.parametrize("D", [1, 2])
@pytest.mark.parametrize("order", [2, 3])
def test_save_txt_other_dimensions(ctx, tmp_path, D, order):
    N = 7
    rng = np.random.default_rng(100 + 10 * D + order)
    v = rng.standard_normal((N * D,) * order) * 10.0 ** rng.integers(-8, 8, (N * D,) * order)
    v[rng.random(v.shape) < 0.6] = 0.0
    v = np.asfortranarray(v)
    d = fc3.DeviceArray.from_host(v, ctx=ctx)
    save = fc3.save_third_order if order == 3 else fc3.save_second_order
    path = save(d, N, D, str(tmp_path), fmt="txt", units="eV")
    _check_file(path, v, N, D, "Ψ" if order == 3 else "Φ", "eV")


def test_save_format_and_dimension_errors(ctx, tmp_path):
    n = 6
    d = fc3.DeviceArray((n, n, n), np.float64, ctx, zero=True)
    with pytest.raises(NotImplementedError):
        fc3.save_third_order(d, 2, 3, str(tmp_path))                  # default fmt = JLD2
    with pytest.raises(ValueError, match="Invalid format"):
        fc3.save_third_order(d, 2, 3, str(tmp_path), fmt="csv")
    with pytest.raises(AssertionError, match="Incorrect dimensions"):
        fc3.save_third_order(d, 3, 3, str(tmp_path), fmt="txt")
    # an all-zero tensor: header only
    path = fc3.save_third_order(d, 2, 3, str(tmp_path), fmt="txt", units="u")
    assert open(path, encoding="utf-8").read() == "#2 atoms, Sum Ψ = 0.0, Units: u\n"
