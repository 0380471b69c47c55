"""Batched snapshots (IFC3 assembly + mcc3 in one graph launch) on the GPU.

Each snapshot's K3 is checked against the FP64 host contraction of the oracle's
mass-weighted Psi for the same positions (relative Frobenius error <= 1e-5 for BF16x3,
2e-6 for FP32).  The graph is checked against the same kernels issued directly
(bitwise equal) and against its own replay (bitwise equal); the tuning table is
checked to survive a save/load round trip.
"""
import numpy as np
import pytest

import fc3
from fc3 import systems as S
from oracle import fc3_oracle as O

pytestmark = pytest.mark.gpu

LJ = S.ARGON_LJ
CALC = S.AnalyticalCalculator(1e-12, LJ.r_cut)
TOL = {"fp32": 2e-6, "bf16x3": 1e-5}


def _snapshots(ctx, spec):
    """spec: list of (cells, seed); sigma_u ~ U[0.02, 0.10] from the seed."""
    out = []
    for cells, seed in spec:
        su = float(np.random.default_rng(seed).uniform(0.02, 0.10))
        sysd = S.fcc(S.ARGON_A, cells, S.ARGON_MASS, su, seed=seed)
        f2 = O.lj_second_order(sysd, LJ, CALC)
        _, phi = O.get_modes(O.dynamical_matrix(f2, sysd.masses))
        phi32 = np.asfortranarray(phi.astype(np.float32))
        n = sysd.n_dof
        out.append((sysd, fc3.DeviceArray.from_host(phi32, ctx=ctx),
                    fc3.DeviceArray((n, n, n), np.float32, ctx), phi32))
    return out


def _ref(sysd, phi32):
    psi = O.lj_third_order(sysd, LJ, CALC)
    O.mass_weight_third_order(psi, sysd.masses)
    return O.mcc3_fp64(psi.astype(np.float32), phi32)


def _rel(a, b):
    return float(np.linalg.norm((a.astype(np.float64) - b).ravel()) / np.linalg.norm(b.ravel()))


@pytest.mark.parametrize("prec", ["bf16x3", "fp32"])
def test_batch_uneven_sizes_match_fp64(ctx, prec):
    spec = [((2, 2, 2), 1), ((3, 3, 3), 2), ((2, 2, 2), 3), ((3, 3, 3), 5)]
    snaps = _snapshots(ctx, spec)
    fc3.batch_third_order_mcc3([s[:3] for s in snaps], LJ, CALC, precision=prec, ctx=ctx)
    ctx.synchronize()
    print(f"{prec}: graph nodes {fc3.batch_graph_nodes(ctx)}", flush=True)
    for (sysd, _, k3, phi32), (cells, seed) in zip(snaps, spec):
        err = _rel(k3.to_host(), _ref(sysd, phi32))
        print(f"{prec} snapshot {cells} seed {seed}: rel fro err {err:.3e}", flush=True)
        assert err <= TOL[prec]


def test_batch_graph_replay_and_direct_launches_agree(ctx):
    spec = [((2, 2, 2), 11), ((3, 3, 3), 12), ((4, 4, 4), 13)]
    snaps = _snapshots(ctx, spec)
    args = [s[:3] for s in snaps]
    fc3.batch_third_order_mcc3(args, LJ, CALC, ctx=ctx)            # tunes, captures, launches
    ctx.synchronize()
    first = [s[2].to_host() for s in snaps]
    nodes = fc3.batch_graph_nodes(ctx)
    for s in snaps:
        s[2].to_host()
    fc3.batch_third_order_mcc3(args, LJ, CALC, ctx=ctx)            # same plan: replay
    ctx.synchronize()
    assert fc3.batch_graph_nodes(ctx) == nodes
    for a, s in zip(first, snaps):
        assert np.array_equal(a, s[2].to_host())
    fc3.batch_third_order_mcc3(args, LJ, CALC, graph=False, ctx=ctx)   # direct launches
    ctx.synchronize()
    for a, s in zip(first, snaps):
        assert np.array_equal(a, s[2].to_host())


def test_autotune_table_roundtrip(ctx, tmp_path):
    snaps = _snapshots(ctx, [((4, 4, 4), 21)])
    fc3.batch_third_order_mcc3([s[:3] for s in snaps], LJ, CALC, ctx=ctx)
    ctx.synchronize()
    v, t, ms = fc3.autotune_query(768, "bf16x3", ctx=ctx)
    print(f"n=768 bf16x3: variant {v} tiles {t} {ms:.3f} ms", flush=True)
    assert v in (2, 3, 4) and ms > 0
    path = str(tmp_path / "tune.txt")
    fc3.autotune_save(path, ctx=ctx)
    other = fc3.Context(ctx.device)
    try:
        fc3.autotune_load(path, ctx=other)
        assert fc3.autotune_query(768, "bf16x3", ctx=other)[:2] == (v, t)
    finally:
        other.close()


def _spread_slabs(w2, n, count):
    """`count` output slabs spread over the spectrum, the three rigid modes left out."""
    rigid = set(np.argsort(np.abs(w2), kind="stable")[:3].tolist())
    cand = np.unique(np.linspace(0, n - 1, count + 6).round().astype(int))
    return [int(l) for l in cand if l not in rigid][:count]


@pytest.mark.parametrize("variant", [4, 3])
def test_batch_forced_variant_timed_sizes_match_fp64(ctx, tmp_path, variant):
    """The sizes of the timed batch (FCC 3x3x3, 4x4x4, 5x5x5: 324, 768, 1500 DoF) with the
    configuration forced through a loaded tuning table instead of the tuner's choice:
    variant 4 (the fused path, no dense Psi) and variant 3 (dense Psi, pre-split BF16x3
    kernels).  Each snapshot's K3 is compared with the FP64 contraction of the oracle's
    mass-weighted FP32 Psi on 12 slabs spread over the spectrum (rigid modes left out)."""
    spec = [((3, 3, 3), 31), ((4, 4, 4), 32), ((5, 5, 5), 33)]
    table = tmp_path / "forced.txt"
    table.write_text("".join(f"{3 * 4 * c[0] ** 3} 1 {variant} 16384 1.0\n" for c, _ in spec))
    other = fc3.Context(ctx.device)
    try:
        fc3.autotune_load(str(table), ctx=other)
        snaps, modes = [], []
        for cells, seed in spec:
            su = float(np.random.default_rng(seed).uniform(0.02, 0.10))
            sysd = S.fcc(S.ARGON_A, cells, S.ARGON_MASS, su, seed=seed)
            f2 = O.lj_second_order(sysd, LJ, CALC)
            w2, phi = O.get_modes(O.dynamical_matrix(f2, sysd.masses))
            phi32 = np.asfortranarray(phi.astype(np.float32))
            n = sysd.n_dof
            snaps.append((sysd, fc3.DeviceArray.from_host(phi32, ctx=other),
                          fc3.DeviceArray((n, n, n), np.float32, other)))
            modes.append((w2, phi32))
        fc3.batch_third_order_mcc3(snaps, LJ, CALC, precision="bf16x3", autotune=False, ctx=other)
        other.synchronize()
        for (sysd, dphi, k3), (w2, phi32), (cells, seed) in zip(snaps, modes, spec):
            n = sysd.n_dof
            assert fc3.autotune_query(n, "bf16x3", ctx=other)[0] == variant
            coo = O.lj_third_order_coo(sysd, LJ, CALC)
            v32 = O.coo_mass_weight(coo, sysd.masses).astype(np.float32).astype(np.float64)
            coo32 = (coo[0], coo[1], coo[2], v32)
            got = k3.to_host()
            num = den = 0.0
            for l in _spread_slabs(w2, n, 12):
                ref = O.mcc3_fp64_slab(coo32, phi32.astype(np.float64), l)
                num += float(np.sum((got[:, :, l].astype(np.float64) - ref) ** 2))
                den += float(np.sum(ref * ref))
            err = (num / den) ** 0.5
            print(f"forced variant {variant}: FCC {cells} (n={n}) rel fro err {err:.3e}", flush=True)
            assert err <= TOL["bf16x3"]
            dphi.free()
            k3.free()
    finally:
        other.close()


def test_batch_device_modes(ctx):
    """FC3_BATCH_DEVICE_MODES: every snapshot's phi from the device eigensolver (batched
    over same-size snapshots) inside the batch call, after a single-snapshot fc3_modes_lj
    call in the same process.  Eigenvalues against the host FP64 get_modes (src/dynmat.jl:
    123-131) of the oracle's IFC2; phi orthonormal and an eigenbasis of D (FP32 rounding);
    K3 against the FP64 contraction of the oracle's Psi with that phi."""
    spec = [((3, 3, 3), 11), ((2, 2, 2), 12), ((3, 3, 3), 13), ((4, 4, 4), 14), ((3, 3, 3), 15)]
    first = S.fcc(S.ARGON_A, (3, 3, 3), S.ARGON_MASS, 0.05, seed=10)
    w_single, phi_single = fc3.modes_device(first, LJ, CALC, ctx=ctx)
    print("modes_device (single snapshot) done; batch with device modes next", flush=True)
    snaps = []
    for cells, seed in spec:
        su = float(np.random.default_rng(seed).uniform(0.02, 0.10))
        sysd = S.fcc(S.ARGON_A, cells, S.ARGON_MASS, su, seed=seed)
        n = sysd.n_dof
        snaps.append((sysd, fc3.DeviceArray((n, n), np.float32, ctx, zero=True),
                      fc3.DeviceArray((n, n, n), np.float32, ctx)))
    freqs = fc3.batch_third_order_mcc3(snaps, LJ, CALC, precision="bf16x3",
                                       device_modes=True, ctx=ctx)
    ctx.synchronize()
    for (sysd, dphi, k3), w, (cells, seed) in zip(snaps, freqs, spec):
        n = sysd.n_dof
        dyn = O.dynamical_matrix(O.lj_second_order(sysd, LJ, CALC), sysd.masses)
        w_ref, _ = O.get_modes(dyn)
        scale = np.abs(w_ref).max()
        assert np.abs(w - w_ref).max() <= 1e-11 * scale
        assert set(np.flatnonzero(w == 0.0)) == set(np.flatnonzero(w_ref == 0.0))
        phi = dphi.to_host().astype(np.float64)
        assert np.abs(phi.T @ phi - np.eye(n)).max() <= 2e-6
        lam = np.linalg.eigvalsh(dyn)                 # unzeroed eigenvalues, ascending
        res = np.linalg.norm(dyn @ phi - phi * lam) / np.linalg.norm(dyn)
        assert res <= 2e-6
        err = _rel(k3.to_host(), _ref(sysd, phi.astype(np.float32)))
        print(f"device modes {cells} seed {seed}: eig ok, K3 rel fro err {err:.3e}", flush=True)
        assert err <= TOL["bf16x3"]
        dphi.free()
        k3.free()
    phi_single.free()
    assert w_single.shape == (81 * 4,)


def test_batch_lanes_with_shared_outputs(ctx):
    """Lanes: snapshots whose K3 outputs are the same buffer form a chain that one lane
    runs in submission order, so each buffer must end with its last writer's K3, bit for
    bit the K3 of that snapshot computed alone.  Eight 96-DoF snapshots rotate through
    three buffers and four 324-DoF snapshots through two: five chains, five lanes (the
    graph's parallel branches); the direct launches give the same bits."""
    spec = [((2, 2, 2), 70 + j) for j in range(8)] + [((3, 3, 3), 80 + j) for j in range(4)]
    snaps = _snapshots(ctx, spec)
    bufs = {96: [fc3.DeviceArray((96,) * 3, np.float32, ctx) for _ in range(3)],
            324: [fc3.DeviceArray((324,) * 3, np.float32, ctx) for _ in range(2)]}
    args, last = [], {}
    for j, (sysd, dphi, own, _) in enumerate(snaps):
        n = sysd.n_dof
        b = bufs[n][(j if n == 96 else j - 8) % len(bufs[n])]
        args.append((sysd, dphi, b))
        last[id(b)] = j
    fc3.batch_third_order_mcc3(args, LJ, CALC, precision="bf16x3", ctx=ctx)
    ctx.synchronize()
    lanes = fc3.batch_graph_lanes(ctx)
    print(f"shared outputs: {lanes} lanes, {fc3.batch_graph_nodes(ctx)} graph nodes", flush=True)
    assert lanes == 5
    got = {id(b): b.to_host() for bl in bufs.values() for b in bl}
    for key, j in last.items():
        sysd, dphi, own, phi32 = snaps[j]
        fc3.batch_third_order_mcc3([(sysd, dphi, own)], LJ, CALC, precision="bf16x3", ctx=ctx)
        ctx.synchronize()
        alone = own.to_host()
        print(f"buffer of snapshot {j} (n={sysd.n_dof}): equal to it alone "
              f"{np.array_equal(got[key], alone)}", flush=True)
        assert np.array_equal(got[key], alone)
        assert _rel(alone, _ref(sysd, phi32)) <= TOL["bf16x3"]
    fc3.batch_third_order_mcc3(args, LJ, CALC, precision="bf16x3", graph=False, ctx=ctx)
    ctx.synchronize()
    for bl in bufs.values():
        for b in bl:
            assert np.array_equal(got[id(b)], b.to_host())
    for s in snaps:
        s[1].free()
        s[2].free()
    for bl in bufs.values():
        for b in bl:
            b.free()


def test_batch_lane_budget_shrinks_lanes(ctx, monkeypatch):
    """The lane planner under a memory cap (FC3_BATCH_LANE_BYTES): with room for every
    lane the batch of four distinct-output snapshots per size runs in 4 + 4 lanes; with
    a cap of 0 bytes the lanes shrink to one per size class and then to a single lane.
    The K3 of every plan is the same bit for bit (each snapshot's kernels do not depend
    on the lane it runs in)."""
    spec = [((2, 2, 2), 90 + j) for j in range(4)] + [((3, 3, 3), 95 + j) for j in range(4)]
    snaps = _snapshots(ctx, spec)
    args = [s[:3] for s in snaps]
    results = {}
    for cap in (None, "0"):
        if cap is None:
            monkeypatch.delenv("FC3_BATCH_LANE_BYTES", raising=False)
        else:
            monkeypatch.setenv("FC3_BATCH_LANE_BYTES", cap)
        fc3.batch_third_order_mcc3(args, LJ, CALC, precision="bf16x3", ctx=ctx)
        ctx.synchronize()
        lanes = fc3.batch_graph_lanes(ctx)
        print(f"lane cap {cap}: {lanes} lanes", flush=True)
        results[cap] = (lanes, [s[2].to_host() for s in snaps])
    monkeypatch.delenv("FC3_BATCH_LANE_BYTES", raising=False)
    assert results[None][0] == 8 and results["0"][0] == 1
    for a, b in zip(results[None][1], results["0"][1]):
        assert np.array_equal(a, b)
    for (sysd, _, k3, phi32), got in zip(snaps, results[None][1]):
        assert _rel(got, _ref(sysd, phi32)) <= TOL["bf16x3"]
    for s in snaps:
        s[1].free()
        s[2].free()
