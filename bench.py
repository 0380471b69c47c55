"""Benchmark: K3 TFLOP/s at 2592 DoF (BF16x3 and FP32) and IFC3 assembly GB/s.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cells 6] [--no-cpu-baseline]

Workload: 12-6 LJ argon (sigma 3.4 A, epsilon 0.24037 kcal/mol, r_cut 8.5 A, 39.95 amu)
on FCC 6x6x6 (864 atoms, 2592 DoF) with seeded Gaussian displacements sigma_u = 0.05 A.
One step of each phase is one pass of the hot path over the whole system:
  assembly:  fused third_order (FP64 Psi, 139 GB) + mass-weighted FP32 Psi (70 GB)
  mcc3:      K3 = Psi x1 phi x2 phi x3 phi (6 n^4 flop) from the FP32 Psi, phi from the
             host FP64 eigendecomposition of the dynamical matrix
Inputs are resident in HBM when the timed region starts.  The JSON line's `value`
is the BF16x3 K3 throughput, 6 n^4 / (time per mcc3).  Errors are relative
Frobenius errors of sampled K3 slabs against an FP64 host contraction of the same
FP32 Psi and phi.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "forceconstants.jl_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "K3 TFLOP/s at 2592 DoF, bf16x3 vs fp32 @≤1e-5 rel err; IFC3 assembly GB/s"
PEAK = {"bf16": 2.5e15, "fp32": 157.3e12, "hbm": 8.0e12}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def profile(ctx, kind):
    import ctypes
    t = ctypes.c_double()
    c = ctypes.c_int()
    from fc3._lib import check
    check(ctx.lib.fc3_profile_query(ctx.h, kind, ctypes.byref(t), ctypes.byref(c)))
    return float(t.value), int(c.value)


def profile_enable(ctx, on=True):
    from fc3._lib import check
    check(ctx.lib.fc3_profile_enable(ctx.h, 1 if on else 0))


def spread_slabs(w2, n, count):
    """`count` K3 output slabs spread evenly over the spectrum, leaving out the three
    rigid-translation modes (the smallest |omega^2|), whose slabs vanish by the acoustic
    sum rule and so carry no information about the contraction error."""
    rigid = set(np.argsort(np.abs(w2), kind="stable")[:3].tolist())
    cand = np.unique(np.linspace(0, n - 1, count + 6).round().astype(int))
    return [int(l) for l in cand if l not in rigid][:count]


def gemm_traffic():
    """HBM bytes per mode-2 k_gemm3 launch at this workload, from the committed rocprofv3
    PMC summary (FETCH_SIZE and WRITE_SIZE per dispatch, calibrated on a copy kernel), or
    None if that summary is absent."""
    path = os.path.join(ROOT, "profiles", "k_gemm3_lj_fcc666_traffic.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def cpu_model():
    """The host CPU's model name as /proc/cpuinfo gives it (None if unavailable)."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.lower().startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def blas_threads():
    try:
        from threadpoolctl import threadpool_info
        return max((i.get("num_threads", 1) for i in threadpool_info()), default=1)
    except Exception:
        return os.cpu_count() or 1


def cpu_baseline_mcc3(coo32, phi32, n, width=24, reps=5):
    """The reference has no CPU mcc3, so this restates its three mode products on the host
    with numpy/OpenBLAS FP32 GEMMs on all cores, on this workload's own mass-weighted FP32
    Psi and phi, as the device orders them (mode 3, then modes 1 and 2 per slab).  Bounded
    sample, the same fraction width / n of each product: mode 3 for the rows (i, j) with j
    in one band of `width` columns (dense, all n output modes: 2 width n^3 flop) and modes
    1 and 2 for `width` K3 slabs (4 width n^3 flop), from the dense slabs U_l = Psi x3
    phi[:, l] prepared beforehand.  Its flop rate is that of the whole 6 n^4 contraction.
    One warm-up, then the median of `reps` runs."""
    import scipy.sparse as sp
    i, j, k, v = coo32
    i = i.astype(np.int64)
    j = j.astype(np.int64)
    k = k.astype(np.int64)
    phi = np.asfortranarray(phi32, dtype=np.float32)
    phiT = np.ascontiguousarray(phi.T)
    c0 = n // 2
    sel = (j >= c0) & (j < c0 + width)
    A = np.zeros((n * width, n), dtype=np.float32, order="F")       # Psi rows of the band
    A[i[sel] + n * (j[sel] - c0), k[sel]] = v[sel].astype(np.float32)
    M = sp.csr_matrix((v, (i + n * j, k)), shape=(n * n, n))
    Ub = np.asarray(M @ phi[:, c0:c0 + width].astype(np.float64), dtype=np.float32)
    Us = [np.ascontiguousarray(Ub[:, c]).reshape(n, n, order="F") for c in range(width)]
    del M, Ub
    times = []
    for r in range(reps + 1):
        t0 = time.perf_counter()
        _ = A @ phi                                   # mode 3 for the band's rows
        for Ul in Us:
            _ = phiT @ Ul @ phi                       # modes 1 and 2 of one slab
        dt = time.perf_counter() - t0
        if r:
            times.append(dt)
    dt = float(np.median(times))
    th = blas_threads()
    return {"value": 6.0 * width * n ** 3 / dt / 1e12, "unit": "TFLOP/s", "cores": int(th),
            "kind": "port", "cpu": cpu_model(),
            "sample": f"FP32 mcc3 (numpy/OpenBLAS, {th} threads) of this workload's Psi and phi "
                      f"at {n} DoF: mode 3 for the rows of {width} Psi columns and modes 1-2 "
                      f"for {width} K3 slabs (6 x {width} n^3 flop, {width}/{n} of the whole), "
                      f"median of {reps} runs after a warm-up: {dt:.2f} s"}


def cpu_baseline_assembly(cells=5, reps=5):
    """The oracle's restatement of the reference loop (dense FP64 third_order!, ASR!,
    apply_tols!) on LJ argon FCC cells^3, threaded over atoms on all OpenMP threads as
    the reference's @tasks loops are (third_order_analytical.jl:14-15, asr.jl:19-40);
    one warm-up, then the median of `reps` runs."""
    from fc3 import systems as S
    from oracle import fc3_oracle as O
    sysd = S.fcc(S.ARGON_A, (cells,) * 3, S.ARGON_MASS, 0.05, seed=768)
    calc = S.AnalyticalCalculator(1e-12, S.ARGON_LJ.r_cut)
    threads = int(O.lib().fco_num_threads())
    n = sysd.n_dof
    log(f"cpu assembly (oracle, FCC {cells}^3, {threads} threads)")
    times = []
    for r in range(reps + 1):
        t0 = time.perf_counter()
        psi = O.lj_third_order(sysd, S.ARGON_LJ, calc)
        dt = time.perf_counter() - t0
        del psi
        if r:
            times.append(dt)
    dt = float(np.median(times))
    return {"value": 8.0 * n ** 3 / dt / 1e9, "unit": "GB/s", "cores": threads, "kind": "port",
            "cpu": cpu_model(),
            "sample": f"LJ FCC {cells}x{cells}x{cells} ({n} DoF) third_order, dense FP64, "
                      f"OpenMP over atoms, median of {reps} runs after a warm-up: {dt:.2f} s"}


def unique_timing(ctx, d32, dphi, k3, n, slabs, a, full_ms):
    """Symmetric-unique K3 (mcc3_unique, BF16x3) on the headline Psi and phi: time per
    call, and a bitwise check of its entries against the full block-sparse BF16x3 K3 on
    the sampled slabs.  (The unique form runs the same block-sparse tiles; the dense form,
    FC3_MCC3_DENSE=1, groups mode 1's K sum differently when n % 16 == 0 and agrees with
    it only to rounding.)"""
    import ctypes
    import fc3
    from fc3._lib import check
    os.environ.pop("FC3_MCC3_DENSE", None)
    fc3.mcc3(d32, dphi, precision="bf16x3", out=k3)   # the full K3 to compare against
    du = fc3.DeviceArray((fc3.unique_count(n),), np.float32, ctx)
    for _ in range(a.warmup):
        fc3.mcc3_unique(d32, dphi, precision="bf16x3", out=du)
    ctx.synchronize()
    profile_enable(ctx)
    ctx.timer_start()
    for _ in range(a.steps):
        fc3.mcc3_unique(d32, dphi, precision="bf16x3", out=du)
    t = ctx.timer_stop() / a.steps
    m3_ms, _ = profile(ctx, 1)
    m1_ms, _ = profile(ctx, 2)
    m2_ms, _ = profile(ctx, 5)
    profile_enable(ctx, False)
    buf = np.empty((n, n), dtype=np.float32, order="F")
    mism = 0
    for l in slabs:
        check(ctx.lib.fc3_memcpy_d2h(ctx.h, buf.ctypes.data_as(ctypes.c_void_p),
                                     ctypes.c_void_p(k3.ptr.value + 4 * n * n * l), 4 * n * n))
        tri = np.ascontiguousarray(buf[: l + 1, : l + 1].T[np.tril_indices(l + 1)])
        got = np.empty_like(tri)
        lo = int(fc3.unique_index(0, 0, l))
        check(ctx.lib.fc3_memcpy_d2h(ctx.h, got.ctypes.data_as(ctypes.c_void_p),
                                     ctypes.c_void_p(du.ptr.value + 4 * lo), 4 * tri.size))
        mism += int(np.count_nonzero(got.view(np.uint32) != tri.view(np.uint32)))
    du.free()
    ctx.release_workspace()   # the n^3 intermediate of the unique form
    res = {"ms_per_step": t, "entries": fc3.unique_count(n),
           "K3_equiv_TFLOP_per_s": 6.0 * n ** 4 / (t * 1e-3) / 1e12,
           "speedup_vs_full_bf16x3": full_ms / t,
           "mode3_ms_per_step": m3_ms / a.steps, "mode1_ms_per_step": m1_ms / a.steps,
           "mode2_ms_per_step": m2_ms / a.steps,
           "bitwise_mismatches": mism, "checked_slabs": slabs,
           "note": "K3[a,b,l] for a <= b <= l (BF16x3): modes 1 and 2 run only the GEMM tiles "
                   "holding such entries; checked bit for bit against the full K3"}
    log(f"unique {res}")
    return res


def sw_assembly(ctx, steps=3):
    """IFC3 assembly for Stillinger-Weber silicon on diamond 3x3x3 (648 DoF, sigma_u =
    0.03 A, seed 648): FP64 Psi + mass-weighted FP32 Psi in one pass, 12 n^3 bytes."""
    import fc3
    from fc3 import systems as S
    pot = S.StillingerWeberSilicon()
    calc = S.AutoDiffCalculator(1e-12, pot.r_cut)
    sysd = S.diamond(S.SI_A, (3, 3, 3), S.SI_MASS, 0.03, seed=648)
    n = sysd.n_dof
    d64 = fc3.DeviceArray((n, n, n), np.float64, ctx)
    d32 = fc3.DeviceArray((n, n, n), np.float32, ctx)
    fc3.third_order_device(sysd, pot, calc, psi64=d64, psi32=d32, masses=sysd.masses, ctx=ctx)
    ctx.synchronize()
    profile_enable(ctx)
    ctx.timer_start()
    for _ in range(steps):
        fc3.third_order_device(sysd, pot, calc, psi64=d64, psi32=d32, masses=sysd.masses, ctx=ctx)
    t = ctx.timer_stop() / steps
    w_ms, w_cnt = profile(ctx, 0)
    profile_enable(ctx, False)
    d64.free()
    d32.free()
    b = 12.0 * n ** 3
    res = {"config": f"SW silicon diamond 3x3x3 ({n} DoF)", "ms_per_step": t,
           "GB_per_s": b / (t * 1e-3) / 1e9, "writer_kernel_ms": w_ms / max(w_cnt, 1),
           "writer_GB_per_s": b / (w_ms / max(w_cnt, 1) * 1e-3) / 1e9, "bytes_per_step": b}
    log(f"sw assembly {res}")
    return res


def sw_fused(ctx, steps=3, cells=4):
    """Stillinger-Weber third_order + mcc3 on diamond 4x4x4 (1536 DoF, sigma_u = 0.03 A,
    seed 1536; cells=3: 648 DoF, seed 648), phi from the library's dynamical matrix and
    host eigensolver: the fused
    call (mode 1 over the reduced block table's nonzeros, modes 2 and 3 BF16x3, no n^3
    Psi) against the dense route (FP32 mass-weighted Psi, then the BF16x3 mcc3), with the
    relative difference of the two K3 on 4 sampled slabs."""
    import ctypes
    import fc3
    from fc3 import systems as S
    from fc3._lib import check
    pot = S.StillingerWeberSilicon()
    calc = S.AutoDiffCalculator(1e-12, pot.r_cut)
    sysd = S.diamond(S.SI_A, (cells, cells, cells), S.SI_MASS, 0.03, seed=3 * 8 * cells ** 3)
    n = sysd.n_dof
    _, phi = fc3.get_modes(fc3.dynamical_matrix(sysd, pot, calc, ctx=ctx))
    dphi = fc3.DeviceArray.from_host(np.asfortranarray(phi.astype(np.float32)), ctx=ctx)
    k3f = fc3.DeviceArray((n, n, n), np.float32, ctx)
    k3d = fc3.DeviceArray((n, n, n), np.float32, ctx)
    d32 = fc3.DeviceArray((n, n, n), np.float32, ctx)

    def dense():
        fc3.third_order_device(sysd, pot, calc, psi32=d32, masses=sysd.masses, ctx=ctx)
        fc3.mcc3(d32, dphi, precision="bf16x3", out=k3d)

    def fused():
        fc3.third_order_mcc3(sysd, pot, calc, dphi, masses=sysd.masses, out=k3f, ctx=ctx)

    res = {"config": f"SW silicon diamond {cells}x{cells}x{cells} ({n} DoF)"}
    for name, fn in (("dense", dense), ("fused", fused)):
        fn()
        ctx.synchronize()
        ctx.timer_start()
        for _ in range(steps):
            fn()
        res[f"{name}_ms_per_step"] = ctx.timer_stop() / steps
    d32.free()
    num = den = 0.0
    a = np.empty((n, n), dtype=np.float32, order="F")
    b = np.empty_like(a)
    for l in np.linspace(3, n - 1, 4).astype(int):
        off = 4 * n * n * int(l)
        check(ctx.lib.fc3_memcpy_d2h(ctx.h, a.ctypes.data_as(ctypes.c_void_p),
                                     ctypes.c_void_p(k3f.ptr.value + off), 4 * n * n))
        check(ctx.lib.fc3_memcpy_d2h(ctx.h, b.ctypes.data_as(ctypes.c_void_p),
                                     ctypes.c_void_p(k3d.ptr.value + off), 4 * n * n))
        num += float(np.sum((a.astype(np.float64) - b) ** 2))
        den += float(np.sum(b.astype(np.float64) ** 2))
    k3f.free()
    k3d.free()
    dphi.free()
    ctx.release_workspace()
    res["speedup_vs_dense"] = res["dense_ms_per_step"] / res["fused_ms_per_step"]
    res["rel_fro_diff_vs_dense"] = float(np.sqrt(num / den))
    res["note"] = ("dense = FP32 mass-weighted Psi (the assembly without the FP64 copy) + "
                   "block-sparse BF16x3 mcc3; fused = the same K3 with no n^3 Psi")
    log(f"sw fused {res}")
    return res


def batched_snapshots(ctx, per_size=32, replays=3, modes_timing=False):
    """The batched configuration: 3 x `per_size` LJ argon snapshots of FCC 3x3x3,
    4x4x4 and 5x5x5 (324, 768, 1500 DoF), sigma_u ~ U[0.02, 0.10] A per snapshot
    (seeded), phi per snapshot from the host FP64 eigendecomposition of its dynamical
    matrix (with modes_timing, the device modes are timed beside it for one snapshot per
    size); assembly + BF16x3 mcc3 of all snapshots in one graph launch.  K3 outputs
    rotate through 8 / 4 / 2 buffers for 324 / 768 / 1500 DoF (36 GB; the whole set
    would need 494 GB): snapshots sharing a buffer form one lane of the graph, the lanes
    run concurrently.  The error is checked on sampled slabs of the last two snapshots
    of each size."""
    import fc3
    from fc3 import systems as S
    from oracle import fc3_oracle as O
    calc = S.AnalyticalCalculator(1e-12, S.ARGON_LJ.r_cut)
    rng = np.random.default_rng(96)
    out_bufs, snaps, meta = {}, [], []
    t0 = time.perf_counter()
    modes = {}
    phi_s = 0.0
    for cells in (3, 4, 5):
        if modes_timing:
            modes[str(3 * 4 * cells ** 3)] = modes_setup_timing(ctx, cells, calc)
        for j in range(per_size):
            seed = int(rng.integers(1 << 30))
            su = float(np.random.default_rng(seed).uniform(0.02, 0.10))
            sysd = S.fcc(S.ARGON_A, (cells,) * 3, S.ARGON_MASS, su, seed=seed)
            n = sysd.n_dof
            tp = time.perf_counter()
            w2, phi = fc3.get_modes(fc3.dynamical_matrix(sysd, S.ARGON_LJ, calc, ctx=ctx))
            dph = fc3.DeviceArray.from_host(np.asfortranarray(phi.astype(np.float32)), ctx=ctx)
            phi_s += time.perf_counter() - tp
            nbuf = {3: 8, 4: 4}.get(cells, 2)
            if n not in out_bufs:
                out_bufs[n] = [fc3.DeviceArray((n, n, n), np.float32, ctx) for _ in range(nbuf)]
            snaps.append((sysd, dph, out_bufs[n][j % nbuf]))
            meta.append((sysd, dph, seed, su, w2))
        log(f"batched: phi for {per_size} snapshots of FCC {cells}^3 ready "
            f"({time.perf_counter() - t0:.0f} s)")
    fc3.batch_third_order_mcc3(snaps, S.ARGON_LJ, calc, precision="bf16x3", ctx=ctx)  # tune + capture
    ctx.synchronize()
    nodes = fc3.batch_graph_nodes(ctx)
    lanes = fc3.batch_graph_lanes(ctx)
    ctx.timer_start()
    for _ in range(replays):
        fc3.batch_third_order_mcc3(snaps, S.ARGON_LJ, calc, precision="bf16x3", ctx=ctx)
    t = ctx.timer_stop() / replays
    nsnap = len(snaps)
    flop = sum(6.0 * (3 * m[0].n_atoms) ** 4 for m in meta)
    tune = {}
    for n in sorted(out_bufs):
        v, tiles, ms = fc3.autotune_query(n, "bf16x3", ctx=ctx)
        tune[str(n)] = {"variant": v, "tiles": tiles, "snapshot_ms": ms}
    # each size class as a batch of its own: the lanes' concurrency for small snapshots
    by_size = {}
    for n in sorted(out_bufs):
        sub = [sn for sn, m in zip(snaps, meta) if 3 * m[0].n_atoms == n]
        fc3.batch_third_order_mcc3(sub, S.ARGON_LJ, calc, precision="bf16x3", ctx=ctx)
        ctx.synchronize()
        ctx.timer_start()
        for _ in range(replays):
            fc3.batch_third_order_mcc3(sub, S.ARGON_LJ, calc, precision="bf16x3", ctx=ctx)
        ts = ctx.timer_stop() / replays
        by_size[str(n)] = {
            "snapshots": len(sub), "graph_lanes": fc3.batch_graph_lanes(ctx), "ms": ts,
            "serial_sum_ms": len(sub) * tune[str(n)]["snapshot_ms"],
            "K3_TFLOP_per_s": len(sub) * 6.0 * n ** 4 / (ts * 1e-3) / 1e12}
    log(f"batched per size: {by_size}")
    # error of the surviving outputs (last two per size) on sampled slabs
    import ctypes
    from fc3._lib import check
    errs = []
    for i in range(nsnap - 1, -1, -1):
        sysd, dph, seed, su, w2 = meta[i]
        n = sysd.n_dof
        if sum(1 for e in errs if e["n"] == n) >= 2:
            continue
        phi32 = dph.to_host()
        coo = O.lj_third_order_coo(sysd, S.ARGON_LJ, calc)
        v32 = O.coo_mass_weight(coo, sysd.masses).astype(np.float32).astype(np.float64)
        k3 = snaps[i][2]
        num = den = 0.0
        buf = np.empty((n, n), dtype=np.float32, order="F")
        for l in spread_slabs(w2, n, 8):
            ref = O.mcc3_fp64_slab((coo[0], coo[1], coo[2], v32), phi32.astype(np.float64), l)
            check(ctx.lib.fc3_memcpy_d2h(ctx.h, buf.ctypes.data_as(ctypes.c_void_p),
                                         ctypes.c_void_p(k3.ptr.value + 4 * n * n * l), 4 * n * n))
            d = buf.astype(np.float64) - ref
            num += float((d * d).sum())
            den += float((ref * ref).sum())
        errs.append({"n": n, "seed": seed, "sigma_u": su, "rel_fro_err": (num / den) ** 0.5})
    # F2: the same batch with every phi from the device eigensolver inside the call
    # (phi buffers are outputs now); wall time of the whole call, second of two
    dev = {}
    for rep in range(2):
        tw = time.perf_counter()
        freqs = fc3.batch_third_order_mcc3(snaps, S.ARGON_LJ, calc, precision="bf16x3",
                                           device_modes=True, ctx=ctx)
        ctx.synchronize()
        dev["ms_per_batch_with_phi"] = (time.perf_counter() - tw) * 1e3
    dev["phi_setup_s"] = max(dev["ms_per_batch_with_phi"] - t, 0.0) * 1e-3
    dev_errs = []
    for i in (nsnap - 1, nsnap - 1 - per_size, nsnap - 1 - 2 * per_size):
        sysd, dph, seed, su, w2 = meta[i]
        n = sysd.n_dof
        w_ref = w2
        phi32 = dph.to_host()
        coo = O.lj_third_order_coo(sysd, S.ARGON_LJ, calc)
        v32 = O.coo_mass_weight(coo, sysd.masses).astype(np.float32).astype(np.float64)
        k3 = snaps[i][2]
        num = den = 0.0
        buf = np.empty((n, n), dtype=np.float32, order="F")
        for l in spread_slabs(w_ref, n, 4):
            ref = O.mcc3_fp64_slab((coo[0], coo[1], coo[2], v32), phi32.astype(np.float64), l)
            check(ctx.lib.fc3_memcpy_d2h(ctx.h, buf.ctypes.data_as(ctypes.c_void_p),
                                         ctypes.c_void_p(k3.ptr.value + 4 * n * n * l), 4 * n * n))
            d = buf.astype(np.float64) - ref
            num += float((d * d).sum())
            den += float((ref * ref).sum())
        dev_errs.append({"n": n, "seed": seed,
                         "eig_max_rel_diff_vs_host": float(np.abs(freqs[i] - w_ref).max() /
                                                           np.abs(w_ref).max()),
                         "rel_fro_err": (num / den) ** 0.5})
    dev["errors_sampled"] = dev_errs
    dev["note"] = ("phi of all snapshots by IFC2 + mass weighting + rocSOLVER dsyevd "
                   "(strided-batched over same-size snapshots) inside the batch call; "
                   "K3 checked against the FP64 contraction with that phi")
    log(f"batched device modes: {dev}")
    for bufs in out_bufs.values():
        for b in bufs:
            b.free()
    for m in meta:
        m[1].free()
    res = {"snapshots": nsnap, "sizes_dof": sorted(out_bufs), "ms_per_batch": t,
           "snapshots_per_s": nsnap / (t * 1e-3), "K3_TFLOP_per_s": flop / (t * 1e-3) / 1e12,
           "graph_nodes": nodes, "graph_lanes": lanes, "replays_timed": replays,
           "per_size": by_size,
           "autotuned": tune,
           # one snapshot after another at the tuner's per-snapshot times: what the
           # batch would take without concurrent lanes
           "serial_sum_ms": sum(tune[str(3 * m[0].n_atoms)]["snapshot_ms"] for m in meta),
           "errors_sampled": errs, "modes_setup_per_snapshot": modes,
           "phi_setup_s": phi_s, "device_modes": dev,
           "note": "assembly (FP64 blocks, mass-weighted FP32 Psi) + BF16x3 mcc3 per snapshot, "
                   "one graph launch per batch"}
    log(f"batched: {res}")
    return res


def modes_setup_timing(ctx, cells, calc):
    """phi of one FCC cells^3 snapshot on the host (IFC2 on the GPU, D and eigh on the host,
    as the reference does) against dynamical matrix + rocSOLVER eigensolver on the GPU."""
    import fc3
    from fc3 import systems as S
    sysd = S.fcc(S.ARGON_A, (cells,) * 3, S.ARGON_MASS, 0.05, seed=cells)
    _, d = fc3.modes_device(sysd, S.ARGON_LJ, calc, ctx=ctx)      # warm-up (library load)
    d.free()
    th = time.perf_counter()
    fc3.get_modes(fc3.dynamical_matrix(sysd, S.ARGON_LJ, calc, ctx=ctx))
    th = time.perf_counter() - th
    td = time.perf_counter()
    _, d = fc3.modes_device(sysd, S.ARGON_LJ, calc, ctx=ctx)
    ctx.synchronize()
    td = time.perf_counter() - td
    d.free()
    return {"host_ms": th * 1e3, "device_ms": td * 1e3}


def adjoint_timing(ctx, cells=5, reps=2):
    """The differentiation configuration: LJ argon FCC 5x5x5 (1500 DoF), sigma_u = 0.05 A
    (seeded), L = sum w K3^2 with w ~ U(0, 1) from Philox4x32-10 (seed 1500), generated on
    the device and regenerable on any host.  Forward = assembly (mass-weighted FP32 Psi)
    + BF16x3 mcc3 + L; backward = the cotangent 2 w K3, the mcc3 pullback (dL/dPsi and
    dL/dphi) and the sigma/epsilon pullback of the assembly.  Gradients are checked in
    tests against FP64 central differences and the dual-number oracle at sizes where those
    are affordable; here dL/depsilon is checked against <dL/dPsi, Psi> / epsilon, exact
    because Psi is linear in epsilon."""
    import fc3
    from fc3 import systems as S
    calc = S.AnalyticalCalculator(1e-12, S.ARGON_LJ.r_cut)
    sysd = S.fcc(S.ARGON_A, (cells,) * 3, S.ARGON_MASS, 0.05, seed=1500)
    n = sysd.n_dof
    _, phi = fc3.get_modes(fc3.dynamical_matrix(sysd, S.ARGON_LJ, calc, ctx=ctx))
    dphi = fc3.DeviceArray.from_host(phi, dtype=np.float32, ctx=ctx)
    psi = fc3.DeviceArray((n, n, n), np.float32, ctx)
    k3 = fc3.DeviceArray((n, n, n), np.float32, ctx)

    seed = 1500

    def forward():
        fc3.third_order_device(sysd, S.ARGON_LJ, calc, psi32=psi, masses=sysd.masses, ctx=ctx)
        fc3.mcc3(psi, dphi, precision="bf16x3", out=k3)
        return fc3.weighted_square_loss(k3, seed, grad=False)[0]

    def backward():
        _, G = fc3.weighted_square_loss(k3, seed)
        dpsi, dph = fc3.mcc3_backward(psi, dphi, G, precision="bf16x3")
        G.free()
        g = fc3.third_order_param_backward(sysd, S.ARGON_LJ, calc, dpsi, masses=sysd.masses)
        return dpsi, dph, g

    loss = forward()
    dpsi, dph, g = backward()
    ctx.synchronize()
    # exact check: dL/depsilon = <dL/dPsi, Psi> / epsilon
    import ctypes
    from fc3._lib import check
    num = 0.0
    a = np.empty((n, n), dtype=np.float32, order="F")
    b = np.empty((n, n), dtype=np.float32, order="F")
    for l in range(n):
        off = 4 * n * n * l
        check(ctx.lib.fc3_memcpy_d2h(ctx.h, a.ctypes.data_as(ctypes.c_void_p),
                                     ctypes.c_void_p(dpsi.ptr.value + off), 4 * n * n))
        check(ctx.lib.fc3_memcpy_d2h(ctx.h, b.ctypes.data_as(ctypes.c_void_p),
                                     ctypes.c_void_p(psi.ptr.value + off), 4 * n * n))
        num += float(np.dot(a.ravel().astype(np.float64), b.ravel().astype(np.float64)))
    eps_check = num / S.ARGON_LJ.epsilon
    dpsi.free()
    dph.free()
    ctx.timer_start()
    for _ in range(reps):
        forward()
    t_fwd = ctx.timer_stop() / reps
    ctx.timer_start()
    for _ in range(reps):
        forward()
        dp, df, _ = backward()
        dp.free()
        df.free()
    t_fb = ctx.timer_stop() / reps
    for d in (psi, k3, dphi):
        d.free()
    res = {"config": f"LJ argon FCC {cells}x{cells}x{cells} ({n} DoF)",
           "loss": "L = sum w K3^2, w ~ U(0,1) Philox4x32-10 seed 1500 (device)", "L": loss,
           "forward_ms": t_fwd,
           "forward_backward_ms": t_fb, "ratio": t_fb / t_fwd,
           "dL_dsigma": g["sigma"], "dL_depsilon": g["epsilon"],
           "dL_depsilon_vs_exact_rel": abs(g["epsilon"] - eps_check) / abs(eps_check),
           "note": "backward = mcc3 pullback in BF16X3 (the seven mode products and the "
                   "three n x n^2 x n contractions, split K, on the BF16x3 kernels) + the "
                   "sigma/epsilon pullback"}
    log(f"adjoint: {res}")
    return res


def vendor_gemm(n, slabs=144, reps=5):
    """The vendor library's BF16 GEMM (torch.matmul, which calls hipBLASLt) on the shape of
    one mode-2 launch of the mcc3 above: (slabs n) x n times n x n, operands drawn from a
    standard normal (random mantissa bits, as the split planes have), BF16 output.  Also
    16384^3 for scale.  Context for k_gemm3's fraction of peak, not part of the path."""
    import torch
    out = {"what": "torch.matmul bf16 (hipBLASLt), randn operands, median of "
                   f"{reps} after 2 warm-ups"}

    def rate(M, N, K):
        a_ = torch.randn(M, K, device="cuda", dtype=torch.bfloat16)
        b_ = torch.randn(K, N, device="cuda", dtype=torch.bfloat16)
        for _ in range(2):
            c_ = a_ @ b_
        ts = []
        for _ in range(reps):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            c_ = a_ @ b_
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        del a_, b_, c_
        torch.cuda.empty_cache()
        ms = float(np.median(ts))
        log(f"vendor gemm {M}x{N}x{K}: {ms:.3f} ms")
        return {"shape_MxNxK": [M, N, K], "ms": ms,
                "TFLOP_per_s": 2.0 * M * N * K / (ms * 1e-3) / 1e12,
                "frac_of_2500": 2.0 * M * N * K / (ms * 1e-3) / 2.5e15}

    out["mode2_shape"] = rate(slabs * n, n, n)
    out["square_16384"] = rate(16384, 16384, 16384)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cells", type=int, default=6)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--fp32-steps", type=int, default=1)
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched-snapshot section (96 snapshots of 324/768/1500 DoF)")
    ap.add_argument("--modes-timing", action="store_true",
                    help="time host against device normal modes for one snapshot per size")
    ap.add_argument("--no-fused", action="store_true",
                    help="skip the fused third_order + mcc3 (no dense Psi) timing")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the forward+backward timing at FCC 5x5x5 (1500 DoF)")
    ap.add_argument("--no-unique", action="store_true",
                    help="skip the symmetric-unique K3 (mcc3_unique) section")
    ap.add_argument("--no-sw", action="store_true",
                    help="skip the Stillinger-Weber assembly timing (diamond 3x3x3, 648 DoF)")
    ap.add_argument("--vendor-gemm", action="store_true",
                    help="only time the vendor BF16 GEMM (torch.matmul: hipBLASLt) on the "
                         "mode-2 GEMM shape and print that JSON line")
    a = ap.parse_args()
    if a.vendor_gemm:
        print(json.dumps(vendor_gemm(3 * 4 * a.cells ** 3)), flush=True)
        return
    if a.modes_timing:
        print("--modes-timing is disabled: a profiled run of device modes followed by the batch "
              "aborted the GPU queue, and that sequence stays off until its cause is fixed",
              file=sys.stderr)
        sys.exit(2)
    if a.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path "
              f"(--gpus {a.gpus} requested)", file=sys.stderr)
        sys.exit(2)

    import fc3
    from fc3 import systems as S
    from oracle import fc3_oracle as O

    ctx = fc3.Context(0)
    cells = (a.cells,) * 3
    sysd = S.fcc(S.ARGON_A, cells, S.ARGON_MASS, 0.05, seed=2592)
    calc = S.AnalyticalCalculator(1e-12, S.ARGON_LJ.r_cut)
    n = sysd.n_dof
    log(f"workload FCC {cells} n={n}")

    # phi: IFC2 on the GPU, FP64 eigendecomposition of the dynamical matrix on the host
    dyn = fc3.dynamical_matrix(sysd, S.ARGON_LJ, calc, ctx=ctx)
    w2, phi = fc3.get_modes(dyn)
    phi32 = np.asfortranarray(phi.astype(np.float32))
    dphi = fc3.DeviceArray.from_host(phi32, ctx=ctx)
    log("phi ready")

    # ---------------- assembly ----------------
    d64 = fc3.DeviceArray((n, n, n), np.float64, ctx)
    d32 = fc3.DeviceArray((n, n, n), np.float32, ctx)
    for _ in range(a.warmup):
        fc3.third_order_device(sysd, S.ARGON_LJ, calc, psi64=d64, psi32=d32,
                               masses=sysd.masses, ctx=ctx)
    ctx.synchronize()
    profile_enable(ctx)
    ctx.timer_start()
    for _ in range(a.steps):
        fc3.third_order_device(sysd, S.ARGON_LJ, calc, psi64=d64, psi32=d32,
                               masses=sysd.masses, ctx=ctx)
    t_asm = ctx.timer_stop() / a.steps
    w_ms, w_cnt = profile(ctx, 0)
    profile_enable(ctx, False)
    asm_bytes = 12.0 * n ** 3
    asm = {"ms_per_step": t_asm, "GB_per_s": asm_bytes / (t_asm * 1e-3) / 1e9,
           "writer_kernel_ms": w_ms / max(w_cnt, 1),
           "writer_GB_per_s": asm_bytes / (w_ms / max(w_cnt, 1) * 1e-3) / 1e9,
           "bytes_per_step": asm_bytes}
    log(f"assembly {asm}")
    d64.free()

    # ---------------- mcc3 ----------------
    k3 = fc3.DeviceArray((n, n, n), np.float32, ctx)
    coo = O.lj_third_order_coo(sysd, S.ARGON_LJ, calc)
    v32 = O.coo_mass_weight(coo, sysd.masses).astype(np.float32).astype(np.float64)
    coo32 = (coo[0], coo[1], coo[2], v32)
    phi64 = phi32.astype(np.float64)
    slabs = spread_slabs(w2, n, 16)
    ref = {l: O.mcc3_fp64_slab(coo32, phi64, l) for l in slabs}
    log("oracle slabs ready")
    flop = 6.0 * n ** 4

    def slab_error():
        """Relative Frobenius error of k3's sampled slabs against the FP64 reference."""
        import ctypes
        from fc3._lib import check
        buf = np.empty((n, n), dtype=np.float32, order="F")
        num = den = 0.0
        for l in slabs:
            check(ctx.lib.fc3_memcpy_d2h(ctx.h, buf.ctypes.data_as(ctypes.c_void_p),
                                         ctypes.c_void_p(k3.ptr.value + 4 * n * n * l), 4 * n * n))
            d = buf.astype(np.float64) - ref[l]
            num += float((d * d).sum())
            den += float((ref[l] * ref[l]).sum())
        return (num / den) ** 0.5

    results = {}
    runs = [("bf16x3", a.steps)] + ([("fp32", a.fp32_steps)] if a.fp32_steps > 0 else [])
    if a.fp32_steps > 0:
        # BF16x3 with every block multiplied: the same arithmetic as the FP32 form, which
        # does not skip the operands' all-zero blocks (the like-for-like precision ratio)
        runs.append(("bf16x3_dense", a.fp32_steps))
    for key, steps in runs:
        prec = "bf16x3" if key.startswith("bf16x3") else key
        if key == "bf16x3_dense":
            os.environ["FC3_MCC3_DENSE"] = "1"
        for _ in range(a.warmup):
            fc3.mcc3(d32, dphi, precision=prec, out=k3)
        ctx.synchronize()
        profile_enable(ctx)
        ctx.timer_start()
        for _ in range(steps):
            fc3.mcc3(d32, dphi, precision=prec, out=k3)
        t = ctx.timer_stop() / steps
        m3_ms, m3_c = profile(ctx, 1)
        m12_ms, m12_c = profile(ctx, 2)
        m2_ms, m2_c = profile(ctx, 5)
        profile_enable(ctx, False)
        os.environ.pop("FC3_MCC3_DENSE", None)
        err = slab_error()
        mul = 3.0 if prec == "bf16x3" else 1.0
        results[key] = {
            "ms_per_step": t, "TFLOP_per_s": flop / (t * 1e-3) / 1e12, "rel_fro_err": err,
            "err_sample": f"K3 slabs {slabs} vs FP64 host contraction of the same FP32 Psi, phi",
            "gemm_ms_per_step": (m3_ms + m12_ms + m2_ms) / steps,
            "gemm_launches_per_step": (m3_c + m12_c + m2_c) // steps,
            "mode3_launches_per_step": m3_c / steps, "mode3_ms_per_step": m3_ms / steps,
            # BF16x3: mode 1 under "mode12" and mode 2 on its own; FP32: both under "mode12"
            "mode12_launches_per_step": m12_c / steps, "mode12_ms_per_step": m12_ms / steps,
            "mode2_launches_per_step": m2_c / steps, "mode2_ms_per_step": m2_ms / steps,
            # MFMA flops the kernels issue (3 bf16 products per FP32-equivalent product)
            "mfma_flop_per_step": mul * flop,
        }
        log(f"mcc3 {key}: {results[key]}")

    bf = results["bf16x3"]
    # Dominant kernel: k_gemm3 in mode 2, the one dense mode product (modes 3 and 1 skip
    # the all-zero blocks of Psi and of U).  Algorithmic work of those launches = 3 bf16
    # products x 2 n^4; achieved = that per launch / their average launch duration, both
    # from HIP events on the library's stream.
    m2 = max(1, int(round(bf["mode2_launches_per_step"])))
    f_launch = 3.0 * 2.0 * n ** 4 / m2
    ms_launch = bf["mode2_ms_per_step"] / m2
    achieved = f_launch / (ms_launch * 1e-3) / 1e12
    traffic = gemm_traffic()
    roofline = {"bound": "mfma", "kernel": "k_gemm3 (BF16x3), the mode-2 launches",
                "achieved": achieved, "peak": PEAK["bf16"] / 1e12, "unit": "TFLOP/s",
                "frac": achieved / (PEAK["bf16"] / 1e12), "traffic": traffic,
                "traffic_source": ("profiles/k_gemm3_lj_fcc666_traffic.md: rocprofv3 --pmc "
                                   "FETCH_SIZE and WRITE_SIZE in separate runs of this workload, "
                                   "copy-kernel calibrated; not measured in this run")
                if traffic is not None else None,
                "flop_per_launch": f_launch, "ms_per_launch": ms_launch, "launches_per_step": m2,
                "other_modes": {
                    "mode1_launches_per_step": bf["mode12_launches_per_step"],
                    "mode1_ms_per_step": bf["mode12_ms_per_step"],
                    "mode3_launches_per_step": bf["mode3_launches_per_step"],
                    "mode3_ms_per_step": bf["mode3_ms_per_step"],
                    "note": "mode 3 multiplies only the 16-wide K blocks of Psi that hold a "
                            "nonzero, and mode 1 only the rows i of U that Psi's (i, j) pattern "
                            "touches, gathered per 16-column group (the skipped products are "
                            "exact zeros; the gathering regroups the K sum, so K3 equals the "
                            "dense form to rounding); dense: all blocks (mcc3.bf16x3_dense)"}}
    roofline_assembly = {"bound": "hbm", "kernel": "k_write_ifc3_pairwise",
                         "achieved": asm["writer_GB_per_s"], "peak": PEAK["hbm"] / 1e9,
                         "unit": "GB/s", "frac": asm["writer_GB_per_s"] / (PEAK["hbm"] / 1e9),
                         "bytes_per_launch": asm["bytes_per_step"],
                         "ms_per_launch": asm["writer_kernel_ms"]}

    out = {
        "metric": METRIC, "value": bf["TFLOP_per_s"], "unit": "TFLOP/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": bf["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16x3",
        "data": "synthetic: seeded FCC LJ argon, sigma_u = 0.05 A (numpy PCG64 seed 2592)",
        "config": {"workload": f"LJ argon FCC {a.cells}x{a.cells}x{a.cells} ({n} DoF): IFC3 "
                               "assembly + mcc3 (BF16x3 headline, FP32 alongside)",
                   "n_dof": n, "n_atoms": sysd.n_atoms},
        "roofline": roofline,
        "roofline_assembly": roofline_assembly,
        "mcc3": results,
        "fp32_vs_bf16x3_speedup": (results["fp32"]["ms_per_step"] / bf["ms_per_step"]
                                   if "fp32" in results else None),
        "fp32_vs_bf16x3_dense_speedup": (
            results["fp32"]["ms_per_step"] / results["bf16x3_dense"]["ms_per_step"]
            if "bf16x3_dense" in results else None),
        "assembly": asm,
    }
    if not a.no_unique:
        out["unique"] = unique_timing(ctx, d32, dphi, k3, n, slabs, a, bf["ms_per_step"])
    d32.free()
    if not a.no_fused:
        # third_order + mcc3 without the dense Psi: sparse mode 1 from the pair blocks,
        # modes 2 and 3 in BF16x3 (same positions, masses, phi and K3 output)
        for _ in range(a.warmup):
            fc3.third_order_mcc3(sysd, S.ARGON_LJ, calc, dphi, masses=sysd.masses, out=k3,
                                 ctx=ctx)
        ctx.synchronize()
        ctx.timer_start()
        for _ in range(a.steps):
            fc3.third_order_mcc3(sysd, S.ARGON_LJ, calc, dphi, masses=sysd.masses, out=k3,
                                 ctx=ctx)
        t_f = ctx.timer_stop() / a.steps
        out["fused"] = {
            "ms_per_step": t_f, "rel_fro_err": slab_error(),
            "dense_ms_per_step": asm["ms_per_step"] + bf["ms_per_step"],
            "speedup_vs_dense": (asm["ms_per_step"] + bf["ms_per_step"]) / t_f,
            "note": "third_order + mass weighting + mcc3 with no n^3 Psi: mode 1 over Psi's "
                    "nonzeros from the pair blocks (FP32), modes 2 and 3 BF16x3; dense = "
                    "assembly (FP64 + FP32 Psi) + BF16x3 mcc3 above",
        }
        log(f"fused: {out['fused']}")
    k3.free()
    if not a.no_batched:
        out["batched"] = batched_snapshots(ctx, modes_timing=a.modes_timing)
    if not a.no_adjoint:
        out["adjoint"] = adjoint_timing(ctx)
    if not a.no_sw:
        out["sw_assembly"] = sw_assembly(ctx)
        out["sw_fused"] = sw_fused(ctx)
        out["sw_fused_648"] = sw_fused(ctx, steps=10, cells=3)
    if not a.no_cpu_baseline:
        log("cpu baselines")
        out["cpu_baseline"] = cpu_baseline_mcc3(coo32, phi32, n)
        out["cpu_baseline_assembly"] = cpu_baseline_assembly()
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
