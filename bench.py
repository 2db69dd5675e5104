"""Benchmark of the Bargmann->Fock path on one MI355X.

Headline workload: 256 three-mode Gaussian kets (A = U diag(tanh r) U^T,
r ~ U[0, 1.2], b ~ CN(0, 0.5^2)), shape (30, 30, 30) each (6.9 M complex128
amplitudes), with a CN(0, 1) cotangent: batched fused forward + VJP returning
G, dA, db, dc.  One step = one pass over the whole batch.  `value` is its
throughput in amplitudes/s with inputs resident in HBM.

Also reported (extra keys): the same batch forward only; the plain path (one
triple per launch, 256 launches per step); the uneven 4,096-ket work-queue
batch; a single 2.56 M-amplitude lossy density matrix; single mid-size
triples ((60, 60, 60), (16,)^4: forward, stable, fused, recompute per call);
and the CPU oracle
(the C restatement of vanilla_numba + vanilla_vjp_numba, OpenMP over
triples) on the headline workload.

Prints one JSON line.  `--gpus` above 1 exits: this build measures one GPU.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Fock amplitudes/s fwd & fwd+VJP, batched vs plain GPU, at rel err ≤1e-12"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def fwd_flops_box(shape) -> int:
    """Exact forward flop count of one box: sum_i (s_i-1) prod_{j>i} s_j (10(D-i)+8)."""
    D = len(shape)
    tot = 0
    for i in range(D):
        tail = int(np.prod(shape[i + 1 :])) if i + 1 < D else 1
        tot += (shape[i] - 1) * tail * (10 * (D - i) + 8)
    return tot


def vjp_flops_per_amp(D: int) -> float:
    return 8 * (D * (D + 1) / 2 + D + 1) + 1.5 * D * D + 4.5 * D


def diag_values_written(cutoffs) -> int:
    """Values the diagonal recursion stores (diagonal_amps.py:19-150 write
    conditions); each is loaded once more by its one consumer."""
    cut = np.asarray(cutoffs)
    M = len(cut)
    P = np.indices(tuple(cut)).reshape(M, -1)
    n = 0
    diag = (cut[0] == 1) | (P[0] < cut[0] - 1)
    for i in range(2 * M):
        ok = diag & (P[i // 2] + 1 < cut[i // 2])
        if i == 1:
            ok &= P[0] + 2 < cut[0]
        n += int(ok.sum())
    zero = np.ones(P.shape[1], bool)
    for d in range(M):
        ok = zero & (P[d] < cut[d] - 1)
        n += int(ok.sum()) + int((ok & (P[d] + 2 < cut[d])).sum())
        for i in range(d + 1, M):
            n += 2 * int((ok & (P[i] + 1 < cut[i])).sum())
        zero &= P[d] == 0
    return n


def traffic_from_profiles(name: str):
    """HBM bytes per launch recorded by the rocprofv3 PMC runs, if committed."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1].strip())
            except ValueError:
                return None
    return None


def counter_scale():
    """FETCH_SIZE / WRITE_SIZE scale factors measured on a 1 GiB device copy
    (recorded with the committed traffic summary), else 2 and 1."""
    path = os.path.join(ROOT, "profiles", "traffic_fused_batch_30x30x30.txt")
    if os.path.exists(path):
        import re

        m = re.search(r"FETCH_SIZE is scaled by ([\d.]+) and WRITE_SIZE by ([\d.]+)", open(path).read())
        if m:
            return float(m.group(1)), float(m.group(2))
    return 2.0, 1.0


def live_traffic(kernel_prefix: str = "void mm::lds_kernel<3, true, 512>", timeout_s: int = 240):
    """HBM bytes per launch of the headline kernel in this run: two rocprofv3
    counter passes (FETCH_SIZE, then WRITE_SIZE) over `bench.py --probe
    headline` in child processes, scaled as calibrated on a plain copy.
    Returns (bytes, note); bytes is None when rocprofv3 is unavailable or a
    pass fails."""
    import shutil
    import signal
    import sqlite3
    import subprocess
    import tempfile

    exe = shutil.which("rocprofv3")
    if exe is None:
        return None, "rocprofv3 not found"
    ff, wf = counter_scale()
    kib = {}
    for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="mm_pmc_")
        cmd = [exe, "--pmc", ctr, "-d", d, "-o", "run", "--", sys.executable, os.path.abspath(__file__), "--probe", "headline"]
        proc = subprocess.Popen(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, start_new_session=True,
                                env=dict(os.environ, TMPDIR="/tmp"))
        try:
            proc.wait(timeout=timeout_s)
        except subprocess.TimeoutExpired:
            os.killpg(proc.pid, signal.SIGKILL)
            proc.wait()
            return None, f"rocprofv3 --pmc {ctr} timed out"
        dbs = [os.path.join(r, f) for r, _, fs in os.walk(d) for f in fs if f.endswith("_results.db")]
        if not dbs:
            return None, f"rocprofv3 --pmc {ctr} wrote no database (exit {proc.returncode})"
        con = sqlite3.connect(dbs[0])
        vals = [v for name, v in con.execute(
            "select kernel_name, sum(value) from counters_collection where counter_name = ? group by dispatch_id", (ctr,))
            if name.startswith(kernel_prefix)]
        con.close()
        if not vals:
            return None, f"no {ctr} record of the headline kernel"
        kib[ctr] = sum(vals) / len(vals)
    traffic = (kib["FETCH_SIZE"] * ff + kib["WRITE_SIZE"] * wf) * 1024.0
    return traffic, (f"measured in this run: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE over bench.py --probe headline, "
                     f"{kib['FETCH_SIZE']:.0f} / {kib['WRITE_SIZE']:.0f} KiB per launch, scaled by {ff:.3f} / {wf:.3f}")


def run_probe(which: str, dev) -> None:
    """Run one workload three times without timing (kernel traces and counters)."""
    import torch

    from mrmustard_amd import workloads as W
    from mrmustard_amd.lattice import strategies as st

    if which == "headline":
        (A, b, c), shape, g = W.batched_kets_30x3()
        A, b, c, g = (torch.as_tensor(x, device=dev) for x in (A, b, c, g))
        for _ in range(3):
            st.vanilla_batch_fused_hip(shape, A, b, c, g)
        n = int(np.prod(shape)) * A.shape[0]
    elif which == "dm40":
        (A, b, c), shape = W.lossy_dm_40x4()
        A, b, c = (torch.as_tensor(np.asarray(x), device=dev) for x in (A, b, c))
        for _ in range(3):
            st.vanilla_hip(shape, A, b, c)
        n = int(np.prod(shape))
    elif which == "dm8":
        (A, b, c), shape = W.dm8_10x8()
        A, b, c = (torch.as_tensor(np.asarray(x), device=dev) for x in (A, b, c))
        gen = torch.Generator(device=dev)
        gen.manual_seed(4)
        g = torch.randn(shape, dtype=torch.complex128, device=dev, generator=gen)
        G = st.vanilla_hip(shape, A, b, c)
        st.vanilla_vjp_hip(G, c, g)
        st.vanilla_vjp_recompute_hip(shape, A, b, c, g)
        n = int(np.prod(shape))
    else:
        from mrmustard_amd.autotune import tuned_plan

        triples, shapes = W.uneven_kets()
        plan, _ = tuned_plan(shapes, fused=False)
        Au, bu, cu = plan.pack(triples)
        Gu = torch.empty(plan.n_amplitudes, dtype=torch.complex128, device=dev)
        for _ in range(3):
            plan.run(Au, bu, cu, G=Gu, graph=True)
        n = plan.n_amplitudes
    torch.cuda.synchronize()
    print(json.dumps({"probe": which, "amplitudes": n, "calls": 3}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--quick", action="store_true", help="headline measurement only (for profiling)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-live-traffic", action="store_true",
                    help="take roofline.traffic from the committed counter summary instead of two rocprofv3 passes in this run")
    ap.add_argument(
        "--copy-probe",
        action="store_true",
        help="only run a plain 1 GiB device-to-device copy (calibrates the HBM byte counters)",
    )
    ap.add_argument(
        "--diagonal-probe",
        action="store_true",
        help="only run the diagonal strategy on the 6-mode density matrix, cutoffs 10, three times (for counters)",
    )
    ap.add_argument(
        "--leftover-probe",
        action="store_true",
        help="only run the one-leftover-mode forward and its VJP on the 3-mode density matrix "
             "(20 x 20 x 15 x 15), three times each (for kernel traces)",
    )
    ap.add_argument(
        "--probe",
        choices=["headline", "dm40", "dm8", "uneven"],
        help="only run one workload three times, untimed (for counter runs): the fused 256 x (30,30,30) batch, "
             "the (40,)^4 forward, the (10,)^8 forward + VJP + recompute VJP, or the 4,096-ket plan",
    )
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build", file=sys.stderr)
        sys.exit(2)

    import torch

    from mrmustard_amd import workloads as W
    from mrmustard_amd.lattice import strategies as st

    torch.cuda.set_device(0)
    dev = torch.device("cuda", 0)
    if args.copy_probe:
        n = 1 << 27  # 2^27 float64 = 1 GiB
        x = torch.ones(n, dtype=torch.float64, device=dev)
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8, "launches": 3}))
        return
    if args.diagonal_probe:
        from mrmustard_amd.lattice.compact_fock import diagonal_hip

        Ad, bd, cd = W.pnr_dm(6)
        Adt, bdt, cdt = (torch.as_tensor(np.asarray(x), device=dev) for x in (Ad, bd, cd))
        for _ in range(3):
            diagonal_hip(Adt, bdt, cdt, (10,) * 6, reordered=True)
        torch.cuda.synchronize()
        print(json.dumps({"diagonal_probe_values_written": diag_values_written((10,) * 6), "calls": 3}))
        return
    if args.leftover_probe:
        from mrmustard_amd.lattice.compact_fock import leftover_hip, leftover_vjp_hip

        Al, bl, cl = W.pnr_dm(3, seed=7)
        Alt, blt, clt = (torch.as_tensor(np.asarray(x), device=dev) for x in (Al, bl, cl))
        cot = torch.ones((20, 20, 15, 15), dtype=torch.complex128, device=dev)
        for _ in range(3):
            leftover_hip(Alt, blt, clt, 19, (14, 14))
        for _ in range(3):
            leftover_vjp_hip(Alt, blt, clt, 19, (14, 14), cot)
        torch.cuda.synchronize()
        print(json.dumps({"leftover_probe_calls": 3, "vjp_calls": 3}))
        return
    if args.probe:
        run_probe(args.probe, dev)
        return
    log = lambda *a: print("[bench]", *a, file=sys.stderr, flush=True)  # noqa: E731

    (A_np, b_np, c_np), shape, g_np = W.batched_kets_30x3()
    B = A_np.shape[0]
    D = len(shape)
    N = int(np.prod(shape))
    amps = B * N
    A = torch.as_tensor(A_np, device=dev)
    b = torch.as_tensor(b_np, device=dev)
    c = torch.as_tensor(c_np, device=dev)
    g = torch.as_tensor(g_np, device=dev)
    G = torch.empty((B, *shape), dtype=torch.complex128, device=dev)
    stream = torch.cuda.current_stream()

    def fused_step():
        return st.vanilla_batch_fused_hip(shape, A, b, c, g, out=G)

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 1e3 / steps  # seconds per step

    # ---- headline: batched fused forward + VJP
    log("headline: batched fused fwd+VJP, 256 x (30,30,30)")
    t_fused = timed(fused_step, args.steps, args.warmup)
    value = amps / t_fused

    # outputs of the measured configuration, checked against the oracle in the CPU leg
    Gd, dA, db, dc = fused_step()
    torch.cuda.synchronize()
    Gh, dAh, dbh, dch = (x.cpu().numpy() for x in (Gd, dA, db, dc))

    bytes_per_launch = 32.0 * amps
    flops_per_launch = B * (fwd_flops_box(shape) + vjp_flops_per_amp(D) * N)
    achieved_gbs = bytes_per_launch / t_fused / 1e9
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved_gbs, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved_gbs / HBM_PEAK_GBS, 4),
        "traffic": traffic_from_profiles("traffic_fused_batch_30x30x30.txt"),
        "traffic_source": "profiles/traffic_fused_batch_30x30x30.txt",
        "kernel": "lds_kernel<3, true, 512> (uniform mode, one workgroup per triple)",
        "fp64_tflops": round(flops_per_launch / t_fused / 1e12, 2),
        "fp64_frac": round(flops_per_launch / t_fused / 1e12 / FP64_PEAK_TFLOPS, 4),
    }
    extra = {}

    if not args.quick:
        # ---- forward only, same batch
        log("batched forward only")
        t_fwd = timed(lambda: st.vanilla_batch_hip(shape, A, b, c, out=G), args.steps, args.warmup)
        extra["batched_fwd_amps_per_s"] = amps / t_fwd
        # ---- plain: one triple per launch
        log("plain: one triple per launch")
        Gk = [G[k] for k in range(B)]

        def plain_step():
            for k in range(B):
                st.vanilla_fused_hip(shape, A[k], b[k], c[k], g[k], out=Gk[k])

        t_plain = timed(plain_step, max(2, args.steps // 50), 1)
        extra["plain_fused_amps_per_s"] = amps / t_plain
        extra["batched_vs_plain_fused"] = round(t_plain / t_fused, 2)

        def plain_fwd_step():
            for k in range(B):
                st.vanilla_hip(shape, A[k], b[k], c[k], out=Gk[k])

        t_plain_fwd = timed(plain_fwd_step, max(2, args.steps // 50), 1)
        extra["plain_fwd_amps_per_s"] = amps / t_plain_fwd
        extra["batched_vs_plain_fwd"] = round(t_plain_fwd / t_fwd, 2)
        # ---- uneven work-queue batch (4,096 kets, m in 1..3, cutoffs 8..60)
        log("uneven 4096-ket work-queue batch")
        from mrmustard_amd.autotune import tuned_plan

        # another draw of the same distribution first: the measured batch then
        # finds every one of its classes in the tuning table
        _, shapes_other = W.uneven_kets(seed=102)
        other, other_info = tuned_plan(shapes_other, fused=False)
        other.close()
        triples, shapes = W.uneven_kets()
        plan, tune_info = tuned_plan(shapes, fused=False)
        tune_info["other_draw"] = {k: other_info[k] for k in ("tune", "cached", "classes", "classes_timed")}
        Au, bu, cu = plan.pack(triples)
        Gu = torch.empty(plan.n_amplitudes, dtype=torch.complex128, device=dev)
        t_unev = timed(lambda: plan.run(Au, bu, cu, G=Gu, graph=True), max(5, args.steps // 10), 2)
        extra["uneven_batch_fwd_amps_per_s"] = plan.n_amplitudes / t_unev
        extra["uneven_batch_amplitudes"] = plan.n_amplitudes
        extra["uneven_batch_tune"] = tune_info
        # ---- plain large lattice: 2-mode lossy density matrix, shape (40,)^4
        log("plain lossy DM (40,)^4")
        (A1, b1, c1), shape1 = W.lossy_dm_40x4()
        A1t, b1t, c1t = (torch.as_tensor(np.asarray(x), device=dev) for x in (A1, b1, c1))
        G1 = torch.empty(shape1, dtype=torch.complex128, device=dev)
        t_dm = timed(lambda: st.vanilla_hip(shape1, A1t, b1t, c1t, out=G1), 20, 2)
        extra["dm_40x4_fwd_amps_per_s"] = G1.numel() / t_dm
        # the all-pivot average (stable_numba, core.py:127-213) on the same lattice
        t_st = timed(lambda: st.stable_hip(shape1, A1t, b1t, c1t, out=G1), 5, 1)
        extra["dm_40x4_stable_fwd_amps_per_s"] = G1.numel() / t_st
        del G1
        # ---- 4-mode lossy density matrix, shape (10,)^8 = 1e8 amplitudes:
        #      forward, VJP of the stored G, and the VJP that recomputes G in a
        #      three-slab window
        log("4-mode DM (10,)^8: forward, VJP, recompute VJP")
        (A4, b4, c4), shape4 = W.dm8_10x8()
        A4t, b4t, c4t = (torch.as_tensor(np.asarray(x), device=dev) for x in (A4, b4, c4))
        G4 = torch.empty(shape4, dtype=torch.complex128, device=dev)
        gen = torch.Generator(device=dev)
        gen.manual_seed(4)
        g4 = torch.randn(shape4, dtype=torch.complex128, device=dev, generator=gen)
        n4 = G4.numel()
        t4f = timed(lambda: st.vanilla_hip(shape4, A4t, b4t, c4t, out=G4), 5, 1)
        t4v = timed(lambda: st.vanilla_vjp_hip(G4, c4t, g4), 5, 1)
        t4r = timed(lambda: st.vanilla_vjp_recompute_hip(shape4, A4t, b4t, c4t, g4), 5, 1)
        r1 = st.vanilla_vjp_hip(G4, c4t, g4)
        r2 = st.vanilla_vjp_recompute_hip(shape4, A4t, b4t, c4t, g4)
        agree = max(float((x - y).abs().max() / x.abs().max()) for x, y in zip(r1, r2))
        dc_identity = abs(complex(r1[2]) * complex(c4) - complex(torch.sum(G4 * g4))) / abs(complex(torch.sum(G4 * g4)))
        del g4, r1, r2
        # contraction of the density matrix, (bra) x (ket) = 1e4 x 1e4, with a
        # projector P = V V^H onto a random 16-dimensional subspace of the
        # 4-mode number basis: P @ rho on the FP64 matrix cores
        from mrmustard_amd.contraction import dm_operator_product, zgemm_hip

        log("4-mode DM (10,)^8: projector contraction on MFMA")
        dmat = 10**4
        V = torch.randn(dmat, 16, dtype=torch.complex128, device=dev, generator=gen)
        for j in range(16):  # Gram-Schmidt: orthonormal columns
            for i in range(j):
                V[:, j] -= torch.vdot(V[:, i], V[:, j]) * V[:, i]
            V[:, j] /= torch.linalg.vector_norm(V[:, j])
        P = zgemm_hip(V, V.conj().T.contiguous())
        Cp = torch.empty_like(G4)
        t4c = timed(lambda: dm_operator_product(P, G4, out=Cp), 3, 1)
        rows = slice(0, 64)
        ref_rows = torch.matmul(P[rows], G4.reshape(dmat, dmat))
        c_err = float((Cp.reshape(dmat, dmat)[rows] - ref_rows).abs().max() / ref_rows.abs().max())
        zflops = 8.0 * dmat**3
        extra["dm8_10x8"] = {
            "amplitudes": n4,
            "fwd_amps_per_s": n4 / t4f,
            "vjp_amps_per_s": n4 / t4v,
            "fwd_plus_vjp_amps_per_s": n4 / (t4f + t4v),
            "recompute_vjp_amps_per_s": n4 / t4r,
            "recompute_resident_bytes": 3 * 16 * n4 // shape4[0],
            "stored_vs_recompute_rel_err": agree,
            "dc_times_c_vs_sum_G_g_rel_err": dc_identity,
            "contraction": {
                "op": "P @ rho, complex128 10^4 x 10^4 x 10^4 (zgemm_mfma_kernel)",
                "ms": t4c * 1e3,
                "tflops": zflops / t4c / 1e12,
                "mfma_fp64_frac": zflops / t4c / 1e12 / FP64_PEAK_TFLOPS,
                "rel_err_first_64_rows_vs_torch_matmul": c_err,
            },
        }
        del Cp, P, V
        # a two-mode operator on modes 0, 1 of the same density matrix (ArrayAnsatz.contract labels):
        # bra side (no permutation: one 100 x 10^6 x 100 product) and ket side (operand and result permuted)
        log("4-mode DM (10,)^8: two-mode operator, bra and ket side (contract)")
        from mrmustard_amd.contraction import contract

        U2 = torch.randn(10, 10, 10, 10, dtype=torch.complex128, device=dev)
        t_bra = timed(lambda: contract(U2, G4, ["A", "B", 0, 1], [0, 1, 2, 3, 4, 5, 6, 7], ["A", "B", 2, 3, 4, 5, 6, 7]), 3, 1)
        t_ket = timed(lambda: contract(G4, U2, [0, 1, 2, 3, 4, 5, 6, 7], [4, 5, "A", "B"], [0, 1, 2, 3, "A", "B", 6, 7]), 3, 1)
        gflop = 8.0 * 100 * 10**6 * 100 / 1e9
        extra["dm8_10x8"]["two_mode_operator"] = {
            "bra_side_ms": t_bra * 1e3,
            "bra_side_tflops": gflop / t_bra / 1e3,
            "ket_side_ms": t_ket * 1e3,
            "ket_side_tflops": gflop / t_ket / 1e3,
            "note": "8e10 flop per product; the ket side reads rho in place as 10^4 products (100 x 100 x 100) over its bra indices",
        }
        del G4
        # the conversion fused with the same kind of contraction: wires 6, 7 of
        # the density matrix contracted with U2 as each slab of the lattice is
        # made, against the whole lattice followed by contract
        log("4-mode DM (10,)^8: conversion fused with a two-mode operator")
        from mrmustard_amd.lattice.fused import fock_contract

        labs, lab2, labo = list(range(8)), [6, 7, "A", "B"], [0, 1, 2, 3, 4, 5, "A", "B"]

        def unfused_op():
            return contract(st.vanilla_hip(shape4, A4t, b4t, c4t), U2, labs, lab2, labo)

        def fused_op():
            return fock_contract(A4t, b4t, c4t, shape4, labs, U2, lab2, labo)

        def peak_bytes(fn):
            torch.cuda.synchronize()
            base = torch.cuda.memory_allocated()
            torch.cuda.reset_peak_memory_stats()
            r = fn()
            torch.cuda.synchronize()
            return r, torch.cuda.max_memory_allocated() - base

        r_un, m_un = peak_bytes(unfused_op)
        r_fu, m_fu = peak_bytes(fused_op)
        fc_err = float((r_fu - r_un).abs().max() / r_un.abs().max())
        del r_un, r_fu
        t_un = timed(unfused_op, 3, 1)
        t_fu = timed(fused_op, 3, 1)
        extra["dm8_10x8"]["fused_conversion_contraction"] = {
            "op": "fock_contract: lattice of the (10,)^8 DM, wires 6, 7 contracted with a (10,10,10,10) operator",
            "fused_ms": t_fu * 1e3,
            "unfused_ms": t_un * 1e3,
            "fused_peak_bytes": int(m_fu),
            "unfused_peak_bytes": int(m_un),
            "rel_err_fused_vs_unfused": fc_err,
            "note": "peak = device memory allocated during the call above what was allocated before it (output included)",
        }
        del U2
        # ---- binomial strategy (fill by weight, L2 stop) on one (30,30,30) ket
        log("binomial: one (30,30,30) ket, max_l2 = 0.99999")
        A0, b0, c0 = A[0], b[0], c[0]
        t_bin = timed(lambda: st.binomial_hip(shape, A0, b0, c0, 0.99999, sum(shape) - len(shape) + 1), 20, 2)
        extra["binomial_30x30x30"] = {"amps_per_s": N / t_bin, "us_per_call": t_bin * 1e6,
                                      "note": "one triple per launch; includes the copy of the final norm to the host"}
        # ---- single mid-size triples: the schedule is chosen per call (one workgroup, or all CUs)
        log("single triples: (60,60,60) and (16,16,16,16)")
        single = {}
        rng_s = np.random.default_rng(11)
        for shp in [(60, 60, 60), (16, 16, 16, 16)]:
            Ds = len(shp)
            As = rng_s.normal(size=(Ds, Ds)) + 1j * rng_s.normal(size=(Ds, Ds))
            As = 0.2 * (As + As.T) / np.max(np.abs(np.linalg.eigvals(As + As.T)))
            bs = 0.3 * (rng_s.normal(size=Ds) + 1j * rng_s.normal(size=Ds))
            Ast, bst, cst = (torch.as_tensor(np.asarray(x), device=dev) for x in (As, bs, 1.0 + 0j))
            gs = torch.ones(shp, dtype=torch.complex128, device=dev)
            Gf = st.vanilla_hip(shp, Ast, bst, cst)
            Gu, *grads_f = st.vanilla_fused_hip(shp, Ast, bst, cst, gs)
            grads_r = st.vanilla_vjp_recompute_hip(shp, Ast, bst, cst, gs)
            rel_g = max(float((x - y).abs().max() / y.abs().max()) for x, y in zip(grads_r, grads_f))
            single["x".join(map(str, shp))] = {
                "fwd_us": timed(lambda: st.vanilla_hip(shp, Ast, bst, cst), 10, 2) * 1e6,
                "stable_fwd_us": timed(lambda: st.stable_hip(shp, Ast, bst, cst), 5, 1) * 1e6,
                "fused_fwd_vjp_us": timed(lambda: st.vanilla_fused_hip(shp, Ast, bst, cst, gs), 10, 2) * 1e6,
                "recompute_vjp_us": timed(lambda: st.vanilla_vjp_recompute_hip(shp, Ast, bst, cst, gs), 10, 2) * 1e6,
                "fused_G_equals_forward": bool(torch.equal(Gu, Gf)),
                "recompute_vs_fused_grads_rel_err": rel_g,
            }
            del gs, Gf, Gu
        single["note"] = ("one triple per call (API calls: workspace allocation included; the recompute call reads dL/dc back); "
                          "before the per-call schedule choice: (60,60,60) fwd 230, stable 2260, fused 825 us; "
                          "(16,)^4 fwd 109, stable 1180, fused 303 us (one workgroup per triple)")
        extra["single_triples"] = single
        # ---- autoshape of 4,096 four-mode lossy Gaussian density matrices
        log("autoshape: 4096 four-mode density matrices")
        from mrmustard_amd.lattice.autoshape import autoshape_batch_hip

        Aas, bas, cas = W.autoshape_states()
        Aat, bat, cat = (torch.as_tensor(x, device=dev) for x in (Aas, bas, cas))
        t_as = timed(lambda: autoshape_batch_hip(Aat, bat, cat, 0.99999, 50, 1), 20, 2)
        shapes_as = autoshape_batch_hip(Aat, bat, cat, 0.99999, 50, 1)
        extra["autoshape_4096x4modes"] = {
            "states_per_s": len(cas) / t_as,
            "ms_per_batch": t_as * 1e3,
            "mean_cutoff": float(shapes_as.mean()),
            "note": "includes the copy of the cutoffs to the host, which the caller needs to size the Fock array",
        }

        # ---- diagonal strategy: joint photon-number probabilities of a 6-mode density matrix
        log("diagonal: 6-mode lossy density matrix, cutoffs 10 per mode")
        from mrmustard_amd.lattice.compact_fock import diagonal_hip

        Ad, bd, cd = W.pnr_dm(6)
        Adt, bdt, cdt = (torch.as_tensor(np.asarray(x), device=dev) for x in (Ad, bd, cd))
        cut_d = (10,) * 6
        t_d = timed(lambda: diagonal_hip(Adt, bdt, cdt, cut_d, reordered=True), 10, 2)
        p_d = diagonal_hip(Adt, bdt, cdt, cut_d, reordered=True)
        alg_d = 32 * diag_values_written(cut_d)
        from mrmustard_amd.lattice.compact_fock import diagonal_vjp_hip

        cot_d = torch.ones(cut_d, dtype=torch.complex128, device=dev)
        t_dv = timed(lambda: diagonal_vjp_hip(Adt, bdt, cdt, cut_d, cot_d, reordered=True), 5, 1)
        extra["diagonal_6modes_cutoff10"] = {
            "probabilities_per_s": float(np.prod(cut_d)) / t_d,
            "ms_per_call": t_d * 1e3,
            "launches": sum(cut_d) - len(cut_d) + 2,
            "algorithmic_bytes": alg_d,
            "achieved_GB_per_s": alg_d / t_d / 1e9,
            "hbm_frac": alg_d / t_d / 1e9 / HBM_PEAK_GBS,
            "sum_of_probabilities": float(p_d.real.sum()),
            "fwd_plus_vjp_ms_per_call": t_dv * 1e3,
            "note": "API call: input validation, workspace, one launch per total photon number; the VJP call "
                    "runs the forward again, then one reverse launch per weight",
        }

        # ---- one leftover mode: heralded conditional density matrices of a 3-mode density matrix
        log("1leftoverMode: 3-mode lossy density matrix, mode 0 up to 19 photons, two PNR modes up to 14")
        from mrmustard_amd.lattice.compact_fock import leftover_hip

        Al, bl, cl = W.pnr_dm(3, seed=7)
        Alt, blt, clt = (torch.as_tensor(np.asarray(x), device=dev) for x in (Al, bl, cl))
        o_l, pnr_l = 19, (14, 14)
        t_l = timed(lambda: leftover_hip(Alt, blt, clt, o_l, pnr_l), 10, 2)
        n_l = (o_l + 1) ** 2 * int(np.prod([k + 1 for k in pnr_l]))
        from mrmustard_amd.lattice.compact_fock import leftover_vjp_hip

        cot_l = torch.ones((o_l + 1, o_l + 1, *(k + 1 for k in pnr_l)), dtype=torch.complex128, device=dev)
        t_lv = timed(lambda: leftover_vjp_hip(Alt, blt, clt, o_l, pnr_l, cot_l), 5, 1)
        extra["leftover_3modes_20x20x15x15"] = {
            "entries_per_s": n_l / t_l,
            "ms_per_call": t_l * 1e3,
            "entries": n_l,
            "launches": 2 * (sum(pnr_l) + 1) + 2,
            "fwd_plus_vjp_ms_per_call": t_lv * 1e3,
            "note": "API call: input validation, workspace, two launches per total detected photon number; the VJP "
                    "call runs the forward again, then two reverse launches per weight",
        }

        # ---- derived variables: the conversion fused with the contraction with c
        log("derived variables: (40,40) kept, (25,25) derived, fused vs lattice + product")
        from mrmustard_amd.lattice.derived import fock_array_derived

        rng_d = np.random.default_rng(11)
        Xd = rng_d.standard_normal((4, 4)) + 1j * rng_d.standard_normal((4, 4))
        Adv = torch.as_tensor(0.15 * (Xd + Xd.T), device=dev)
        bdv = torch.as_tensor(0.3 * (rng_d.standard_normal(4) + 1j * rng_d.standard_normal(4)), device=dev)
        cdv = torch.as_tensor(rng_d.standard_normal((25, 25)) + 1j * rng_d.standard_normal((25, 25)), device=dev)
        t_fu = timed(lambda: fock_array_derived(Adv, bdv, cdv, (40, 40), fused=True), 10, 2)
        t_un = timed(lambda: fock_array_derived(Adv, bdv, cdv, (40, 40), fused=False), 10, 2)
        r_fu = fock_array_derived(Adv, bdv, cdv, (40, 40), fused=True)
        r_un = fock_array_derived(Adv, bdv, cdv, (40, 40), fused=False)
        extra["derived_40x40_25x25"] = {
            "lattice_amplitudes": 40 * 40 * 25 * 25,
            "fused_ms_per_call": t_fu * 1e3,
            "unfused_ms_per_call": t_un * 1e3,
            "fused_resident_bytes": 3 * 40 * 25 * 25 * 16,
            "unfused_resident_bytes": 40 * 40 * 25 * 25 * 16,
            "rel_err_fused_vs_unfused": float((r_fu - r_un).abs().max() / r_un.abs().max()),
            "note": "one triple; fused = mm_hermite_fwd_contract (three-slab window, each slab contracted as made); "
                    "unfused = whole lattice, then a product on the matrix cores",
        }

    if not args.quick and not args.no_live_traffic:
        log("HBM traffic of the headline kernel: rocprofv3 FETCH_SIZE / WRITE_SIZE passes in child processes")
        torch.cuda.synchronize()
        t_bytes, note = live_traffic()
        if t_bytes is not None:
            roofline["traffic"] = round(t_bytes, 1)
            roofline["traffic_source"] = note
        else:
            roofline["traffic_source"] += f" (live measurement unavailable: {note})"

    cpu = None
    parity = None
    if not args.no_cpu and not args.quick:
        # ---- CPU leg: the oracle as the checker of the measured outputs, then timed
        import oracle
        from oracle import autoshape as oracle_autoshape

        log("parity of the headline batch against the oracle (sample of triples)")

        def rel(x, r):
            return float(np.max(np.abs(x - r)) / np.max(np.abs(r)))

        def rel_elementwise(x, r):
            big = np.abs(r) > 1e-3 * np.max(np.abs(r))
            return float(np.max(np.abs(x[big] - r[big]) / np.abs(r[big])))

        worst = {"G": 0.0, "grads": 0.0, "G_elementwise": 0.0, "grads_elementwise": 0.0}
        checked = list(range(0, B, 37))
        for k in checked:
            ref = oracle.vanilla(shape, A_np[k], b_np[k], c_np[k])
            worst["G"] = max(worst["G"], rel(Gh[k], ref))
            worst["G_elementwise"] = max(worst["G_elementwise"], rel_elementwise(Gh[k], ref))
            for x, r in zip((dAh[k], dbh[k], dch[k]), oracle.vanilla_vjp(ref, c_np[k], g_np[k], kahan=True)):
                worst["grads"] = max(worst["grads"], rel(np.atleast_1d(x), np.atleast_1d(r)))
                worst["grads_elementwise"] = max(worst["grads_elementwise"], rel_elementwise(np.atleast_1d(x), np.atleast_1d(r)))
        parity = {
            "max_rel_err_G": worst["G"],
            "max_rel_err_grads": worst["grads"],
            "max_rel_err_G_elementwise_above_1e-3_max": worst["G_elementwise"],
            "max_rel_err_grads_elementwise_above_1e-3_max": worst["grads_elementwise"],
            "checked_triples": len(checked),
        }
        if "autoshape_4096x4modes" in extra:
            n_as = 256
            t0 = time.perf_counter()
            cuts = [oracle_autoshape.autoshape(Aas[i], bas[i], cas[i], 0.99999, 50, 1) for i in range(n_as)]
            t_cpu_as = time.perf_counter() - t0
            extra["autoshape_4096x4modes"]["cpu_oracle_states_per_s"] = n_as / t_cpu_as
            extra["autoshape_4096x4modes"]["cpu_oracle_sample"] = f"first {n_as} states, numpy restatement, one core"
            extra["autoshape_4096x4modes"]["equal_to_oracle_on_sample"] = bool(np.array_equal(np.stack(cuts), shapes_as[:n_as]))
        if "diagonal_6modes_cutoff10" in extra:
            from oracle import diagonal as oracle_diagonal

            cut_s = (4,) * 6
            Ar_s, br_s = oracle_diagonal.reorder_ab(Ad, bd)
            t0 = time.perf_counter()
            p_ref = oracle_diagonal.diagonal(Ar_s, br_s, cd, cut_s)
            t_cpu_d = time.perf_counter() - t0
            p_gpu = diagonal_hip(Adt, bdt, cdt, cut_s, reordered=True).cpu().numpy()
            extra["diagonal_6modes_cutoff10"]["cpu_oracle_probabilities_per_s"] = float(np.prod(cut_s)) / t_cpu_d
            extra["diagonal_6modes_cutoff10"]["cpu_oracle_sample"] = "same state at cutoffs 4 per mode, Python restatement, one core"
            extra["diagonal_6modes_cutoff10"]["rel_err_vs_oracle_on_sample"] = rel(p_gpu, p_ref)
        if "leftover_3modes_20x20x15x15" in extra:
            from oracle import leftover as oracle_leftover

            o_s, pnr_s = 5, (4, 4)
            t0 = time.perf_counter()
            l_ref = oracle_leftover.leftover(Al, bl, cl, o_s, pnr_s)
            t_cpu_l = time.perf_counter() - t0
            l_gpu = leftover_hip(Alt, blt, clt, o_s, pnr_s).cpu().numpy()
            ex = extra["leftover_3modes_20x20x15x15"]
            ex["cpu_oracle_entries_per_s"] = l_ref.size / t_cpu_l
            ex["cpu_oracle_sample"] = "same state, mode 0 up to 5 photons, PNR modes up to 4, Python restatement, one core"
            ex["rel_err_vs_oracle_on_sample"] = rel(l_gpu, l_ref)
        log("CPU oracle baseline")
        threads = oracle.max_threads()
        triples = [(A_np[k], b_np[k], c_np[k]) for k in range(B)]
        shapes = [shape] * B
        gs = [g_np[k] for k in range(B)]
        reps, t0 = 0, time.perf_counter()
        oracle.uneven(triples[:2], shapes[:2], gs[:2], host_tuned=True)  # build the host-tuned copy untimed
        while True:
            oracle.uneven(triples, shapes, gs, host_tuned=True)
            reps += 1
            if time.perf_counter() - t0 > 5.0:
                break
        wall = time.perf_counter() - t0
        cpu = {
            "value": reps * amps / wall,
            "unit": "amplitudes/s",
            "cores": threads,
            "kind": "port",
            "sample": f"the full headline batch (256 x (30,30,30), forward + VJP) x {reps} repetitions, "
            f"{wall:.1f} s wall, OpenMP over triples; gcc -O3 -march=native -ffp-contract=off built on this host",
        }

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "amplitudes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_fused * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "batched fused forward+VJP: 256 three-mode Gaussian kets, shape (30,30,30), complex128",
            "triples": B,
            "shape": list(shape),
            "amplitudes_per_step": amps,
            "parallelism": "single GPU",
        },
        "parity": parity,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
