#!/usr/bin/env python3
"""Benchmark: PCG solve to 1e-8 relative residual on the 128^3 heterogeneous
7-point operator (log10 k ~ N(0, 1.5^2), harmonic transmissibilities,
Dirichlet faces), hypre-default hierarchy (aggressive HMIS + HMIS, ext+i
trunc 0.3) with l1-Jacobi V(1,1), FP64 outer PCG.

One step = one PCG solve from x = 0 with the hierarchy and b resident in HBM.
The headline value is seconds per solve with the FP32 cycle; the FP64-cycle
and BF16-value-cycle solve times and speedups, iteration counts, V-cycle time
and bandwidth per precision, the roofline of the dominant kernel, each GPU
solution's error against the CPU oracle's FP64 solution, and a CPU baseline
(the oracle's OpenMP restatement of the same solve: one warm-up, then the
median of full solves) ride along in the line.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 128] [--precision fp32|bf16]
       python bench.py --kernel-only   (setup, a 256 MiB calibration copy and the
                                        dominant-kernel launches only: the run the
                                        rocprofv3 --pmc passes profile; with
                                        --workload batched48 the batched pass)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "draugr.jl_amd"))
sys.path.insert(0, ROOT)

METRIC = "s/PCG solve to 1e-8 rel resid, 128³ 7-pt; FP32-cycle vs FP64 speedup; GB/s"
HBM_PEAK_GBS = 8000.0
SEED = 20261202


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--sigma", type=float, default=1.5)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "bf16"])
    ap.add_argument("--cpu-solves", type=int, default=5,
                    help="timed full CPU oracle PCG solves (after one warm-up); the baseline is their median")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--workload", default="hetero128",
                    choices=["hetero128", "poisson256", "resetup192", "batched48", "adjoint96"],
                    help="hetero128 (default): the headline line; the others print their own JSON line")
    ap.add_argument("--kernel-only", action="store_true",
                    help="only the calibration copy and the dominant-kernel launches (for counter runs)")
    ap.add_argument("--kernel", default="auto", choices=["auto", "stream", "rowseq", "pipe", "stream16", "sorted", "stream16pipe",
                                                  "sortedpipe", "sortedquad", "vector", "pattern"],
                    help="sparse-pass kernel (config key kernel); a counter run forces the kernel the "
                         "autotuner picks without a profiler, whose timings it would perturb")
    ap.add_argument("--fuse", default="auto", choices=["auto", "off", "on"],
                    help="zero-guess and one-pass sweep fusions (config key fuse); fixed, like --kernel, "
                         "in profiled runs")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "stream_jacobi_level1_128cubed_traffic.md"),
                    help="summary of the FETCH_SIZE / WRITE_SIZE passes holding the measured HBM bytes per launch")
    return ap.parse_args()


def measured_traffic(path, kernel_tag):
    """HBM bytes per launch of the dominant kernel from the committed counter
    summary (a line 'traffic_bytes_per_launch <tag> <bytes>'), or None."""
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel_tag:
                    return float(parts[2])
    except OSError:
        pass
    return None


def dominant_kernel(torch, h, prec, reps):
    """The largest launch of the cycle: one Jacobi / l1-Jacobi step on level 1
    (the densest operator of the hierarchy), in the cycle's precision, timed
    with events on torch's current stream, which is the stream the library
    launches it on.  Returns (seconds per launch, algorithmic bytes, bytes as
    stored, tag)."""
    n = h.level_info(1)["n"]
    xa = torch.randn(n, dtype=torch.float32, device="cuda")
    ba = torch.randn(n, dtype=torch.float32, device="cuda")
    oa = torch.empty(n, dtype=torch.float32, device="cuda")
    # back-to-back launches between two events, issued by the library in one
    # call (no host round trip between them): the per-launch time a cycle sees,
    # and the average kernel duration a kernel trace of this run reports
    h.level_op_repeat(1, 2, prec, 2, xa, ba, oa)
    t = min(event_time(torch, lambda: h.level_op_repeat(1, 2, prec, reps, xa, ba, oa), 1) / reps for _ in range(3))
    # algorithmic bytes: A_1 once with int32 indices (the count is of the work,
    # whatever the index form), x, b, dinv read, x_out written; the bytes the
    # tuned kernel actually reads (16-bit block-relative columns) are reported
    # beside them
    nbytes = int(h.op_bytes(1, 0, prec, as_stored=False)) + 4 * n * 4
    stored = int(h.op_bytes(1, 0, prec, as_stored=True)) + 4 * n * 4
    kind = "stream"
    for line in h.tune_report().splitlines():  # the kernel the autotuner chose for this pass
        if line.startswith("%s level 1: A=" % prec):
            kind = line.split("A=")[1].split()[0]
    return t, nbytes, stored, "level1_jacobi_%s_%s_grid%d" % (prec, kind, h_blocks(h, 1))


KERNEL_OF_KIND = {"stream": "k_csr_stream", "stream16": "k_csr_stream", "sorted": "k_csr_stream",
                  "rowseq": "k_csr_rowseq", "pipe": "k_csr_pipe", "stream16pipe": "k_stream_pipe",
                  "sortedpipe": "k_stream_pipe", "sortedquad": "k_csr_stream_quad", "vector": "k_csr_vector",
                  "pattern": "k_csr_pattern"}


def h_blocks(h, l):
    return int(h.level_info(l)["nblk"])


def calibration_copy(torch):
    """A plain 256 MiB device-to-device copy: the counter runs scale
    FETCH_SIZE / WRITE_SIZE by what they read for these known bytes."""
    src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()


def cpu_model():
    """CPU model name (lscpu, else /proc/cpuinfo) and the logical CPUs visible."""
    name = None
    try:
        import subprocess
        for line in subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout.splitlines():
            if line.startswith("Model name:"):
                name = line.split(":", 1)[1].strip()
                break
    except Exception:
        pass
    if name is None:
        try:
            with open("/proc/cpuinfo") as f:
                for line in f:
                    if line.startswith("model name"):
                        name = line.split(":", 1)[1].strip()
                        break
        except OSError:
            name = "unknown"
    return "%s (%d logical CPUs visible)" % (name, os.cpu_count() or 0)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_solves(dh, torch, h, cfg, b, steps, warmup, tol):
    x = torch.zeros_like(b)
    its = []
    for _ in range(warmup):
        x.zero_()
        dh.pcg_solve_(x, b, h, cfg, tol=tol, maxiter=500)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        x.zero_()
        _, it, rel = dh.pcg_solve_(x, b, h, cfg, tol=tol, maxiter=500)
        its.append(it)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    return dt, its, rel, x


def event_time(torch, fn, reps):
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / 1e3 / reps


def timed(torch, fn):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t0, out


def workload_poisson256(args, torch, dh):
    """2-D 5-point Poisson 256^2, aggregation (theta 0.25), Jacobi (2/3) V(1,1),
    max_coarse_size 50, RHS U(-1,1); FP64 PCG to 1e-8 on the GPU and the
    oracle's CPU PCG on the same hierarchy."""
    P = dh.problems
    A = P.poisson2d(256)
    b_host = 2.0 * P.splitmix_uniform(SEED + 0, A.n) - 1.0
    cfg = dh.AMGConfig(coarsening=dh.AggregationCoarsening(0.25), smoother=dh.JacobiSmootherType(),
                       jacobi_omega=2.0 / 3.0, max_coarse_size=50)
    h = dh.amg_setup(A, cfg)
    b = torch.tensor(b_host, device="cuda")
    t, its, rel, x = time_solves(dh, torch, h, cfg, b, args.steps, args.warmup, args.tol)
    out = {"workload": "poisson256", "n": A.n, "nnz": A.nnz, "levels": h.nlevels, "gpu_s_per_solve": t,
           "iterations": its[-1], "relres": rel}
    if not args.skip_cpu:
        from oracle import oracle as orc
        levels = [h.level(l, with_rmap=False) for l in range(h.nlevels)]
        c = h.coarsest()
        oh = orc.OracleHierarchy(levels, c["rp"], c["cv"], c["nz"], smoother="jacobi")
        t0 = time.perf_counter()
        xo, ito, _ = oh.pcg(np.zeros(A.n), b_host, args.tol, 500)
        out.update({"cpu_oracle_s_per_solve": time.perf_counter() - t0, "cpu_iterations": ito,
                    "rel_err_gpu_vs_cpu": float(np.linalg.norm(x.cpu().numpy() - xo) / np.linalg.norm(xo))})
    print(json.dumps(out))


def workload_resetup192(args, torch, dh):
    """3-D heterogeneous 7-point 192^3 (sigma 1.5): one setup, then 10 partial
    resetups with k <- k exp(N(0, 0.3^2)) assembled on the device; per resetup
    the FP64-, FP32- and BF16-cycle PCG solves to 1e-8 with their true
    residuals and errors against the FP64 solution."""
    from draugr_hip import adjoint as dadj
    P = dh.problems
    m = 192
    logk = P.log_permeability(m, 1.5, SEED + 3)
    A = P.heterogeneous3d(m, logk)
    cfg = dh.hypre_default_config(smoother=dh.L1JacobiSmootherType())
    t0 = time.perf_counter()
    h = dh.amg_setup(A, cfg)
    t_setup = time.perf_counter() - t0
    log(h.summary())
    b = torch.tensor(P.splitmix_normal(SEED + 13, A.n), device="cuda")
    lk = torch.tensor(logk, device="cuda")
    rows = []
    for r in range(10):
        eps = torch.tensor(0.3 * P.splitmix_normal(SEED + 100 + r, A.n), device="cuda")
        lk = lk + eps
        k = torch.exp(lk)
        t_res, _ = timed(torch, lambda: dh.amg_resetup_(h, dadj.assemble_7pt_(h, m, k), cfg))
        row = {"assemble_and_resetup_s": t_res}
        xs = {}
        for prec in ("fp64", "fp32", "bf16"):
            c = dh.hypre_default_config(smoother=dh.L1JacobiSmootherType(), precision=prec)
            x = torch.zeros_like(b)
            dh.pcg_solve_(x, b, h, c, tol=args.tol, maxiter=500)  # warm (tuning, mirrors, graphs)
            x.zero_()
            t, (_, it, rel) = timed(torch, lambda: dh.pcg_solve_(x, b, h, c, tol=args.tol, maxiter=500))
            xs[prec] = x
            row[prec] = {"s": t, "iterations": it, "relres": rel}
        nz = dadj.assemble_7pt_(h, m, k)
        import scipy.sparse as sp
        Ah = sp.csr_matrix((nz.cpu().numpy(), A.colval, A.rowptr), shape=(A.n, A.n))
        bh = b.cpu().numpy()
        x64 = xs["fp64"].cpu().numpy()
        for prec in ("fp64", "fp32", "bf16"):
            xp = xs[prec].cpu().numpy()
            row[prec]["true_relres"] = float(np.linalg.norm(bh - Ah @ xp) / np.linalg.norm(bh))
            row[prec]["rel_err_vs_fp64"] = float(np.linalg.norm(xp - x64) / np.linalg.norm(x64))
        log("resetup %d: %s" % (r, json.dumps(row)))
        rows.append(row)
    avg = lambda key: float(np.mean([r_[key]["s"] for r_ in rows]))
    print(json.dumps({"workload": "resetup192", "n": A.n, "nnz": A.nnz, "levels": h.nlevels, "setup_s": t_setup,
                      "assemble_and_resetup_s_mean": float(np.mean([r_["assemble_and_resetup_s"] for r_ in rows])),
                      "pcg_s_mean": {p_: avg(p_) for p_ in ("fp64", "fp32", "bf16")},
                      "speedup_fp32": avg("fp64") / avg("fp32"), "speedup_bf16": avg("fp64") / avg("bf16"),
                      "per_resetup": rows}))


def workload_batched48(args, torch, dh):
    """48 independent 3-D heterogeneous 7-point systems, edges uniform in
    {32..96} (sigma 1.0), 16 right-hand sides N(0,1) each; batched FP32-cycle
    (and FP64-cycle) PCG to 1e-8, one graph replay per cycle."""
    P = dh.problems
    u = P.splitmix_uniform(SEED + 4, 48)
    edges = [32 + int(v * 65) for v in u]
    mats = [P.heterogeneous3d(m, P.log_permeability(m, 1.0, SEED + 1000 + s)) for s, m in enumerate(edges)]
    K = 16
    cfg = dh.hypre_default_config(smoother=dh.L1JacobiSmootherType())
    t0 = time.perf_counter()
    batch = dh.amg_batch_setup(mats, cfg, nrhs=K)
    t_setup = time.perf_counter() - t0
    info = batch.info()
    log("batch: %d systems, %d rows, %d rhs, deepest %d levels, setup %.1fs, resident %.1f GB"
        % (info["nsys"], info["rows"], K, info["max_levels"], t_setup, info["resident_bytes"] / 1e9))
    if args.kernel_only:  # the calibration copy and the dominant pass only (for counter runs)
        li = batch.level_info(0)
        xa = torch.randn(li["n_vec"] * K, dtype=torch.float32, device="cuda")
        ba = torch.randn(li["n_vec"] * K, dtype=torch.float32, device="cuda")
        oa = torch.empty_like(xa)
        batch.jacobi_pass(0, "fp32", xa, ba, oa)  # tunes the FP32 mappings
        torch.cuda.synchronize()
        calibration_copy(torch)
        t_k = event_time(torch, lambda: batch.jacobi_pass(0, "fp32", xa, ba, oa), 10)
        print(json.dumps({"kernel": "batched48_level0_jacobi_fp32", "avg_launch_us": t_k * 1e6,
                          "mapping": batch.tune_report().splitlines()[0] if batch.tune_report() else ""}))
        return
    b = torch.tensor(P.splitmix_normal(SEED + 14, info["rows"] * K), device="cuda")
    out = {"workload": "batched48", "systems": 48, "rhs": K, "rows": info["rows"],
           "edges": edges, "setup_s": t_setup}
    for prec in ("fp64", "fp32"):
        c = dh.hypre_default_config(smoother=dh.L1JacobiSmootherType(), precision=prec)
        x = torch.zeros_like(b)
        batch.pcg_solve_(x, b, c, tol=args.tol, maxiter=500)
        ts = []
        for _ in range(max(1, args.steps // 3)):
            x.zero_()
            t, (_, its, rel) = timed(torch, lambda: batch.pcg_solve_(x, b, c, tol=args.tol, maxiter=500))
            ts.append(t)
        t = float(np.median(ts))
        out[prec] = {"s_per_batched_solve": t, "solves_per_s": 48 * K / t, "max_iterations": int(its.max()),
                     "mean_iterations": float(its.mean()), "max_relres": float(rel.max())}
        if prec == "fp32":
            xs = batch.unpack(x)
            bs = batch.unpack(b)
            worst = 0.0
            for s_ in range(0, 48, 8):
                for j in range(0, K, 5):
                    r = bs[s_][:, j] - mats[s_].matvec(xs[s_][:, j])
                    worst = max(worst, float(np.linalg.norm(r) / np.linalg.norm(bs[s_][:, j])))
            out[prec]["true_relres_worst_of_sample"] = worst
    out["speedup_fp32"] = out["fp64"]["s_per_batched_solve"] / out["fp32"]["s_per_batched_solve"]
    log(batch.tune_report())
    # dominant pass: the K-wide Jacobi step on packed level 0 (all 48 systems),
    # FP32, timed back to back with events on the stream it runs on
    li = batch.level_info(0)
    n0, nnz0 = li["n_smooth"], li["nnz"]
    xa = torch.randn(li["n_vec"] * K, dtype=torch.float32, device="cuda")
    ba = torch.randn(li["n_vec"] * K, dtype=torch.float32, device="cuda")
    oa = torch.empty_like(xa)
    t_k = min(event_time(torch, lambda: batch.jacobi_pass(0, "fp32", xa, ba, oa), 20) for _ in range(3))
    kbytes = nnz0 * (4 + 4) + (n0 + 1) * 4 + n0 * 4 + 3 * K * n0 * 4  # A, d, then x, b, out (K wide)
    out["roofline"] = {"bound": "hbm", "achieved": kbytes / t_k / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": kbytes / t_k / 1e9 / HBM_PEAK_GBS,
                       "traffic": measured_traffic(os.path.join(ROOT, "profiles", "batched48_kernel_stats.md"),
                                                   "batched48_level0_jacobi_fp32"),
                       "kernel": "K-wide l1-Jacobi step on packed level 0 (%d rows x %d rhs, %d nnz), FP32"
                                 % (n0, K, nnz0),
                       "bytes_per_launch": kbytes, "avg_launch_us": t_k * 1e6}
    print(json.dumps(out))


def workload_adjoint96(args, torch, dh):
    """Loss 1/2 |x - x*|^2, A(k) x = b on 96^3 (sigma 1.5): time of one
    forward (assembly, resetup, PCG) plus backward (adjoint PCG, gradient
    kernels) evaluation, FP64 and FP32 cycle."""
    from draugr_hip import adjoint as dadj
    P = dh.problems
    m = 96
    logk = P.log_permeability(m, 1.5, SEED + 5)
    cfg = dh.hypre_default_config(smoother=dh.L1JacobiSmootherType())
    h = dh.amg_setup(P.heterogeneous3d(m, logk), cfg)
    b = torch.tensor(P.splitmix_normal(SEED + 105, m ** 3), device="cuda")
    xs = torch.tensor(P.splitmix_normal(SEED + 205, m ** 3), device="cuda")
    out = {"workload": "adjoint96", "n": m ** 3}
    for prec in ("fp64", "fp32"):
        c = dh.hypre_default_config(smoother=dh.L1JacobiSmootherType(), precision=prec)

        def evaluate():
            lk = torch.tensor(logk, device="cuda").requires_grad_(True)
            x = dadj.solve_7pt(lk, b, h, m, c, tol=1e-10)
            (0.5 * ((x - xs) ** 2).sum()).backward()
            return lk.grad

        evaluate()
        ts = [timed(torch, evaluate)[0] for _ in range(3)]
        out[prec] = {"s_per_loss_and_gradient": float(np.median(ts))}
    print(json.dumps(out))


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the design for sharding is in "
              "the documentation); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    import draugr_hip as dh
    other = {"poisson256": workload_poisson256, "resetup192": workload_resetup192,
             "batched48": workload_batched48, "adjoint96": workload_adjoint96}
    if args.workload in other:
        other[args.workload](args, torch, dh)
        return

    P = dh.problems
    m = args.size
    t0 = time.perf_counter()
    logk = P.log_permeability(m, args.sigma, SEED + 1)
    A = P.heterogeneous3d(m, logk)
    b_host = P.splitmix_normal(SEED + 2, A.n)
    log("problem %d^3: n=%d nnz=%d (%.1fs)" % (m, A.n, A.nnz, time.perf_counter() - t0))
    cfg64 = dh.hypre_default_config(smoother=dh.L1JacobiSmootherType(), kernel=args.kernel, fuse=args.fuse)
    cfg32 = dh.hypre_default_config(smoother=dh.L1JacobiSmootherType(), precision=args.precision, kernel=args.kernel,
                                    fuse=args.fuse)
    t0 = time.perf_counter()
    h = dh.amg_setup(A, cfg64)
    t_setup = time.perf_counter() - t0
    log(h.summary())
    log("setup + upload %.2fs, resident %.1f MB" % (t_setup, h.resident_bytes() / 1e6))

    b = torch.tensor(b_host, device="cuda")
    if args.kernel_only:
        z = torch.zeros_like(b)
        dh.amg_cycle_(z, b, h, cfg32, zero_guess=True)  # the autotuner picks the pass kernels
        torch.cuda.synchronize()
        calibration_copy(torch)
        t_k, kbytes, kstored, tag = dominant_kernel(torch, h, args.precision, 20)
        print(json.dumps({"kernel": tag, "bytes_per_launch": kbytes, "bytes_as_stored": kstored,
                          "avg_launch_us": t_k * 1e6, "achieved_GBps": kbytes / t_k / 1e9}))
        return
    t64, its64, rel64, x64 = time_solves(dh, torch, h, cfg64, b, args.steps, args.warmup, args.tol)
    log("FP64 cycle: %.4f s/solve, iterations %s, relres %.2e" % (t64, its64, rel64))
    t32, its32, rel32, x32 = time_solves(dh, torch, h, cfg32, b, args.steps, args.warmup, args.tol)
    log("%s cycle: %.4f s/solve, iterations %s, relres %.2e" % (args.precision, t32, its32, rel32))
    other = "bf16" if args.precision == "fp32" else "fp32"
    cfg_o = dh.hypre_default_config(smoother=dh.L1JacobiSmootherType(), precision=other, kernel=args.kernel,
                                    fuse=args.fuse)
    t_o, its_o, rel_o, x_o = time_solves(dh, torch, h, cfg_o, b, args.steps, args.warmup, args.tol)
    log("%s cycle: %.4f s/solve, iterations %s, relres %.2e" % (other, t_o, its_o, rel_o))
    log(h.tune_report())
    xr = x64.cpu().numpy()
    x3 = x32.cpu().numpy()
    xo3 = x_o.cpu().numpy()
    true = lambda v: float(np.linalg.norm(b_host - A.matvec(v)) / np.linalg.norm(b_host))  # noqa: E731
    relerr = lambda u, v: float(np.linalg.norm(u - v) / np.linalg.norm(v))  # noqa: E731
    true32 = true(x3)
    err32 = relerr(x3, xr)

    # V-cycle time and bandwidth (zero-guess preconditioner application)
    z = torch.zeros_like(b)
    # best of three runs of 20 back-to-back cycles (the first run after host
    # work can see the GPU still clocking up)
    cyc = {}
    for name, c in (("fp64", cfg64), (args.precision, cfg32), (other, cfg_o)):
        # the cycle as the PCG applies it (launched as the config says: eager
        # by default, graph replays with graphs=true) on its own vectors, 20
        # back to back from one call
        h.precondition_repeat(b, c, 2)
        t_c = min(event_time(torch, lambda: h.precondition_repeat(b, c, 20), 1) / 20 for _ in range(3))
        # the same cycle through the C-ABI call on caller vectors (adds the
        # copies / precision conversions of x and b around it)
        t_api = min(event_time(torch, lambda: dh.amg_cycle_(z, b, h, c, zero_guess=True), 20) for _ in range(3))
        nb = h.cycle_bytes(c)
        cyc[name] = {"ms": t_c * 1e3, "bytes": nb, "GBps": nb / t_c / 1e9,
                     "frac_of_hbm_peak": nb / t_c / 1e9 / HBM_PEAK_GBS, "api_call_ms": t_api * 1e3}
    t_cyc32, t_cyc64 = cyc[args.precision]["ms"] / 1e3, cyc["fp64"]["ms"] / 1e3
    bytes32, bytes64 = cyc[args.precision]["bytes"], cyc["fp64"]["bytes"]

    # dominant kernel, timed live
    t_k, kbytes, kstored, tag = dominant_kernel(torch, h, args.precision, 50)
    ach = kbytes / t_k / 1e9
    traffic = measured_traffic(args.traffic, tag)

    out = {
        "metric": METRIC,
        "value": t32,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t32 * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 cycle / f64 PCG" if args.precision == "fp32" else "bf16-value cycle / f64 PCG",
        "data": "synthetic",
        "config": {
            "workload": "3D 7-point heterogeneous diffusion %d^3 (n=%d, nnz=%d), log10 k ~ N(0,%.1f^2), "
                        "hypre_default_config + l1-Jacobi V(1,1), PCG to %.0e" % (m, A.n, A.nnz, args.sigma,
                                                                                args.tol),
            "n": A.n, "nnz": A.nnz, "levels": h.nlevels, "seed": SEED,
        },
        "fp64_cycle_s_per_solve": t64,
        "speedup_fp32_cycle_vs_fp64": t64 / t32,
        "%s_cycle_s_per_solve" % other: t_o,
        "speedup_%s_cycle_vs_fp64" % other: t64 / t_o,
        "iters_fp64": its64[-1],
        "iters_mixed": its32[-1],
        "iters_%s" % other: its_o[-1],
        "mixed_true_relres": true32,
        "mixed_rel_err_vs_fp64_solution": err32,
        "%s_true_relres" % other: true(xo3),
        "fp64_true_relres": true(xr),
        "vcycle_ms": {"mixed": t_cyc32 * 1e3, "fp64": t_cyc64 * 1e3},
        "vcycle_GBps_algorithmic": {"mixed": bytes32 / t_cyc32 / 1e9, "fp64": bytes64 / t_cyc64 / 1e9},
        "vcycle_bytes": {"mixed": bytes32, "fp64": bytes64},
        "vcycle_by_precision": cyc,
        "setup_s": t_setup,
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "%s: %s EpiJacobi, l1-Jacobi step on level 1 (%d rows, %d nnz)"
                               % (tag, KERNEL_OF_KIND.get(tag.split("_")[3], "?"), h.level_info(1)["n"],
                                  h.level_info(1)["nnz"]),
                     "bytes_per_launch": kbytes, "bytes_as_stored": kstored, "avg_launch_us": t_k * 1e6},
    }

    if not args.skip_cpu:
        # The oracle (OpenMP C restatement of Draugr's CPU cycle and a textbook
        # FP64 PCG) on the same hierarchy and right-hand side, after the GPU
        # timing: its solution anchors the GPU solutions, and one warm-up plus
        # the median of full solves is the CPU baseline.
        from oracle import oracle as orc
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        levels = [h.level(l, with_rmap=False) for l in range(h.nlevels)]
        c = h.coarsest()
        oh = orc.OracleHierarchy(levels, c["rp"], c["cv"], c["nz"], smoother="l1_jacobi")
        ts = []
        for rep in range(1 + args.cpu_solves):
            t0 = time.perf_counter()
            xo, itc, _ = oh.pcg(np.zeros(A.n), b_host, args.tol, 500)
            if rep:
                ts.append(time.perf_counter() - t0)
        log("CPU oracle: %d iterations, %s s" % (itc, ["%.3f" % t for t in ts]))
        out["cpu_oracle_iterations"] = itc
        out["rel_err_vs_cpu_oracle"] = {"fp64": relerr(xr, xo), args.precision: relerr(x3, xo), other: relerr(xo3, xo)}
        out["cpu_baseline"] = {"value": float(np.median(ts)), "unit": "s", "cores": threads, "kind": "port",
                               "cpu": cpu_model(),
                               "sample": "1 warm-up + median of %d full FP64 PCG solves to %.0e (%d iterations; "
                                         "oracle OpenMP restatement of the CPU cycle, same hierarchy and b); "
                                         "min %.3f s, max %.3f s" % (args.cpu_solves, args.tol, itc, min(ts),
                                                                     max(ts))}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
