"""Benchmark: seconds to relative residual 1e-10 and HBM GB/s for FP64 PCG on
the 3D 7-point graph Laplacian 256^3 + 1e-8 I (n = 16,777,216, nnz =
117,047,296), b ~ N(0,1) seeded and mean-subtracted, Jacobi preconditioner.

One step = one complete PCG solve (x0 = 0 to ||r|| < 1e-10 ||r0||) with the
matrix, preconditioner and right-hand side already resident in HBM.  Prints
one JSON line.  Single GPU only.

    python bench.py [--steps K] [--warmup W] [--size 256] [--prec line_jacobi]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def spmv_bytes(n: int, nnz: int) -> int:
    """Algorithmic bytes of one FP64 CSR SpMV with beta = 0:
    12 B per nonzero (int32 column + FP64 value), 4 B per row pointer,
    8 B to read x and 8 B to write y per row."""
    return 12 * nnz + 4 * (n + 1) + 16 * n


def jacobi_iteration_bytes(n: int, nnz: int) -> int:
    """Bytes of one Jacobi PCG iteration: the SpMV + 80 B per row (K_B reads
    r, q, 1/diag and writes r; K_C reads x, p, r, 1/diag and writes x, p)."""
    return spmv_bytes(n, nnz) + 80 * n


def measured_traffic(size: int):
    """HBM bytes per launch of the fused SpMV on size^3 as measured by
    rocprofv3 FETCH_SIZE + WRITE_SIZE (each scaled by the counter's reading on
    a 1 GiB copy), from the committed profile summary; None when absent."""
    path = os.path.join(ROOT, "profiles", f"k_spmv_pq_{size}cubed_traffic.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":", 1)[1])
    return None


def host_cores() -> int:
    """CPUs this process may run on (its affinity mask), not the machine's."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:  # not Linux
        return os.cpu_count() or 1


def host_cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(Ah, b, prec: str, tol: float, gpu_iters: int, sample_iters: int):
    """The reference's OpenMP cg() (compiled from its sources into
    oracle/_ref/) timed on this host for `sample_iters` iterations of the same
    system, inside the bracket cg_driver2.c:392-408 puts around cg(); seconds
    to the GPU's iteration count are extrapolated from the per-iteration time
    (a full solve at ~60 ms per iteration would take ~95 s).  The build is the
    reference's as written: SpMV on its hard-coded 16 threads
    (openmp_blas.c:37), dot on OMP_NUM_THREADS, other vector ops serial, and
    the SpMV's shared loop variable `col` raced on as in openmp_blas.c:35-51.
    Beside it, the all-cores figure: the oracle's OpenMP restatement with the
    SpMV and dot on every core this process may use.  Falls back to the
    oracle for the first figure when the reference build is absent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    cores = host_cores()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or cores
    maxit = max(sample_iters - 1, 1)
    orc = oracle.Oracle()
    if os.path.exists(oracle.REF_OMP):
        ref = oracle.Reference(oracle.REF_OMP)
        pdk = ref.make_pdata(Ah, prec)
        t0 = time.perf_counter()
        _, it, _, _ = ref.cg(Ah, b, prec, tol=tol, maxit=maxit, pdk=pdk)
        dt = time.perf_counter() - t0
        kind = "reference"
        what = ("the reference's OpenMP cg() (openmp_blas.c build: SpMV on its hard-coded 16 "
                f"threads with the racy shared `col`, dot on OMP_NUM_THREADS={threads}, other "
                "vector ops serial)")
    else:
        t0 = time.perf_counter()
        _, it, _, _ = orc.cg(Ah, b, prec, tol=tol, maxit=maxit, omp=(16, threads))
        dt = time.perf_counter() - t0
        kind = "port"
        what = "the oracle's OpenMP restatement of the reference CPU path"
    per_iter = dt / it
    prec_h = orc.make_prec(Ah, prec)
    t0 = time.perf_counter()
    _, it_all, _, _ = orc.cg(Ah, b, prec, tol=tol, maxit=maxit, prec=prec_h, omp=(cores, cores))
    dt_all = time.perf_counter() - t0
    per_all = dt_all / it_all
    return {
        "value": per_iter * gpu_iters,
        "unit": "s",
        "cores": min(cores, max(16, threads)),
        "kind": kind,
        "sample": (f"{it} PCG iterations of {what} on the same 256^3 system "
                   f"({dt:.2f} s, {per_iter * 1e3:.1f} ms/iteration), scaled to the "
                   f"GPU's {gpu_iters} iterations"),
        "host_cores": cores,
        "host_cpu": host_cpu_model(),
        "all_cores": {
            "value": per_all * gpu_iters, "unit": "s", "cores": cores, "kind": "port",
            "sample": (f"{it_all} PCG iterations of the oracle's OpenMP restatement with SpMV "
                       f"and dot on all {cores} cores ({dt_all:.2f} s, "
                       f"{per_all * 1e3:.1f} ms/iteration), scaled to {gpu_iters} iterations"),
        },
    }


def bench_c1(S, torch, np, reps: int = 3) -> dict:
    """The reference's own CPU-runnable case: 2D 5-point Laplacian 128x128 +
    1e-8*I, b ~ U(-1,1) seed 12345 mean-subtracted, Jacobi PCG to rel-res 1e-8.
    The GPU solve (median of `reps`, inputs resident) beside the reference's
    serial simple_blas cg() on one core, timed around cg() as
    cg_driver2.c:392-408 does (median of `reps`; the oracle's serial
    restatement when the reference build is absent)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from lap_amd.problems import grid_laplacian, rhs_uniform

    Ah = grid_laplacian([128, 128], 1e-8)
    b_h = rhs_uniform(Ah.n, 12345)
    A = S.DeviceCSR.from_csr(Ah)
    b = S.to_dev(b_h)
    P = S.make_preconditioner(A, Ah, "line_jacobi")
    S.pcg(A, b, P, tol=1e-8, maxit=25000)
    walls, r = [], None
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = S.pcg(A, b, P, tol=1e-8, maxit=25000)
        torch.cuda.synchronize()
        walls.append(time.perf_counter() - t0)
    cpu, it_c = [], None
    if oracle.have_ref():
        ref = oracle.Reference(oracle.REF_CPU)
        pdk = ref.make_pdata(Ah, "line_jacobi")
        kind = "reference"
        for _ in range(reps):
            t0 = time.perf_counter()
            _, it_c, _, _ = ref.cg(Ah, b_h, "line_jacobi", tol=1e-8, maxit=25000, pdk=pdk)
            cpu.append(time.perf_counter() - t0)
    else:
        orc = oracle.Oracle()
        prec = orc.make_prec(Ah, "line_jacobi")
        kind = "port"
        for _ in range(reps):
            t0 = time.perf_counter()
            _, it_c, _, _ = orc.cg(Ah, b_h, "line_jacobi", tol=1e-8, maxit=25000, prec=prec)
            cpu.append(time.perf_counter() - t0)
    gpu_s, cpu_s = float(np.median(walls)), float(np.median(cpu))
    return {
        "workload": "2D 5-point graph Laplacian 128x128 + 1e-8*I, b~U(-1,1) seed 12345 "
                    "mean-subtracted, Jacobi PCG to rel-res 1e-8",
        "gpu": {"value": gpu_s, "unit": "s", "iterations": r.iterations, "flag": r.flag,
                "device_ms": r.gpu_ms},
        "cpu_serial": {"value": cpu_s, "unit": "s", "iterations": it_c, "cores": 1,
                       "kind": kind, "sample": f"full solve, median of {reps}"},
        "speedup": cpu_s / gpu_s,
    }


def block_iteration_bytes(n: int, nnz: int, k: int) -> int:
    """Block CG with k right-hand sides: one SpMM (matrix read once) plus
    104 B per row and column of block-vector traffic."""
    return 12 * nnz + 4 * (n + 1) + 104 * k * n


def bench_preconditioners(S, torch, A, Ah, b, tol: float) -> dict:
    """The same 256^3 system with the IC0 and GS_std preconditioners (M = K =
    1): setup (host analysis + device factorisation) and one timed solve."""
    out = {}
    for kind in ("ichol", "GS_std"):
        t0 = time.perf_counter()
        P = S.make_preconditioner(A, Ah, kind)
        torch.cuda.synchronize()
        setup = time.perf_counter() - t0
        t0 = time.perf_counter()
        r = S.pcg(A, b, P, tol=tol, maxit=25000)
        wall = time.perf_counter() - t0
        out[kind] = {"value": wall, "unit": "s", "iterations": r.iterations, "flag": r.flag,
                     "device_ms": r.gpu_ms, "setup_s": setup,
                     "ms_per_iteration": r.gpu_ms / max(r.iterations, 1)}
        if kind == "ichol":  # FP32 inner PCG with the FP32 factor, FP64 refinement
            S.pcg_mixed(A, b, P, tol=tol, maxit=25000)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            m = S.pcg_mixed(A, b, P, tol=tol, maxit=25000)
            torch.cuda.synchronize()
            mw = time.perf_counter() - t0
            out[kind]["mixed_precision"] = {
                "value": mw, "unit": "s", "iterations": m.iterations,
                "refinements": m.refinements, "flag": m.flag,
                "true_rel_res": m.true_res / m.res0, "speedup_vs_fp64": wall / mw,
                "scheme": "FP32 inner vectors, FP32 IC0 factor (chunked FP32 solves), FP32 "
                          "off-diagonals; the near-null constant mode deflated"}
        del P
    return out


def bench_block(S, torch, np, side: int = 128, k: int = 64, tol: float = 1e-10) -> dict:
    """k-RHS block PCG (line_jacobi) on side^3 + 1e-8 I against k separate
    single-RHS solves of the same columns."""
    from lap_amd.problems import diagonal, grid_laplacian
    Ah = grid_laplacian([side] * 3, 1e-8)
    A = S.DeviceCSR.from_csr(Ah)
    d_r = S.to_dev(1.0 / diagonal(Ah))
    Bh = np.random.default_rng(12345).standard_normal((Ah.n, k))
    Bh -= Bh.mean(axis=0)
    B = S.to_dev(Bh)
    S.pcg_block(A, B, d_r, tol=tol)  # warm-up, picks the SpMM tile
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = S.pcg_block(A, B, d_r, tol=tol)
    wall = time.perf_counter() - t0
    P = S.make_preconditioner(A, Ah, "line_jacobi")
    cols = [B[:, j].contiguous() for j in range(k)]
    S.pcg(A, cols[0], P, tol=tol)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    single_it = 0
    for c in cols:
        single_it += S.pcg(A, c, P, tol=tol).iterations
    seq = time.perf_counter() - t0
    its = int(res.iterations.max())
    return {
        "workload": f"{side}^3 7-point Laplacian + 1e-8*I, {k} right-hand sides N(0,1) "
                    f"mean-subtracted, line_jacobi, tol {tol:g}",
        "value": wall, "unit": "s", "device_ms": res.gpu_ms,
        "iterations_min_max": [int(res.iterations.min()), its],
        "flags": sorted(set(res.flags.tolist())),
        "sequential_single_rhs_s": seq, "sequential_iterations": single_it,
        "speedup_vs_sequential": seq / wall,
        "iteration_gbs": block_iteration_bytes(A.n, A.nnz, k) * its / (res.gpu_ms * 1e-3) / 1e9,
    }


def bench_batched(S, torch, np, count: int = 256, tol: float = 1e-10) -> dict:
    """count 2D grid Laplacians (side ~ U{32..448}) + 1e-8 I in one persistent
    launch, autotuned, against sequential single-system solves of 16 of them
    (scaled to count)."""
    from lap_amd.problems import grid_laplacian, rhs_normal
    rng = np.random.default_rng(12345)
    sides = rng.integers(32, 449, count)
    systems = [grid_laplacian([int(s), int(s)], 1e-8) for s in sides]
    rhs = [rhs_normal(Ah.n, 1000 + i) for i, Ah in enumerate(systems)]
    Bt = S.Batch.from_systems(systems, rhs)
    cfg = S.batched_tune(Bt, 20)
    S.pcg_batched(Bt, tol=tol)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x, it, fl, _, _, ms = S.pcg_batched(Bt, tol=tol)
    wall = time.perf_counter() - t0
    sel = list(range(0, count, max(1, count // 16)))
    seq_ms = 0.0
    for i in sel:
        A = S.DeviceCSR.from_csr(systems[i])
        P = S.make_preconditioner(A, systems[i], "line_jacobi")
        b = S.to_dev(rhs[i])
        S.pcg(A, b, P, tol=tol)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        S.pcg(A, b, P, tol=tol)
        seq_ms += (time.perf_counter() - t0) * 1e3
    seq_est = seq_ms * count / len(sel) / 1e3
    return {
        "workload": f"{count} 2D 5-point grids, side U{{32..448}} seed 12345, + 1e-8*I, "
                    f"b~N(0,1) mean-subtracted, line_jacobi, tol {tol:g}",
        "value": wall, "unit": "s", "device_ms": ms, "rows": int(Bt.A.n),
        "config": {"block": cfg[0], "threads_per_row": cfg[1]},
        "iterations_min_max": [int(it.min()), int(it.max())],
        "flags": sorted(set(fl.tolist())),
        "sequential_single_solves_s_est": seq_est,
        "sequential_sample": f"{len(sel)} of {count} systems solved one by one, scaled",
        "speedup_vs_sequential": seq_est / wall,
    }


def bench_sweeps(S, torch, A, Ah, b, tol: float) -> dict:
    """The 256^3 system with the SpMV-only preconditioners at M/K values of
    the reference's experiment table (script_hip/run_power_cg_hip.sh):
    it_jacobi K=2, GS_it2 M=1 K=1 and M=1 K=3, each solved with one launch
    per sweep (fuse_sweeps=1, the default) and with the reference's launch
    sequence (fuse_sweeps=0); both give the same iterates bit for bit."""
    out = {}
    for kind, m, k in (("it_jacobi", 1, 2), ("GS_it2", 1, 1), ("GS_it2", 1, 3)):
        P = S.make_preconditioner(A, Ah, kind, m=m, k=k)
        res = {}
        try:
            for fuse in (0, 1):
                S.set_option("fuse_sweeps", fuse)
                S.pcg(A, b, P, tol=tol, maxit=25000)  # warm (graph, plans)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                r = S.pcg(A, b, P, tol=tol, maxit=25000)
                res[fuse] = (time.perf_counter() - t0, r)
        finally:
            S.set_option("fuse_sweeps", 1)
        (w0, r0), (w1, r1) = res[0], res[1]
        out[f"{kind}_M{m}_K{k}"] = {
            "value": w1, "unit": "s", "iterations": r1.iterations, "flag": r1.flag,
            "ms_per_iteration": r1.gpu_ms / max(r1.iterations, 1),
            "unfused_s": w0, "unfused_ms_per_iteration": r0.gpu_ms / max(r0.iterations, 1),
            "speedup_vs_unfused": w0 / w1,
            "bitwise_equal_to_unfused": bool(torch.equal(r0.x, r1.x)
                                             and r0.iterations == r1.iterations),
        }
        del P
    return out


def bench_energy(S, torch, A, b, P, tol: float, solves: int = 3) -> dict:
    """Energy to solution of the FP64 and mixed-precision Jacobi solves on the
    256^3 system: the GPU's energy accumulator read around `solves`
    back-to-back solves, the socket power sampled every 20 ms beside it (the
    reference's time x mean-power figure), and the idle power over 2 s after 3 s
    of settling (right after load the board's power decays for a second or
    more, and the sampled power, an average over the firmware's own window,
    lags the accumulator there)."""
    from lap_amd.energy import EnergyMeter
    out = {}
    torch.cuda.synchronize()
    time.sleep(3.0)
    with EnergyMeter() as idle:
        time.sleep(2.0)
    if not idle.available:
        return idle.report()
    out["idle"] = idle.report()
    for name, fn in (("fp64", lambda: S.pcg(A, b, P, tol=tol, maxit=25000)),
                     ("mixed", lambda: S.pcg_mixed(A, b, P, tol=tol, maxit=25000))):
        fn()
        torch.cuda.synchronize()
        with EnergyMeter() as em:
            for _ in range(solves):
                fn()
            torch.cuda.synchronize()
        out[name] = em.report(solves)
    if "fp64" in out and "mixed" in out and "energy_j_per_solve" in out["mixed"]:
        f, m = out["fp64"], out["mixed"]
        out["mixed_energy_saving"] = f["energy_j_per_solve"] / m["energy_j_per_solve"]
        if "energy_j_per_solve_time_x_power" in f and "energy_j_per_solve_time_x_power" in m:
            out["mixed_energy_saving_time_x_power"] = (f["energy_j_per_solve_time_x_power"]
                                                      / m["energy_j_per_solve_time_x_power"])
            # the counter-based saving stands only when both readings agree
            # with their sampled power
            out["counter_readings_consistent"] = bool(f.get("sources_agree")
                                                      and m.get("sources_agree"))
    return out


def bench_power_law(S, torch, np, tol: float, n: int = 1 << 22, reps: int = 2,
                    kinds=("line_jacobi", "ichol"), maxit: int = 25000) -> dict:
    """The mixed-precision workload on a power-law graph: weighted Chung-Lu
    Laplacian, n = 2^22, degree exponent 2.1, mean degree 16, degrees capped
    at 20000, lognormal(0, 1) weights, diagonal (1 + 1e-6) x the weight sum;
    b ~ N(0,1) (seed 12345) mean-subtracted.  FP64 PCG against the
    mixed-precision solve (FP32 matrix values with row sums kept, FP64
    refinement to `tol`) with FP64 inner vectors (mixed_vectors 1) and, for
    Jacobi, FP32 inner vectors (mixed_vectors 0); Jacobi and IC0.  Also the
    fused SpMV's rate on this matrix (gather-bound: its columns are random
    beyond the hub prefix).  maxit below the solves' iteration counts
    (profiling runs) stops every solve there: the times are then not solve
    times, and the record says so."""
    from lap_amd.problems import chung_lu, rhs_normal

    t0 = time.perf_counter()
    Ah = chung_lu(n)
    out = {"workload": ("weighted Chung-Lu Laplacian n=2^22, gamma 2.1, mean degree 16, cap "
                        "20000, lognormal(0,1) weights, diag (1+1e-6)*sum w; b~N(0,1) seed "
                        f"12345 mean-subtracted; PCG to rel-res {tol:g}"),
           "n": n, "nnz": Ah.nnz, "max_row": int(np.diff(Ah.ia).max()),
           "build_s": time.perf_counter() - t0}
    if maxit < 25000:
        out["maxit_capped"] = maxit
    A = S.DeviceCSR.from_csr(Ah)
    b = S.to_dev(rhs_normal(n, 12345))
    x = torch.randn(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    ms = S.spmv_timed(A, x, y, 20, fused=True)
    gbs = spmv_bytes(n, Ah.nnz) / (ms * 1e-3) / 1e9
    out["spmv"] = {"kernel": "k_spmv_pq", "avg_ms": ms, "achieved_gbs": gbs,
                   "frac": gbs / HBM_PEAK_GBS}
    del x, y

    def timed(fn, count):
        res, t = [], time.perf_counter()
        for _ in range(count):
            res.append(fn())
        torch.cuda.synchronize()
        return res, (time.perf_counter() - t) / count

    prev = S.set_option("mixed_vectors", 0)
    try:
        for kind in kinds:
            count = reps if kind == "line_jacobi" else 1
            t = time.perf_counter()
            P = S.make_preconditioner(A, Ah, kind)
            torch.cuda.synchronize()
            rec = {"setup_s": time.perf_counter() - t}
            fp64 = lambda: S.pcg(A, b, P, tol=tol, maxit=maxit)
            if count > 1:
                fp64()
            res, wall = timed(fp64, count)
            d = res[-1]
            rec["fp64"] = {"s": wall, "iterations": d.iterations, "flag": d.flag,
                           "device_ms": d.gpu_ms, "true_rel_res": d.true_res / d.res0}
            log(f"power_law {kind} fp64: {d.iterations} iterations {wall:.3f} s")
            for mv in ((1, 0) if kind == "line_jacobi" else (1,)):
                S.set_option("mixed_vectors", mv)
                mixed = lambda: S.pcg_mixed(A, b, P, tol=tol, maxit=maxit)
                if count > 1:
                    mixed()
                res, mwall = timed(mixed, count)
                m = res[-1]
                err = float(torch.linalg.norm(m.x - d.x) / torch.linalg.norm(d.x))
                rec[f"mixed_vectors_{mv}"] = {
                    "s": mwall, "iterations": m.iterations, "refinements": m.refinements,
                    "flag": m.flag, "device_ms": m.gpu_ms, "true_rel_res": m.true_res / m.res0,
                    "speedup_vs_fp64": wall / mwall, "rel_err_vs_fp64_solution": err}
                log(f"power_law {kind} mixed_vectors {mv}: {m.iterations} iterations "
                    f"{mwall:.3f} s ({wall / mwall:.2f}x)")
            S.set_option("mixed_vectors", 0)
            if kind == "ichol":
                # long rows summed as workgroup trees (sptrsv_long 1): no serial
                # chains, tolerance parity instead of bitwise
                prev_long = S.set_option("sptrsv_long", 1)
                try:
                    res, twall = timed(fp64, 1)
                    t64 = res[-1]
                    S.set_option("mixed_vectors", 1)
                    res, tmwall = timed(lambda: S.pcg_mixed(A, b, P, tol=tol, maxit=maxit), 1)
                    tm = res[-1]
                finally:
                    S.set_option("sptrsv_long", prev_long)
                    S.set_option("mixed_vectors", 0)
                rec["tree_long_rows"] = {
                    "fp64": {"s": twall, "iterations": t64.iterations, "flag": t64.flag,
                             "true_rel_res": t64.true_res / t64.res0,
                             "rel_diff_vs_sequential_solution": float(
                                 torch.linalg.norm(t64.x - d.x) / torch.linalg.norm(d.x))},
                    "mixed_vectors_1": {"s": tmwall, "iterations": tm.iterations,
                                        "refinements": tm.refinements, "flag": tm.flag,
                                        "speedup_vs_fp64_tree": twall / tmwall},
                    "speedup_vs_sequential_fp64": wall / twall}
                log(f"power_law ichol tree rows: fp64 {t64.iterations} iterations {twall:.3f} s, "
                    f"mixed {tm.iterations} iterations {tmwall:.3f} s")
            out[kind] = rec
            del P
    finally:
        S.set_option("mixed_vectors", prev)
    A.forget()
    return out


def bench_value_indexed(S, torch, A, b, P, tol: float, ref, reps: int) -> dict:
    """The headline solve with the SpMV reading value-indexed CSR (option
    spmv_format 1: one-byte indices into the matrix's 5 distinct doubles,
    5 B per nonzero instead of 12).  Iterates are bitwise the CSR solve's.
    Its SpMV kernel is timed live; its roofline uses its own algorithmic
    bytes, 5*nnz + 4(n+1) + 16n.  The mixed-precision solve with its FP32
    off-diagonal copy value-indexed is timed too, against both FP64 solves."""
    try:
        S.set_option("spmv_format", 1)
        S.pcg(A, b, P, tol=tol, maxit=25000)  # warm
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = S.pcg(A, b, P, tol=tol, maxit=25000)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        x = torch.randn(A.n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        ms = S.spmv_timed(A, x, y, reps, fused=True)
        S.pcg_mixed(A, b, P, tol=tol, maxit=25000)  # warm
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        m = S.pcg_mixed(A, b, P, tol=tol, maxit=25000)
        torch.cuda.synchronize()
        mwall = time.perf_counter() - t0
    finally:
        S.set_option("spmv_format", 0)
    vib = 5 * A.nnz + 4 * (A.n + 1) + 16 * A.n
    ach = vib / (ms * 1e-3) / 1e9
    return {
        "value": wall, "unit": "s", "iterations": r.iterations, "flag": r.flag,
        "device_ms_per_solve": r.gpu_ms,
        "ms_per_iteration": r.gpu_ms / max(r.iterations, 1),
        "bitwise_equal_to_csr_solve": bool(torch.equal(r.x, ref.x)
                                           and r.iterations == ref.iterations),
        "speedup_vs_csr": ref.gpu_ms / r.gpu_ms,
        "mixed_precision": {
            "value": mwall, "unit": "s", "iterations": m.iterations,
            "refinements": m.refinements, "flag": m.flag, "device_ms_per_solve": m.gpu_ms,
            "true_rel_res": m.true_res / m.res0,
            "speedup_vs_fp64_csr": ref.gpu_ms / m.gpu_ms,
            "speedup_vs_fp64_value_indexed": r.gpu_ms / m.gpu_ms,
            "scheme": "the mixed_precision solve with its FP32 off-diagonal copy value-indexed",
        },
        "spmv_roofline": {"bound": "hbm", "kernel": "k_spmv_pq_vi", "achieved": ach,
                          "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                          "algorithmic_bytes": vib, "avg_ms": ms},
        "delta_indexed": bench_delta_indexed(S, torch, A, b, P, tol, ref, reps),
    }


def bench_delta_indexed(S, torch, A, b, P, tol: float, ref, reps: int) -> dict:
    """The headline solve with the SpMV reading delta-indexed CSR (option
    spmv_format 2: per nonzero a 16-bit code, one byte indexing the 5
    distinct values, one the 7 distinct column offsets col - row; 2 B per
    nonzero).  Iterates bitwise the CSR solve's; the SpMV's roofline uses its
    own algorithmic bytes, 2*nnz + 4(n+1) + 16n.  The mixed-precision solve
    with its FP32 off-diagonal copy delta-indexed is timed too."""
    try:
        S.set_option("spmv_format", 2)
        S.pcg(A, b, P, tol=tol, maxit=25000)  # warm
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = S.pcg(A, b, P, tol=tol, maxit=25000)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        x = torch.randn(A.n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        ms = S.spmv_timed(A, x, y, reps, fused=True)
        S.pcg_mixed(A, b, P, tol=tol, maxit=25000)  # warm
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        m = S.pcg_mixed(A, b, P, tol=tol, maxit=25000)
        torch.cuda.synchronize()
        mwall = time.perf_counter() - t0
    finally:
        S.set_option("spmv_format", 0)
    dib = 2 * A.nnz + 4 * (A.n + 1) + 16 * A.n
    ach = dib / (ms * 1e-3) / 1e9
    return {
        "value": wall, "unit": "s", "iterations": r.iterations, "flag": r.flag,
        "device_ms_per_solve": r.gpu_ms,
        "ms_per_iteration": r.gpu_ms / max(r.iterations, 1),
        "bitwise_equal_to_csr_solve": bool(torch.equal(r.x, ref.x)
                                           and r.iterations == ref.iterations),
        "speedup_vs_csr": ref.gpu_ms / r.gpu_ms,
        "mixed_precision": {
            "value": mwall, "unit": "s", "iterations": m.iterations,
            "refinements": m.refinements, "flag": m.flag, "device_ms_per_solve": m.gpu_ms,
            "true_rel_res": m.true_res / m.res0,
            "speedup_vs_fp64_csr": ref.gpu_ms / m.gpu_ms,
            "speedup_vs_fp64_delta_indexed": r.gpu_ms / m.gpu_ms,
            "scheme": "the mixed_precision solve with its FP32 off-diagonal copy delta-indexed",
        },
        "spmv_roofline": {"bound": "hbm", "kernel": "k_spmv_pq_dvi", "achieved": ach,
                          "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                          "algorithmic_bytes": dib, "avg_ms": ms},
    }


def bench_adjoint(torch, np, n: int = 1_000_000, tol: float = 1e-10) -> dict:
    """Forward + adjoint solve of L_w x = b on a random geometric graph (n
    points U[0,1]^2, r = sqrt(10/(pi n)), w = exp(-d^2/sigma^2), sigma = r/2)
    + 1e-4 I, loss 0.5||x - x*||^2 with x* the solution for w*exp(0.1 xi).
    The points come in random order; the solver renumbers the nodes by
    reverse Cuthill-McKee inside (the default), and the same step in the
    caller's order is timed beside it."""
    from lap_amd import graph as GR
    from lap_amd.problems import random_geometric_graph
    g = random_geometric_graph(n, 12345)
    rng = np.random.default_rng(12345)
    t0 = time.perf_counter()
    G = GR.GraphLaplacian(g.n, g.src, g.dst, 1e-4)
    setup = time.perf_counter() - t0
    b = torch.from_numpy(rng.standard_normal(n)).cuda()
    wstar = torch.from_numpy(g.w * np.exp(0.1 * rng.standard_normal(g.m))).cuda()
    xstar = GR.laplacian_solve(G, wstar, b, tol=tol).detach()

    def step(GG):
        w = torch.from_numpy(g.w).cuda().requires_grad_(True)
        bb = b.clone().requires_grad_(True)
        x = GR.laplacian_solve(GG, w, bb, tol=tol)
        loss = 0.5 * ((x - xstar) ** 2).sum()
        loss.backward()
        return w.grad, bb.grad

    def timed(GG):
        step(GG)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        gw, gb = step(GG)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, gw, gb

    wall, gw, gb = timed(G)
    out = {
        "workload": f"random geometric graph n={n} m={g.m} + 1e-4*I, line_jacobi, tol {tol:g}, "
                    "loss 0.5||x - x*||^2, dl/dw and dl/db; nodes renumbered by reverse "
                    "Cuthill-McKee inside the solver",
        "value": wall, "unit": "s", "pattern_setup_s": setup,
        "forward_iterations": G.last_forward.iterations,
        "adjoint_iterations": G.last_adjoint.iterations,
        "forward_device_ms": G.last_forward.gpu_ms, "adjoint_device_ms": G.last_adjoint.gpu_ms,
        "grad_w_norm": float(torch.linalg.norm(gw)), "grad_b_norm": float(torch.linalg.norm(gb)),
    }
    Gn = GR.GraphLaplacian(g.n, g.src, g.dst, 1e-4, order="natural")
    nwall, ngw, ngb = timed(Gn)
    out["natural_order"] = {
        "s": nwall, "forward_iterations": Gn.last_forward.iterations,
        "adjoint_iterations": Gn.last_adjoint.iterations,
        "forward_device_ms": Gn.last_forward.gpu_ms,
        "grad_w_rel_diff": float(torch.linalg.norm(ngw - gw) / torch.linalg.norm(gw)),
        "grad_b_rel_diff": float(torch.linalg.norm(ngb - gb) / torch.linalg.norm(gb))}
    out["speedup_vs_natural_order"] = nwall / wall
    return out


def _warm(path: str) -> None:
    """Reads a file once so every timed reader finds it in the page cache."""
    with open(path, "rb") as f:
        while f.read(1 << 26):
            pass


def bench_input(S, torch, np, A, Ah, size: int) -> dict:
    """Input path on the size^3 matrix.  The matrix is written as a Matrix
    Market file (lower triangle, %.17g) and read back: to a device CSR
    (lap_mm_load on all host cores + lap_coo_to_csr on the device), split into
    the GS_std triangles on the device (lap_csr_split), through the library's
    io_utils.h definitions on the host (read_mm_file + coo_to_csr + split),
    and as a binary CSR streamed to the device.  Every file is read once
    before its timing, so all readers see the page cache."""
    import shutil
    import tempfile

    from lap_amd import io as IO
    d = tempfile.mkdtemp(prefix="lap_io_")
    try:
        path = os.path.join(d, f"grid{size}.mtx")
        t0 = time.perf_counter()
        IO.write_mm(path, Ah)
        t_write = time.perf_counter() - t0
        fbytes = os.path.getsize(path)
        _warm(path)
        D, info, t = IO.load_mm(path, times=True)
        same = bool(torch.equal(D.ia, A.ia) and torch.equal(D.ja, A.ja)
                    and torch.equal(D.a, A.a))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        Lt, dv, Ut = IO.split(D, augmented=True)
        torch.cuda.synchronize()
        t_split = time.perf_counter() - t0
        del Lt, Ut, dv
        t0 = time.perf_counter()
        host = IO.host_load(path)
        t_host = time.perf_counter() - t0
        host_same = bool(np.array_equal(host[0].ja, Ah.ja) and np.array_equal(host[0].a, Ah.a))
        del host
        bpath = os.path.join(d, f"grid{size}.lapcsr")
        IO.save_csr(bpath, D)
        _warm(bpath)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        E = IO.load_csr(bpath)
        torch.cuda.synchronize()
        t_bin = time.perf_counter() - t0
        bin_same = bool(torch.equal(E.ja, A.ja) and torch.equal(E.a, A.a))
        bbytes = os.path.getsize(bpath)
        del D, E
        return {
            "workload": f"{size}^3 Laplacian as Matrix Market (lower triangle, {info.nnz} "
                        f"entries, {fbytes / 1e9:.2f} GB) and as binary CSR "
                        f"({bbytes / 1e9:.2f} GB), page cache warm",
            "value": t.total_s, "unit": "s",
            "device_csr_s": {"parse": t.parse_s, "upload_coo": t.upload_s,
                             "coo_to_csr": t.convert_s},
            "parse_gbs": fbytes / t.parse_s / 1e9,
            "split_augmented_device_s": t_split,
            "host_io_utils_s": t_host,
            "binary_csr_to_device_s": t_bin,
            "binary_gbs": bbytes / t_bin / 1e9,
            "write_mm_s": t_write,
            "identical_to_generated": same and host_same and bin_same,
            "host_threads": os.cpu_count(),
        }
    finally:
        shutil.rmtree(d, ignore_errors=True)


def cpu_input_baseline(np, size: int = 128) -> dict:
    """The reference's read_mm_file + coo_to_csr + split (io_utils.c compiled
    into oracle/_ref, one thread) on a size^3 Matrix Market file, beside the
    library's device path and io_utils.h definitions on the same file."""
    import shutil
    import tempfile

    import torch

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    from lap_amd import io as IO
    from lap_amd.problems import grid_laplacian
    if not os.path.exists(oracle.REF_CPU):
        return {"error": "reference CPU path not built (oracle/_ref)"}
    Ah = grid_laplacian([size] * 3, 1e-8)
    d = tempfile.mkdtemp(prefix="lap_io_ref_")
    try:
        path = os.path.join(d, f"grid{size}.mtx")
        IO.write_mm(path, Ah)
        _warm(path)
        ref = oracle.IoUtils(oracle.Reference(oracle.REF_CPU).lib)
        t0 = time.perf_counter()
        R = ref.load(path)
        t_ref = time.perf_counter() - t0
        t0 = time.perf_counter()
        D, _ = IO.load_mm(path)
        IO.split(D, augmented=True)
        torch.cuda.synchronize()
        t_dev = time.perf_counter() - t0
        t0 = time.perf_counter()
        IO.host_load(path)
        t_host = time.perf_counter() - t0
        same = bool(np.array_equal(R["A"].ja, Ah.ja) and np.array_equal(R["A"].a, Ah.a))
        return {"value": t_ref, "unit": "s", "cores": 1, "kind": "reference",
                "sample": f"read_mm_file + coo_to_csr + split on the {size}^3 file "
                          f"({os.path.getsize(path) / 1e9:.2f} GB, page cache warm)",
                "library_device_s": t_dev, "library_host_io_utils_s": t_host,
                "speedup_device": t_ref / t_dev, "speedup_host": t_ref / t_host,
                "reference_matches": same}
    finally:
        shutil.rmtree(d, ignore_errors=True)


def pmc_probe(S, torch, size: int, reps: int, walk: int = 0, bound: bool = False) -> int:
    """Launches for rocprofv3 --pmc: 4 elementwise copies (dst = src * 1) of a
    1 GiB float64 buffer (read 1 GiB, write 1 GiB each: the calibration of
    FETCH_SIZE and WRITE_SIZE) and `reps` launches of the fused SpMV on the
    size^3 Laplacian, its row blocks walked in order `walk` (option
    spmv_walk).  `bound`: the same arrays with every column index replaced by
    its own row (same bytes and the same LDS two-phase structure, but each
    row block reads only its own slice of x, once), which bounds what any
    x-reuse scheme could gain."""
    import numpy as np
    from lap_amd.problems import CSR, grid_laplacian
    src = torch.randn(1 << 27, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(4):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    log("copy calibration: 4 x 1 GiB")
    del src, dst
    Ah = grid_laplacian([size] * 3, 1e-8)
    if bound:
        rows = np.repeat(np.arange(Ah.n, dtype=np.int32), np.diff(Ah.ia))
        Ah = CSR(Ah.n, Ah.ia, rows, Ah.a)
    prev = S.set_option("spmv_walk", walk)
    try:
        A = S.DeviceCSR.from_csr(Ah)
        x = torch.randn(A.n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        ms = S.spmv_timed(A, x, y, reps, fused=True)
    finally:
        S.set_option("spmv_walk", prev)
    log(f"fused SpMV x{reps} (spmv_walk {walk}{', ja = row' if bound else ''}): {ms:.4f} ms "
        f"per launch, algorithmic bytes {spmv_bytes(A.n, A.nnz)}")
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--prec", default="line_jacobi")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--cpu-iters", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--spmv-reps", type=int, default=50)
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the preconditioner, multi-RHS, batched, adjoint and input-path "
                         "measurements")
    ap.add_argument("--only", default="",
                    help="comma-separated extras to run (preconditioners, value_indexed, energy, "
                         "power_law, sweeps, block, batched, adjoint, input); default all")
    ap.add_argument("--power-law-kinds", default="line_jacobi,ichol",
                    help="preconditioners of the power_law extra")
    ap.add_argument("--power-law-maxit", type=int, default=25000,
                    help="iteration cap of the power_law solves (profiling runs only)")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="only run the kernels the HBM-traffic counters are read on: a 1 GiB "
                         "device copy (calibration) and the fused SpMV, then exit")
    ap.add_argument("--pmc-walk", type=int, default=0,
                    help="--pmc-probe: the SpMV's row-block walk (option spmv_walk)")
    ap.add_argument("--pmc-bound", action="store_true",
                    help="--pmc-probe: every column index replaced by its row (x read once)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU (--gpus 1): the solver has no multi-GPU path", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from lap_amd import solver as S
    from lap_amd.problems import grid_laplacian, rhs_normal

    torch.cuda.set_device(0)
    if args.pmc_probe:
        return pmc_probe(S, torch, args.size, args.spmv_reps, args.pmc_walk, args.pmc_bound)
    t0 = time.perf_counter()
    Ah = grid_laplacian([args.size] * 3, 1e-8)
    b_h = rhs_normal(Ah.n, 12345)
    log(f"built {args.size}^3 Laplacian n={Ah.n} nnz={Ah.nnz} in {time.perf_counter() - t0:.1f}s")
    A = S.DeviceCSR.from_csr(Ah)
    b = S.to_dev(b_h)
    P = S.make_preconditioner(A, Ah, args.prec)
    log("setup done")

    for w in range(args.warmup):
        r = S.pcg(A, b, P, tol=args.tol, maxit=25000)
        log(f"warmup {w}: {r.iterations} iterations, {r.gpu_ms:.1f} ms, flag {r.flag}")
    torch.cuda.synchronize()
    results = []
    t_start = time.perf_counter()
    for s in range(args.steps):
        results.append(S.pcg(A, b, P, tol=args.tol, maxit=25000))
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t_start) / max(args.steps, 1)
    its = results[-1].iterations
    gpu_ms = float(np.mean([r.gpu_ms for r in results]))
    flags = {r.flag for r in results}
    log(f"timed {args.steps} solves: {its} iterations, wall {wall * 1e3:.1f} ms, "
        f"device {gpu_ms:.1f} ms, flags {flags}")

    # dominant kernel: the fused SpMV q = A p (+ p.q), timed live with HIP
    # events on the library stream
    x = torch.randn(A.n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    spmv_ms = S.spmv_timed(A, x, y, args.spmv_reps, fused=True)
    bytes_spmv = spmv_bytes(A.n, A.nnz)
    kname = "k_spmv_pq (fused CSR SpMV q=A p with p.q)"
    achieved = bytes_spmv / (spmv_ms * 1e-3) / 1e9
    it_bytes = jacobi_iteration_bytes(A.n, A.nnz)
    log(f"dominant kernel {spmv_ms:.4f} ms -> {achieved:.0f} GB/s")

    out = {
        "metric": "s to rel-res 1e-10 + HBM GB/s, FP64 PCG vs mixed-prec PCG, 256^3 Laplacian",
        "value": wall,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"3D 7-point graph Laplacian {args.size}^3 + 1e-8*I, b~N(0,1) seed "
                         f"12345 mean-subtracted, FP64 PCG ({args.prec}) from x0=0 to "
                         f"rel-res {args.tol:g}"),
            "n": A.n, "nnz": A.nnz, "preconditioner": args.prec,
            "iterations": its, "flag": sorted(flags),
        },
        "device_ms_per_solve": gpu_ms,
        "iteration_gbs": it_bytes * its / (gpu_ms * 1e-3) / 1e9,
        "roofline": {
            "bound": "hbm",
            "kernel": kname,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(args.size),
            "algorithmic_bytes": bytes_spmv,
            "avg_ms": spmv_ms,
        },
    }
    if not args.no_mixed:
        # mixed precision: FP32 inner PCG in FP64 reliable-update refinement
        for w in range(args.warmup):
            S.pcg_mixed(A, b, P, tol=args.tol, maxit=25000)
        torch.cuda.synchronize()
        mres = []
        t_start = time.perf_counter()
        for s in range(args.steps):
            mres.append(S.pcg_mixed(A, b, P, tol=args.tol, maxit=25000))
        torch.cuda.synchronize()
        mwall = (time.perf_counter() - t_start) / max(args.steps, 1)
        m = mres[-1]
        xd, xm = results[-1].x, m.x
        xd_c, xm_c = xd - xd.mean(), xm - xm.mean()
        err = float(torch.linalg.norm(xm_c - xd_c) / torch.linalg.norm(xd_c))
        out["mixed_precision"] = {
            "value": mwall, "unit": "s", "iterations": m.iterations,
            "refinements": m.refinements, "flag": m.flag,
            "device_ms_per_solve": float(np.mean([r.gpu_ms for r in mres])),
            "true_rel_res": m.true_res / m.res0,
            "speedup_vs_fp64": wall / mwall,
            "rel_err_vs_fp64_solution": err,
            "scheme": "FP32 inner PCG (FP32 off-diagonals, FP64 diagonal and row sums) in "
                      "FP64 reliable-update refinement, delta=0.01",
        }
        log(f"mixed: {m.iterations} iterations, {m.refinements} refinements, wall "
            f"{mwall * 1e3:.1f} ms, speedup {wall / mwall:.2f}x, err vs fp64 {err:.1e}")
    if not args.no_extras:
        for name, fn in (("preconditioners",
                          lambda: bench_preconditioners(S, torch, A, Ah, b, args.tol)),
                         ("value_indexed",
                          lambda: bench_value_indexed(S, torch, A, b, P, args.tol, results[-1],
                                                      args.spmv_reps)),
                         ("energy", lambda: bench_energy(S, torch, A, b, P, args.tol)),
                         ("power_law", lambda: bench_power_law(
                             S, torch, np, args.tol, kinds=args.power_law_kinds.split(","),
                             maxit=args.power_law_maxit)),
                         ("sweeps", lambda: bench_sweeps(S, torch, A, Ah, b, args.tol)),
                         ("block", lambda: bench_block(S, torch, np)),
                         ("batched", lambda: bench_batched(S, torch, np)),
                         ("adjoint", lambda: bench_adjoint(torch, np)),
                         ("input", lambda: bench_input(S, torch, np, A, Ah, args.size)),
                         ("c1", lambda: bench_c1(S, torch, np))):
            if args.only and name not in args.only.split(","):
                continue
            t0 = time.perf_counter()
            try:
                out[name] = fn()
                log(f"{name}: {json.dumps(out[name])} (measured in "
                    f"{time.perf_counter() - t0:.1f}s)")
            except Exception as exc:  # reported, never hides the main line
                out[name] = {"error": f"{type(exc).__name__}: {exc}"}
                log(f"{name} failed: {exc}")
    if not args.no_cpu:
        t0 = time.perf_counter()
        out["cpu_baseline"] = cpu_baseline(Ah, b_h, args.prec, args.tol, its, args.cpu_iters)
        log(f"cpu baseline in {time.perf_counter() - t0:.1f}s: {out['cpu_baseline']['value']:.1f} s")
        if not args.no_extras and "input" in out:
            t0 = time.perf_counter()
            try:
                out["input"]["cpu_baseline"] = cpu_input_baseline(np)
            except Exception as exc:  # reported, never hides the main line
                out["input"]["cpu_baseline"] = {"error": f"{type(exc).__name__}: {exc}"}
            log(f"input cpu baseline in {time.perf_counter() - t0:.1f}s: "
                f"{json.dumps(out['input']['cpu_baseline'])}")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
