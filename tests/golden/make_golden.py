"""Writes the golden fixtures in tests/golden/ from the reference's own CPU
path, compiled from its sources by oracle/build_ref.sh into oracle/_ref/.

Two independent routes per case:
  * the reference's cg() in liblapref_cpu.so called on the in-memory CSR
    (full residual history and solution);
  * the reference's solver binary lap_cpu_cg run on a Matrix Market file and
    RHS file written here (iteration count and the summary's 17-digit final
    residual), which also exercises its MM reader and setup.
Run in this container:  python tests/golden/make_golden.py [--big]

`--spread` adds to every existing fixture the spread of the reference CPU
path's own arithmetic on that input: the oracle's OpenMP restatement of
openmp_blas.c (dot products summed by 2..16 threads, SpMV on 16 threads,
the racy shared column index made private) run for exactly the fixture's
iteration count (tol 0, maxit = iterations - 1), compared with the serial
solution through the same samples and sketch the GPU tests use
(`omp_spread`) and residual norm by residual norm (`omp_spread_history`,
relative to each norm).  The GPU equal-iteration tests hold x to 1e-10
relative and the history to 1e-7, or to twice these spreads where larger.
"""
from __future__ import annotations

import argparse
import hashlib
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle")]

import oracle  # noqa: E402
from lap_amd.problems import grid_laplacian, rhs_normal, rhs_uniform, write_mm, write_rhs  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
SKETCH_ROWS = 16


def input_hash(A, b) -> str:
    h = hashlib.sha256()
    for arr in (A.ia, A.ja, A.a, b):
        h.update(np.ascontiguousarray(arr).tobytes())
    return h.hexdigest()


def sketch(x: np.ndarray, seed: int = 777) -> np.ndarray:
    """Random +-1 projections of x (size-independent solution fingerprint)."""
    rng = np.random.default_rng(seed)
    out = np.zeros(SKETCH_ROWS)
    for k in range(SKETCH_ROWS):
        s = rng.integers(0, 2, x.shape[0]) * 2.0 - 1.0
        out[k] = s @ x
    return out


def run_binary(A, b, kind, tol, maxit, m, k):
    exe = os.path.join(ROOT, "oracle", "_ref", "lap_cpu_cg")
    with tempfile.TemporaryDirectory() as tmp:
        mm, rhs = os.path.join(tmp, "A.mtx"), os.path.join(tmp, "b.mtx")
        write_mm(mm, A)
        write_rhs(rhs, b)
        out = subprocess.run([exe, mm, kind, repr(tol), str(maxit), str(m), str(k), rhs],
                             capture_output=True, text=True, check=True).stdout
    iters = int(re.search(r"Iters\s*:\s*(\d+)", out).group(1))
    res = float(re.search(r"Res\. norm\s*:\s*(\S+)", out).group(1))
    return iters, res


def case(name, A, b, kind, tol, maxit=25000, m=1, k=1, full_x=False, binary=True):
    ref = oracle.Reference()
    x, it, flag, hist = ref.cg(A, b, kind, tol=tol, maxit=maxit, m=m, k=k)
    data = dict(iterations=it, flag=flag, history=hist, sketch=sketch(x), xnorm=np.linalg.norm(x),
                xmean=x.mean(), x_sample=x[:: max(1, A.n // 512)], tol=tol, maxit=maxit, m=m,
                k=k, kind=kind, input_sha256=input_hash(A, b))
    if full_x:
        data["x"] = x
    if binary:
        bit, bres = run_binary(A, b, kind, tol, maxit, m, k)
        assert bit == it, (name, bit, it)
        data["binary_iterations"] = bit
        data["binary_res"] = bres
    np.savez_compressed(os.path.join(OUT, name + ".npz"), **data)
    print(f"{name}: {kind} it={it} flag={flag} res={hist[-1]:.3e}")


def fixture_problem(name):
    """The system of a fixture, from its name."""
    if name.startswith("c1_"):
        A = grid_laplacian([128, 128], 1e-8)
        return A, rhs_uniform(A.n, 12345)
    if name.startswith("g32_"):
        A = grid_laplacian([32, 32, 32], 1e-8)
        return A, rhs_normal(A.n, 12345)
    if name.startswith("c2_"):
        A = grid_laplacian([256, 256, 256], 1e-8)
        return A, rhs_normal(A.n, 12345)
    from lap_amd.problems import chung_lu
    A = chung_lu(1 << 22)
    return A, rhs_normal(A.n, 12345)


def solution_dev(x, g, centre: bool):
    """Relative deviation of x from a fixture's solution through its 512
    samples and 16-row sketch (means removed when `centre`); the larger."""
    n = x.shape[0]
    step = max(1, n // 512)
    rng = np.random.default_rng(777)
    ones = np.array([(rng.integers(0, 2, n) * 2.0 - 1.0).sum() for _ in range(SKETCH_ROWS)])
    xm, gm = (x.mean(), float(g["xmean"])) if centre else (0.0, 0.0)
    xs, gs = x[::step] - xm, g["x_sample"] - gm
    sk, gk = sketch(x) - xm * ones, g["sketch"] - gm * ones
    return max(np.linalg.norm(xs - gs) / np.linalg.norm(gs),
               np.linalg.norm(sk - gk) / np.linalg.norm(gk))


def add_spread(names):
    orc = oracle.Oracle()
    for name in names:
        path = os.path.join(OUT, name + ".npz")
        g = dict(np.load(path, allow_pickle=False))
        A, b = fixture_problem(name)
        assert input_hash(A, b) == str(g["input_sha256"]), name
        it = int(g["iterations"])
        threads = (8,) if name.startswith(("c2_", "c3_")) else (2, 4, 8, 16)
        devs, hdevs = [], []
        for th in threads:
            x, its, _, h = orc.cg(A, b, str(g["kind"]), tol=0.0, maxit=it - 1, m=int(g["m"]),
                                  k=int(g["k"]), omp=(16, th))
            assert its == it, (name, its, it)
            devs.append(solution_dev(x, g, centre=not name.startswith("c3_")))
            hdevs.append(float(np.max(np.abs(h - g["history"]) / g["history"])))
        g["omp_spread"] = max(devs)
        g["omp_spread_history"] = max(hdevs)
        g["omp_spread_threads"] = np.array(threads)
        np.savez_compressed(path, **g)
        print(f"{name}: OpenMP-vs-serial spread {max(devs):.2e}, residual history "
              f"{max(hdevs):.2e} (threads {threads})", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--big", action="store_true", help="also the 256^3 Jacobi case (slow)")
    ap.add_argument("--big-prec", action="store_true",
                    help="only the 256^3 GS_std and IC0 cases (tens of minutes)")
    ap.add_argument("--power-law", action="store_true",
                    help="only the 2^22 power-law (Chung-Lu) Jacobi case (minutes)")
    ap.add_argument("--spread", nargs="*", metavar="NAME",
                    help="add the OpenMP-vs-serial spread to these fixtures (default: all)")
    args = ap.parse_args()
    if args.spread is not None:
        import glob
        names = args.spread or sorted(os.path.basename(p)[:-4]
                                      for p in glob.glob(os.path.join(OUT, "*.npz")))
        add_spread(names)
        return
    if not oracle.have_ref():
        sys.exit("build oracle/_ref first: bash oracle/build_ref.sh")
    if args.power_law:
        from lap_amd.problems import chung_lu
        A = chung_lu(1 << 22)
        b = rhs_normal(A.n, 12345)
        case("c3_line_jacobi", A, b, "line_jacobi", 1e-10, binary=False)
        return
    if args.big_prec:
        A = grid_laplacian([256, 256, 256], 1e-8)
        b = rhs_normal(A.n, 12345)
        for kind in ("GS_std", "ichol"):
            case(f"c2_{kind}", A, b, kind, 1e-10, binary=False)
        return
    # 2D 5-point 128x128 + 1e-8 I, b ~ U(-1,1) mean-subtracted, tol 1e-8
    A = grid_laplacian([128, 128], 1e-8)
    b = rhs_uniform(A.n, 12345)
    case("c1_line_jacobi", A, b, "line_jacobi", 1e-8, full_x=True)
    for kind, m, k in (("GS_std", 1, 1), ("ichol", 1, 1), ("it_jacobi", 1, 2),
                       ("GS_it2", 1, 2), ("none", 1, 1)):
        case(f"c1_{kind}", A, b, kind, 1e-8, m=m, k=k)
    # 3D 7-point 32^3 + 1e-8 I, b ~ N(0,1) mean-subtracted, tol 1e-10
    A = grid_laplacian([32, 32, 32], 1e-8)
    b = rhs_normal(A.n, 12345)
    for kind in ("line_jacobi", "GS_std", "ichol"):
        case(f"g32_{kind}", A, b, kind, 1e-10)
    if args.big:
        A = grid_laplacian([256, 256, 256], 1e-8)
        b = rhs_normal(A.n, 12345)
        case("c2_line_jacobi", A, b, "line_jacobi", 1e-10, binary=False)


if __name__ == "__main__":
    main()
