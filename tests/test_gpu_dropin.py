"""Drop-in check: the reference's own host code (cg_driver2.c, cg.c, prec.c,
GS.c, it_jacobi.c, line_jacobi.c, blas.c, io_utils.c compiled with -DHIP=1 by
oracle/build_ref.sh) linked against liblapgfx.so instead of hip_blas.o +
rocSPARSE/rocBLAS, run on a Matrix Market input, agrees with the reference
CPU binary on the same files: same exit reason, iteration count within +-1.

The same for the reference's Laplacian driver (cg_driver.c): an adjacency
file becomes the unweighted graph Laplacian (io_utils.c:80-130, :364-420),
b = (-1)^i is projected off the constant vector (cg_driver.c:188-206) and the
singular system is solved, with the projection's dot/scal/axpy running on
the library too.

Each driver is also linked with io_utils.c left out (lap_dropin_io_*): the
file reading, COO -> CSR, split and Laplacian construction then come from the
library's io_utils.h definitions."""
import os
import re
import subprocess

import pytest

from cases import ring_chords
from lap_amd.problems import grid_edges, grid_laplacian, rhs_uniform, write_adjacency, write_mm, write_rhs

pytestmark = pytest.mark.gpu

REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref")
DROPIN = os.path.join(REF, "lap_dropin_cg")
CPU = os.path.join(REF, "lap_cpu_cg")
DROPIN_LAP = os.path.join(REF, "lap_dropin_laplacian")
DROPIN_IO = os.path.join(REF, "lap_dropin_io_cg")
DROPIN_IO_LAP = os.path.join(REF, "lap_dropin_io_laplacian")
IO_VARIANTS = {"io_utils.c": (DROPIN, DROPIN_LAP), "library_io": (DROPIN_IO, DROPIN_IO_LAP)}
CPU_LAP = os.path.join(REF, "lap_cpu_laplacian")


def run(exe, args):
    out = subprocess.run([exe, *args], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr[-2000:]
    text = out.stdout
    iters = int(re.search(r"Iters\s*:\s*(\d+)", text).group(1))
    res = float(re.search(r"Res\. norm\s*:\s*(\S+)", text).group(1))
    reason = re.search(r"Reason for exiting : (.*)", text).group(1).strip()
    return iters, res, reason, text


@pytest.mark.skipif(not (os.path.exists(DROPIN) and os.path.exists(CPU)),
                    reason="drop-in binary not built (needs the reference sources)")
@pytest.mark.parametrize("io", sorted(IO_VARIANTS))
@pytest.mark.parametrize("kind", ["line_jacobi", "GS_std", "ichol", "it_jacobi"])
def test_reference_driver_on_gfx950_backend(tmp_path, kind, io):
    print(f"[case] dropin kind={kind} io={io}", flush=True)
    exe = IO_VARIANTS[io][0]
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    A = grid_laplacian([128, 128], 1e-8)
    b = rhs_uniform(A.n, 12345)
    mm, rhs = str(tmp_path / "A.mtx"), str(tmp_path / "b.mtx")
    write_mm(mm, A)
    write_rhs(rhs, b)
    k = "2" if kind == "it_jacobi" else "1"
    args = [mm, kind, "1e-8", "25000", "1", k, rhs]
    gi, gres, greason, gtext = run(exe, args)
    ci, cres, creason, _ = run(CPU, args)
    print(f"[diag] dropin {kind}: gpu it {gi} res {gres:.6e} ({greason}); "
          f"cpu it {ci} res {cres:.6e} ({creason})", flush=True)
    assert greason == creason == "CG converged"
    assert abs(gi - ci) <= 1
    assert "TRUE Norm of r" in gtext


GRAPHS = {
    "grid2d_64": lambda: grid_edges([64, 64]),
    "grid3d_16": lambda: grid_edges([16, 16, 16]),
    "ring_chords_3000": lambda: ring_chords(3000, 3, 5),
}


@pytest.mark.skipif(not (os.path.exists(DROPIN_LAP) and os.path.exists(CPU_LAP)),
                    reason="Laplacian drop-in binary not built (needs the reference sources)")
@pytest.mark.parametrize("io", sorted(IO_VARIANTS))
@pytest.mark.parametrize("kind", ["line_jacobi", "GS_std", "it_jacobi", "GS_it2"])
@pytest.mark.parametrize("graph", sorted(GRAPHS))
def test_reference_laplacian_driver_on_gfx950_backend(tmp_path, graph, kind, io):
    print(f"[case] dropin laplacian graph={graph} kind={kind} io={io}", flush=True)
    exe = IO_VARIANTS[io][1]
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    n, src, dst = GRAPHS[graph]()
    mm = str(tmp_path / "adj.mtx")
    write_adjacency(mm, n, src, dst)
    k = "2" if kind in ("it_jacobi", "GS_it2") else "1"
    args = [mm, kind, "1e-10", "25000", "1", k]
    gi, gres, greason, gtext = run(exe, args)
    ci, cres, creason, _ = run(CPU_LAP, args)
    print(f"[diag] dropin laplacian {graph}/{kind}: gpu it {gi} res {gres:.6e} ({greason}); "
          f"cpu it {ci} res {cres:.6e} ({creason})", flush=True)
    assert greason == creason == "CG converged"
    assert abs(gi - ci) <= 1


DROPIN_SPMV = os.path.join(REF, "lap_dropin_spmv")
CPU_SPMV = os.path.join(REF, "lap_cpu_spmv")


def run_spmv(exe, mm, trials):
    out = subprocess.run([exe, mm, str(trials)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr[-2000:]
    norm = float(re.search(r"Norm of A\*x\s*:\s*(\S+)", out.stdout).group(1))
    per = float(re.search(r"Time \(av\., per one\)\s*:\s*(\S+)", out.stdout).group(1))
    return norm, per


@pytest.mark.skipif(not (os.path.exists(DROPIN_SPMV) and os.path.exists(CPU_SPMV)),
                    reason="SpMV drop-in binary not built (needs the reference sources)")
@pytest.mark.parametrize("shift", [0.0, 1e-8])
@pytest.mark.parametrize("dims", [[128, 128], [40, 30, 20]])
def test_reference_spmv_driver_on_gfx950_backend(tmp_path, dims, shift):
    """mm_driver.c (the reference's SpMV driver) linked against the library:
    y = A*1 through analyze_spmv + csr_matvec (mm_driver.c:94-117), then
    ||y|| through dot (:128).  Every row of an unshifted graph Laplacian sums
    to exactly 0 in CSR order (integer partial sums), so ||A*1|| is exactly
    0, as on the CPU binary; with a 1e-8 shift each row sums to ~1e-8 and
    the norm matches the CPU binary's within the dot's summation order."""
    print(f"[case] dropin spmv dims={dims} shift={shift:g}", flush=True)
    A = grid_laplacian(dims, shift)
    mm = str(tmp_path / "A.mtx")
    write_mm(mm, A)
    gnorm, gper = run_spmv(DROPIN_SPMV, mm, 10)
    cnorm, _ = run_spmv(CPU_SPMV, mm, 10)
    print(f"[diag] dropin spmv {dims} shift {shift:g}: gpu ||A*1|| {gnorm:.16e} "
          f"({gper * 1e6:.1f} us per SpMV incl. sync), cpu {cnorm:.16e}", flush=True)
    if shift == 0.0:
        assert gnorm == 0.0 and cnorm == 0.0
    else:
        assert cnorm > 0.0 and abs(gnorm - cnorm) <= 1e-12 * cnorm
