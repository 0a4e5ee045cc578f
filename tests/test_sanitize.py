"""Sanitizer builds of the host code (CPU only, no GPU call).

* The input path (lap_amd/csrc/io.hip with core.hip): the threaded Matrix
  Market parser, the atomic COO -> CSR of the io_utils.h drop-in, split and
  create_L_and_split, and lap_mm_load / lap_mm_coo, built host-only with
  AddressSanitizer and with ThreadSanitizer and driven by
  tests/sanitize/io_harness.c.  Every array must equal the shipped
  library's on the same file, and neither sanitizer may report.  The harness
  links only io.hip and core.hip; the plan-registry and scratch symbols those
  files reference from other sources are left unresolved on purpose (the
  harness never reaches the functions that call them).
* The oracle (oracle/lap_oracle.c) with AddressSanitizer and
  UndefinedBehaviorSanitizer, driven by tests/sanitize/oracle_harness.c: PCG
  with every preconditioner kind, bitwise equal to the regular oracle build,
  and the OpenMP variant (not under ThreadSanitizer: libgomp is not
  instrumented and would only give false reports).
"""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

from lap_amd import lib as L
from lap_amd.problems import grid_edges, grid_laplacian, write_adjacency, write_mm

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "lap_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CLANG = "/opt/rocm/lib/llvm/bin/clang"

needs_hipcc = pytest.mark.skipif(not (os.path.exists(HIPCC) and os.path.exists(CLANG)),
                                 reason="hipcc / clang not installed")


def _run(cmd, **kw):
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=600, **kw)
    assert res.returncode == 0, f"{' '.join(cmd)}\n{res.stdout[-2000:]}\n{res.stderr[-4000:]}"
    return res


def _io_harness(tmp, san):
    """Builds io_harness with -fsanitize=<san> on the host side only."""
    objs = []
    for src in ("io.hip", "core.hip"):
        obj = os.path.join(tmp, f"{src[:-4]}_{san}.o")
        _run([HIPCC, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fPIC",
              "--cuda-host-only", "-Xarch_host", f"-fsanitize={san}", "-c",
              os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    h = os.path.join(tmp, f"harness_{san}.o")
    _run([CLANG, "-O1", "-g", "-fno-omit-frame-pointer", f"-fsanitize={san}", "-c",
          os.path.join(HERE, "sanitize", "io_harness.c"), "-o", h])
    exe = os.path.join(tmp, f"io_harness_{san}")
    _run([HIPCC, "-Xarch_host", f"-fsanitize={san}", "-Wl,--unresolved-symbols=ignore-all",
          "-o", exe, h, *objs])
    return exe


def _read_out(path):
    raw = open(path, "rb").read()
    head = np.frombuffer(raw[:40], np.int32)
    n, m, nnz_coo, nnz_csr, sym, nl, nu, nd, mm_nnz, mm_nnz_csr = (int(v) for v in head)
    off = 40
    out = {"fields": dict(n=n, m=m, nnz=nnz_coo, nnz_unpacked=nnz_csr, symmetric=sym,
                          L_nnz=nl, U_nnz=nu, D_nnz=nd), "mm": (mm_nnz, mm_nnz_csr)}

    def take(count, dtype):
        nonlocal off
        a = np.frombuffer(raw, dtype, count, off)
        off += a.nbytes
        return a

    for name, k in (("A", nnz_csr), ("L", nl), ("U", nu)):
        out[name] = (take(n + 1, np.int32), take(k, np.int32), take(k, np.float64))
    out["D_vals"] = take(n, np.float64)
    out["coo"] = (take(mm_nnz, np.int32), take(mm_nnz, np.int32), take(mm_nnz, np.float64))
    assert off == len(raw)
    return out


def _sanitizer_clean(res):
    text = res.stderr
    assert "ERROR: AddressSanitizer" not in text, text[-4000:]
    assert "WARNING: ThreadSanitizer" not in text, text[-4000:]
    assert "runtime error:" not in text, text[-4000:]


@needs_hipcc
@pytest.mark.skipif(not os.path.exists(L.LIB_PATH), reason="library not built")
@pytest.mark.parametrize("san", ["address", "thread"])
def test_input_path_under_sanitizer(tmp_path, san):
    import oracle as O
    exe = _io_harness(str(tmp_path), san)
    io = O.IoUtils(C.CDLL(L.LIB_PATH))
    mm = str(tmp_path / "a.mtx")
    write_mm(mm, grid_laplacian([60, 50, 40], 1e-8))
    n, src, dst = grid_edges([50, 40])
    adj = str(tmp_path / "adj.mtx")
    write_adjacency(adj, n, src, dst)
    for mode, path, weighted in (("mm", mm, None), ("adj", adj, 0), ("adj", adj, 1)):
        out = str(tmp_path / f"out_{mode}_{weighted}.bin")
        args = [exe, mode, path, out] + ([] if weighted is None else [str(weighted)])
        res = _run(args, env={**os.environ, "TSAN_OPTIONS": "halt_on_error=1",
                              "ASAN_OPTIONS": "detect_leaks=1"})
        _sanitizer_clean(res)
        got = _read_out(out)
        ref = io.load(path, adjacency=(mode == "adj"), weighted=weighted)
        for k, v in ref["fields"].items():
            assert got["fields"][k] == v, (mode, k)
        for name in ("A", "L", "U"):
            R = ref[name]
            for x, y in zip(got[name], (R.ia, R.ja, R.a)):
                assert np.array_equal(x, y), (mode, name)
        assert np.array_equal(got["D_vals"], ref["D_vals"])
        n_, m_, nnz_, sym_, rows, cols, vals = io.read(path, adjacency=(mode == "adj"))
        assert got["mm"][0] == nnz_
        for x, y in zip(got["coo"], (rows, cols, vals)):
            assert np.array_equal(x, y), (mode, "coo")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not installed")
def test_oracle_under_address_and_undefined_sanitizers(tmp_path, orc):
    exe = str(tmp_path / "oracle_harness")
    _run(["gcc", "-std=c99", "-O1", "-g", "-ffp-contract=off", "-fopenmp",
          "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-o", exe,
          os.path.join(HERE, "sanitize", "oracle_harness.c"),
          os.path.join(ROOT, "oracle", "lap_oracle.c"), "-lm"])
    A = grid_laplacian([48, 40], 1e-8)
    mm = str(tmp_path / "a.mtx")
    write_mm(mm, A)
    b = np.sin(np.arange(1, A.n + 1, dtype=np.float64))
    kinds = {0: "none", 1: "line_jacobi", 2: "it_jacobi", 3: "GS_std", 4: "GS_it",
             5: "GS_it2", 6: "ichol"}
    for omp in (False, True):
        res = _run([exe, mm] + (["omp"] if omp else []))
        _sanitizer_clean(res)
        lines = [ln.split() for ln in res.stdout.splitlines() if ln.startswith("kind")]
        assert len(lines) == (1 if omp else 7)
        for ln in lines:
            kind, it, flag, rn = kinds[int(ln[1])], int(ln[3]), int(ln[5]), float.fromhex(ln[7])
            if omp:
                _, it_r, flag_r, hist = orc.cg(A, b, kind, tol=1e-8, maxit=5000, m=1, k=2,
                                               omp=(4, 4))
                assert abs(it - it_r) <= 1 and flag == flag_r
            else:
                _, it_r, flag_r, hist = orc.cg(A, b, kind, tol=1e-8, maxit=5000, m=1, k=2)
                assert (it, flag) == (it_r, flag_r), kind
                assert rn == hist[-1], kind
