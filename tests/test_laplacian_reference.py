"""The composed reference Laplacian pipeline (tests/lapref.py), the checker of
the device Laplacian mode, reproduces the reference's own lap_cpu_laplacian
binary (cg_driver.c built as MakefileCPU): same iteration count on every
graph and preconditioner of the GPU test (no GPU needed)."""
import os
import re
import subprocess

import pytest

from lapref import GRAPHS, reference_pipeline
from lap_amd.problems import write_adjacency

REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref")
CPU_LAP = os.path.join(REF, "lap_cpu_laplacian")


@pytest.mark.skipif(not os.path.exists(CPU_LAP), reason="reference binary not built")
@pytest.mark.parametrize("kind,m,k", [("line_jacobi", 1, 1), ("GS_std", 1, 1),
                                      ("it_jacobi", 1, 2), ("GS_it2", 1, 2)])
@pytest.mark.parametrize("graph", sorted(GRAPHS))
def test_composed_pipeline_is_the_reference_driver(ref, tmp_path, graph, kind, m, k):
    n, src, dst = GRAPHS[graph]()
    path = str(tmp_path / "adj.mtx")
    write_adjacency(path, n, src, dst)
    _, it, flag, hist = reference_pipeline(ref, path, False, kind, m, k, 1e-10)
    out = subprocess.run([CPU_LAP, path, kind, "1e-10", "25000", str(m), str(k)],
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0
    assert int(re.search(r"Iters\s*:\s*(\d+)", out.stdout).group(1)) == it
    assert flag == 0 and "CG converged" in out.stdout
    res = float(re.search(r"Res\. norm\s*:\s*(\S+)", out.stdout).group(1))
    assert abs(res - hist[-1]) <= 1e-12 * max(abs(res), 1e-300) or res == hist[-1]
