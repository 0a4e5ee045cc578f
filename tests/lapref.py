"""cg_driver.c's CPU pipeline composed from the reference's compiled
functions (oracle/_ref/liblapref_cpu.so) -- TEST INFRASTRUCTURE: the checker
for lap_amd.laplacian.  tests/test_laplacian_reference.py pins it to the
reference's own lap_cpu_laplacian binary."""
import ctypes as C

import numpy as np

from cases import ring_chords
from lap_amd.problems import grid_edges


GRAPHS = {
    "grid3d_16": lambda: grid_edges([16, 16, 16]),
    "ring_chords_3000": lambda: ring_chords(3000, 3, 5),
}


def reference_pipeline(ref, path, weighted, kind, m, k, tol):
    """cg_driver.c's CPU path through the reference's compiled functions."""
    import oracle as O
    io = O.IoUtils(ref.lib)
    R = io.load(path, adjacency=True, weighted=int(weighted))
    A = R["A"]
    n = A.n
    lib = ref.lib
    PD = C.POINTER(C.c_double)
    for nm in ("scal",):
        getattr(lib, nm).argtypes = [C.c_int, C.c_double, PD]
    lib.dot.restype = C.c_double
    lib.dot.argtypes = [C.c_int, PD, PD]
    lib.axpy.argtypes = [C.c_int, C.c_double, PD, PD]
    lib.vector_sqrt.argtypes = [C.c_int, PD, PD]
    lib.vec_vec.argtypes = [C.c_int, PD, PD, PD]
    dp = O.dp
    e = np.ones(n)
    b = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    d = R["D_vals"].copy()
    if not weighted:
        lib.scal(n, 1.0 / np.sqrt(float(n)), dp(e))
        be = lib.dot(n, dp(e), dp(b))
        lib.axpy(n, -1.0 * be, dp(e), dp(b))
    else:
        aux = np.zeros(n)
        lib.vector_sqrt(n, dp(d), dp(aux))
        lib.vec_vec(n, dp(aux), dp(e), dp(aux))
        de = 1.0 / np.sqrt(lib.dot(n, dp(aux), dp(aux)))
        lib.scal(n, de, dp(aux))
        bwe = lib.dot(n, dp(b), dp(aux))
        lib.axpy(n, bwe * -1.0, dp(aux), dp(b))
    x, it, flag, hist = ref.cg(A, b, kind, tol=tol, maxit=25000, m=m, k=k)
    return x, it, flag, hist


def null_free(x, v):
    """x without its component along the null vector v."""
    v = v / np.linalg.norm(v)
    return x - (v @ x) * v
