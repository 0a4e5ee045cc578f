"""ctypes binding of liblapgfx.so (the C ABI in include/lap_gfx950.h).

The library is loaded from this package directory.  There is no fallback:
if liblapgfx.so is missing or cannot be loaded, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblapgfx.so")

VP = C.c_void_p
PI = C.c_void_p  # device pointers travel as void*
PD = C.c_void_p


class LapError(RuntimeError):
    pass


class PData(C.Structure):
    """lap_pdata: same layout as the reference's pdata (common.h:29-50)."""
    _fields_ = [("lia", VP), ("lja", VP), ("la", VP), ("lnnz", C.c_int),
                ("uia", VP), ("uja", VP), ("ua", VP), ("unnz", C.c_int),
                ("ichol_vals", VP), ("d", VP), ("d_r", VP), ("n", C.c_int),
                ("aux_vec1", VP), ("aux_vec2", VP), ("aux_vec3", VP),
                ("prec_op", C.c_char_p), ("m", C.c_int), ("k", C.c_int)]


class SolveInfo(C.Structure):
    _fields_ = [("iterations", C.c_int), ("flag", C.c_int), ("res0", C.c_double),
                ("res_final", C.c_double), ("true_res", C.c_double), ("gpu_ms", C.c_double),
                ("refinements", C.c_int)]


class MmInfo(C.Structure):
    """lap_mm_info (include/lap_io.h)."""
    _fields_ = [("n", C.c_int), ("m", C.c_int), ("nnz", C.c_int), ("nnz_csr", C.c_int),
                ("symmetric", C.c_int), ("pattern", C.c_int)]


IP = C.POINTER(C.c_int)

# name -> (restype, argtypes)
_SIGS = {
    "initialize_handles": (None, []),
    "analyze_spmv": (None, [C.c_int, C.c_int, PI, PI, PD, PD, PD, C.c_char_p]),
    "initialize_and_analyze_L_and_U_solve": (None, [C.c_int, C.c_int, PI, PI, PD, C.c_int, PI,
                                                    PI, PD]),
    "hip_dot": (C.c_double, [C.c_int, PD, PD]),
    "lap_dot_noalloc": (C.c_int, [C.c_int, PD, PD, PD]),
    "hip_scal": (None, [C.c_int, C.c_double, PD]),
    "hip_axpy": (None, [C.c_int, C.c_double, PD, PD]),
    "hip_csr_matvec": (None, [C.c_int, C.c_int, PI, PI, PD, PD, PD, C.POINTER(C.c_double),
                              C.POINTER(C.c_double), C.c_char_p]),
    "hip_lower_triangular_solve": (None, [C.c_int, C.c_int, PI, PI, PD, PD, PD, PD]),
    "hip_upper_triangular_solve": (None, [C.c_int, C.c_int, PI, PI, PD, PD, PD, PD]),
    "hip_vec_vec": (None, [C.c_int, PD, PD, PD]),
    "hip_vector_reciprocal": (None, [C.c_int, PD, PD]),
    "hip_vector_sqrt": (None, [C.c_int, PD, PD]),
    "hip_vec_copy": (None, [C.c_int, PD, PD]),
    "hip_vec_zero": (None, [C.c_int, PD]),
    "hip_vec_set": (None, [C.c_int, C.c_double, PD]),
    "initialize_ichol": (None, [C.c_int, C.c_int, PI, PI, PD]),
    "hip_ichol": (None, [PI, PI, PD, C.c_int, C.POINTER(PData), PD, PD]),
    "mallocForDevice": (VP, [VP, C.c_int, C.c_int]),
    "memcpyDevice": (None, [VP, VP, C.c_int, C.c_int, C.c_char_p]),
    "freeDevice": (None, [VP]),
    "cg": (None, [C.c_int, C.c_double, PI, PI, PD, PD, PD, C.c_double, C.POINTER(PData), C.c_int,
                  C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "prec_function": (None, [PI, PI, PD, C.c_int, C.POINTER(PData), PD, PD]),
    "lap_last_status": (C.c_int, []),
    "lap_last_error": (C.c_char_p, []),
    "lap_clear_error": (None, []),
    "lap_release_all": (None, []),
    "lap_forget": (None, [VP]),
    "lap_pcg": (C.c_int, [C.c_int, C.c_int, PI, PI, PD, PD, PD, C.c_double, C.POINTER(PData),
                          C.c_int, C.POINTER(C.c_double), C.POINTER(SolveInfo)]),
    "lap_pcg_mixed": (C.c_int, [C.c_int, C.c_int, PI, PI, PD, PD, PD, C.c_double,
                                C.POINTER(PData), C.c_int, C.c_double, C.POINTER(C.c_double),
                                C.POINTER(SolveInfo)]),
    "lap_pcg_batched": (C.c_int, [C.c_int, PI, C.c_int, C.c_int, PI, PI, PD, PD, PD, PD,
                                  C.c_double, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int),
                                  C.POINTER(C.c_int), C.POINTER(C.c_double),
                                  C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "lap_pcg_batched_tune": (C.c_int, [C.c_int, PI, C.c_int, C.c_int, PI, PI, PD, PD, PD,
                                       C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "lap_pcg_block": (C.c_int, [C.c_int, C.c_int, PI, PI, PD, PD, C.c_int, PD, PD, C.c_double,
                                C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int),
                                C.POINTER(C.c_int), C.POINTER(C.c_double),
                                C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "lap_graph_pattern": (C.c_int, [C.c_int, C.c_int, PI, PI, PI, PI, PI, PI]),
    "lap_graph_values": (C.c_int, [C.c_int, C.c_int, PI, PI, PI, PI, PD, C.c_double, PD, PD]),
    "lap_edge_grad": (C.c_int, [C.c_int, C.c_int, PI, PI, PD, PD, PD]),
    "lap_pcg_adjoint": (C.c_int, [C.c_int, C.c_int, PI, PI, PD, C.POINTER(PData), PD, PD,
                                  C.c_double, C.c_int, C.c_int, PI, PI, PD, PD,
                                  C.POINTER(C.c_double), C.POINTER(SolveInfo)]),
    "lap_csr_matmat": (C.c_int, [C.c_int, C.c_int, PI, PI, PD, C.c_int, PD, PD, C.c_int]),
    "lap_spmv_timed": (C.c_double, [C.c_int, C.c_int, PI, PI, PD, PD, PD, C.c_int]),
    "lap_pcg_spmv_timed": (C.c_double, [C.c_int, C.c_int, PI, PI, PD, PD, PD, C.c_int]),
    "lap_set_option": (C.c_int, [C.c_char_p, C.c_int]),
    # input path (include/lap_io.h)
    "lap_mm_load": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(VP), C.POINTER(MmInfo)]),
    "lap_mm_coo": (C.c_int, [VP, VP, VP, VP, C.c_int]),
    "lap_mm_free": (None, [VP]),
    "lap_rhs_load": (C.c_int, [C.c_char_p, C.c_int, VP, C.c_int]),
    "lap_mm_write": (C.c_int, [C.c_char_p, C.c_int, VP, VP, VP]),
    "lap_csr_save": (C.c_int, [C.c_char_p, C.c_int, VP, VP, VP]),
    "lap_csr_info": (C.c_int, [C.c_char_p, IP, IP]),
    "lap_csr_load": (C.c_int, [C.c_char_p, VP, VP, VP, C.c_int]),
    "lap_coo_to_csr": (C.c_int, [C.c_int, C.c_int, C.c_int, PI, PI, PD, C.c_int, C.c_int, PI,
                                 PI, PD, IP]),
    "lap_csr_split": (C.c_int, [C.c_int, PI, PI, PD, C.c_int, C.c_int, PI, PI, PD, C.c_int,
                                PI, PI, PD, PD, IP, IP]),
    "lap_adjacency_laplacian": (C.c_int, [C.c_int, PI, PI, PD, C.c_int, PI]),
}

# Symbols the header declares; the CPU test checks the library exports each.
EXPORTED = sorted(list(_SIGS) + [
    "hip_ichol", "GS_std", "GS_it", "GS_it2", "it_jacobi", "line_jacobi", "dot", "axpy", "scal",
    "csr_matvec", "lower_triangular_solve", "upper_triangular_solve", "ichol", "vec_vec",
    "vector_reciprocal", "vector_sqrt", "vec_copy", "vec_zero",
    # io_utils.h drop-in (include/lap_io.h)
    "read_mm_file", "read_adjacency_file", "coo_to_csr", "read_rhs", "split",
    "create_L_and_split"])

_lib = None


def load(path: str = LIB_PATH):
    """Loads liblapgfx.so once; raises LapError when it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise LapError(f"{path} is not built (python -m lap_amd.build)")
    lib = C.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check():
    """Raises LapError if the last library call recorded an error."""
    lib = load()
    st = lib.lap_last_status()
    if st != 0:
        msg = lib.lap_last_error().decode()
        lib.lap_clear_error()
        raise LapError(f"lapgfx status {st}: {msg}")
