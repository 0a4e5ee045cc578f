"""lap_set_option (host-only, no GPU needed): every option takes its valid
range, returns the previous value, and leaves the setting unchanged with an
argument error outside it; unknown names return -1."""
import os

import pytest

from lap_amd import lib as L

pytestmark = pytest.mark.skipif(not os.path.exists(L.LIB_PATH), reason="library not built")

# name: (valid values, an invalid value or None, default)
OPTIONS = {
    "spmv_rows": ([1, 256, 512], 513, 256),
    "spmv_walk": ([0, 3], 1, 0),
    "fuse_sweeps": ([0, 1], 2, 1),
    "spmv_format": ([0, 1, 2], 3, 0),
    "mixed_vectors": ([0, 1], 2, 0),
    "sptrsv_mode": ([0, 3], 1, 3),
    "sptrsv_long": ([0, 1], -1, 0),
    "spmm_wg_per_cu": ([1, 16], 17, None),
    "spmv_wg_per_cu": ([1, 16], 0, 8),
    "vec_grid": ([1, 4096], 4097, 768),
    "graph_iters": ([1, 4096], 0, 16),
    "use_graph": ([0, 1], 2, 1),
}


@pytest.mark.parametrize("name", sorted(OPTIONS))
def test_option_range(name):
    lib = L.load()
    valid, invalid, default = OPTIONS[name]
    key = name.encode()
    lib.lap_clear_error()
    start = lib.lap_set_option(key, valid[0])
    if default is not None:
        assert start == default
    try:
        prev = valid[0]
        for v in valid[1:]:
            assert lib.lap_set_option(key, v) == prev
            prev = v
        assert lib.lap_last_status() == 0
        # out of range: refused, setting kept, argument error raised
        assert lib.lap_set_option(key, invalid) == prev
        assert lib.lap_last_status() != 0
        assert name.encode() in lib.lap_last_error() or b"out of range" in lib.lap_last_error()
        lib.lap_clear_error()
        assert lib.lap_set_option(key, prev) == prev
    finally:
        lib.lap_set_option(key, start)
        lib.lap_clear_error()


def test_unknown_option():
    lib = L.load()
    assert lib.lap_set_option(b"no_such_option", 1) == -1
    # options of measured-and-dropped variants are gone
    for name in (b"spmv_variant", b"jacobi_variant", b"sf_wg_per_cu", b"sf_naps"):
        assert lib.lap_set_option(name, 0) == -1
