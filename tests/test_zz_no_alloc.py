"""lap_dot_noalloc: hip_dot for callers that cannot allocate.  Its partial sums
live in a static device array instead of the library's workspace; the
kernels and grid are hip_dot's, so every call must give hip_dot's bits.
Two calls in a row, the second smaller, catch partial sums left over from the
first; argument errors are refused on the host before any launch."""
import ctypes as C

import numpy as np
import pytest
import torch

from lap_amd import lib as L

pytestmark = pytest.mark.gpu


def _vectors(n, seed):
    g = torch.Generator().manual_seed(seed)
    v = torch.randn(n, dtype=torch.float64, generator=g)
    w = torch.randn(n, dtype=torch.float64, generator=g)
    return v, w


@pytest.mark.parametrize("sizes", [(3_000_017, 40_001), (1_000_000, 0), (257, 1)])
def test_noalloc_dot_matches_hip_dot_on_consecutive_calls(sizes):
    lib = L.load()
    lib.lap_clear_error()
    out = torch.full((2,), float("nan"), dtype=torch.float64, device="cuda:0")
    for i, n in enumerate(sizes):
        print(f"[case] lap_dot_noalloc call {i} n={n}", flush=True)
        v, w = _vectors(max(n, 1), 1000 + n)
        vd, wd = v[:n].to("cuda:0"), w[:n].to("cuda:0")
        torch.cuda.synchronize()
        pv, pw = (vd.data_ptr(), wd.data_ptr()) if n else (None, None)
        rc = lib.lap_dot_noalloc(n, pv, pw, out[i:].data_ptr())
        assert rc == 0, lib.lap_last_error()
        ref = lib.hip_dot(n, pv, pw) if n else 0.0
        L.check()
        got = out[i].item()
        # hip_dot's kernels and grid: the same bits
        assert np.float64(got).tobytes() == np.float64(ref).tobytes(), (n, got, ref)
        exact = float(np.dot(v[:n].numpy(), w[:n].numpy())) if n else 0.0
        assert abs(got - exact) <= 1e-12 * max(1.0, float(np.abs(v[:n].numpy()) @
                                                     np.abs(w[:n].numpy())))


def test_noalloc_dot_refuses_bad_arguments():
    lib = L.load()
    lib.lap_clear_error()
    d = torch.ones(16, dtype=torch.float64, device="cuda:0")
    res = torch.zeros(1, dtype=torch.float64, device="cuda:0")
    host = (C.c_double * 16)()
    assert lib.lap_dot_noalloc(-1, d.data_ptr(), d.data_ptr(), res.data_ptr()) != 0
    lib.lap_clear_error()
    assert lib.lap_dot_noalloc(16, d.data_ptr(), d.data_ptr(), C.addressof(host)) != 0
    lib.lap_clear_error()
    assert lib.lap_dot_noalloc(16, C.addressof(host), d.data_ptr(), res.data_ptr()) != 0
    lib.lap_clear_error()
    # still working after the refusals
    assert lib.lap_dot_noalloc(16, d.data_ptr(), d.data_ptr(), res.data_ptr()) == 0
    assert res.item() == 16.0
