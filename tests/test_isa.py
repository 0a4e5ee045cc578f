"""CPU check of the gfx950 assembly: no kernel of the library has a
workgroup barrier inside a divergent loop (a hang on the device).

The checker itself is pinned on two small kernels compiled here: a
persistent work-queue loop with thread-0 branches on both sides of the back
edge (the compiler turns it into a loop that only lane 0 leaves; it must be
flagged) and the same loop with the slot made uniform and a barrier closing
each round (it must pass)."""
import os
import shutil
import subprocess

import pytest

from lap_amd import build, isa_check

pytestmark = pytest.mark.skipif(shutil.which(build.HIPCC) is None
                                and not os.path.exists(build.HIPCC), reason="no hipcc")

QUEUE_LOOP = r"""
#include <hip/hip_runtime.h>
__global__ __launch_bounds__(256) void k(int nsys, int *queue, const double *v, double *out) {
  __shared__ int slot;
  __shared__ double red[4];
  for (;;) {
    if (threadIdx.x == 0) slot = atomicAdd(queue, 1);
    __syncthreads();
    const int s = %(READ)s;
    __syncthreads();
    if (s >= nsys) return;
    double acc = v[s * 256 + threadIdx.x];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x / 64] = acc;
    __syncthreads();
    if (threadIdx.x == 0) out[s] = red[0] + red[1] + red[2] + red[3];
    %(TAIL)s
  }
}
"""


def _asm(tmp_path, src):
    p = tmp_path / "k.hip"
    p.write_text(src)
    out = tmp_path / "k.s"
    flags = [f for f in build.CFLAGS if f != "-fPIC"]
    subprocess.run([build.HIPCC, *flags, "--offload-device-only", "-S", str(p), "-o", str(out)],
                   check=True, capture_output=True)
    return out.read_text()


def test_checker_flags_thread0_threaded_queue_loop(tmp_path):
    asm = _asm(tmp_path, QUEUE_LOOP % {"READ": "slot", "TAIL": ""})
    assert isa_check.check_asm(asm), "expected a divergent loop around the barrier"


def test_checker_passes_uniform_queue_loop(tmp_path):
    asm = _asm(tmp_path, QUEUE_LOOP % {"READ": "__builtin_amdgcn_readfirstlane(slot)",
                                       "TAIL": "__syncthreads();"})
    assert isa_check.check_asm(asm) == {}


def test_library_has_no_barrier_in_divergent_loop():
    assert isa_check.check_sources() == {}
