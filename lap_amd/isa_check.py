"""Static check of the gfx950 assembly of this library's kernels: no
workgroup barrier inside a divergent loop.

A loop whose exit is taken per lane (the compiler masks leaving lanes out of
exec and keeps iterating with the rest) must not contain ``s_barrier``: the
remaining lanes would keep arriving at the barrier while the masked lanes
wait for them, and a persistent kernel would spin for ever.  The compiler can
create such a loop from uniform source when it threads two branches on the
same lane-dependent condition (``threadIdx.x == 0`` before and after a
loop's back edge); the kernels avoid it with ``readfirstlane`` on
workgroup-uniform values and barriers between such branches, and this check
keeps it that way.

    python -m lap_amd.isa_check          # compile each source with -S, check

Loops are read from the compiler's own block annotations (``; in Loop:
Header=BBx Depth=d`` / ``Parent Loop``); a loop is divergent when one of its
own blocks (innermost loop == that loop) executes ``s_andn2_b64 exec, exec``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import re
import subprocess
import sys
import tempfile

from .build import CFLAGS, HIPCC, sources

_LABEL = re.compile(r"^(?:\.LBB(\d+_\d+)|; %bb\.(\d+)):(.*)")
_HDR = re.compile(r"Header=BB(\d+_\d+)")
_ENCL = re.compile(r"(?:Parent Loop|Header=)\s*BB(\d+_\d+)")


def _kernels(asm: str):
    """(name, lines) of every function body in an assembly listing."""
    name, body = None, []
    for line in asm.splitlines():
        m = re.match(r"^([A-Za-z_][\w.$]*):\s*(;.*)?$", line)
        if m and not line.startswith(".") and "@" in line:
            if name:
                yield name, body
            name, body = m.group(1), []
            continue
        if name:
            body.append(line)
            if "s_endpgm" in line and body:
                pass
    if name:
        yield name, body


def divergent_barrier_loops(lines: list[str]) -> list[str]:
    """Loop headers of divergent loops that contain s_barrier."""
    blocks, cur = [], None
    for i, line in enumerate(lines):
        m = _LABEL.match(line)
        if m:
            label = m.group(1)
            rest = m.group(3)
            inner = None
            if label and "Loop Header" in rest:
                inner = label
            else:
                h = _HDR.search(rest)
                if h:
                    inner = h.group(1)
            encl = set(_ENCL.findall(rest))
            j = i + 1
            while j < len(lines) and lines[j].lstrip().startswith(";"):
                encl.update(_ENCL.findall(lines[j]))
                j += 1
            if inner:
                encl.add(inner)
            cur = {"inner": inner, "all": encl, "lines": []}
            blocks.append(cur)
            continue
        if cur is not None:
            cur["lines"].append(line)
    loops: dict[str, list] = {}
    for b in blocks:
        for h in b["all"]:
            loops.setdefault(h, []).append(b)
    bad = []
    for h, bs in sorted(loops.items()):
        if not any("s_barrier" in x for b in bs for x in b["lines"]):
            continue
        own = [b for b in bs if b["inner"] == h]
        if any("s_andn2_b64 exec, exec" in x for b in own for x in b["lines"]):
            bad.append(h)
    return bad


def check_asm(asm: str) -> dict[str, list[str]]:
    """{kernel: [loop headers]} for every offending kernel of a listing."""
    out = {}
    for name, body in _kernels(asm):
        bad = divergent_barrier_loops(body)
        if bad:
            out[name] = bad
    return out


def _asm_of(src: str, outdir: str) -> str:
    out = os.path.join(outdir, os.path.basename(src)[:-4] + ".s")
    flags = [f for f in CFLAGS if f != "-fPIC"]
    res = subprocess.run([HIPCC, *flags, "--offload-device-only", "-S", src, "-o", out],
                         capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc -S failed for {src}:\n{res.stderr}")
    with open(out) as f:
        return f.read()


def check_sources(srcs=None, jobs: int = 8) -> dict[str, dict[str, list[str]]]:
    """{source: {kernel: [loop headers]}} for every source with a finding."""
    srcs = list(srcs or sources())
    with tempfile.TemporaryDirectory() as tmp, cf.ThreadPoolExecutor(jobs) as ex:
        asms = list(ex.map(lambda s: _asm_of(s, tmp), srcs))
    found = {}
    for src, asm in zip(srcs, asms):
        r = check_asm(asm)
        if r:
            found[os.path.basename(src)] = r
    return found


if __name__ == "__main__":
    res = check_sources()
    for src, ks in res.items():
        for k, loops in ks.items():
            print(f"{src}: {k}: barrier inside divergent loop(s) {loops}")
    print("ok" if not res else "FOUND")
    sys.exit(1 if res else 0)
