"""Builds vaco_amd/libvaco_av1pf.so for gfx950 with hipcc (no cmake/ninja)."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libvaco_av1pf.so")
SOURCES = ["av1pf_cdef.hip", "av1pf_superres.hip", "av1pf_lr.hip", "av1pf_chain.hip", "av1pf_stage.hip",
           "av1pf_host.cpp", "av1pf_replay.cpp", "av1pf_gen.cpp", "av1pf_serve.cpp", "av1pf_syntax.cpp",
           "av1pf_static.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # keep hipcc from forming packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # split LDS accesses the compiler cannot prove aligned into aligned ones:
    # merged unaligned ds_read_b64/b96 stall the LDS pipeline on gfx950
    "-Xclang", "-target-feature", "-Xclang", "-unaligned-access-mode",
    # kernel arguments preloaded into SGPRs at wave launch: the descriptor
    # loads at a kernel's start wait one scalar round trip less (1080p
    # restoration 11.73 to 11.49 us)
    "-mllvm", "-amdgpu-kernarg-preload-count=16",
    "-Wall", "-Wno-unused-command-line-argument",
]


def build(verbose: bool = False) -> str:
    """Compile every source to an object in parallel, then link the library."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for s in SOURCES:
        obj = os.path.join(objdir, s.rsplit(".", 1)[0] + ".o")
        objs.append(obj)
        cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, s), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = []
    for s, p in procs:
        out = p.communicate()[0].decode(errors="replace")
        lines = [ln for ln in out.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
        if p.returncode != 0:
            failed.append((s, "\n".join(lines)))
        elif lines and verbose:
            print("\n".join(lines))
    if failed:
        raise RuntimeError("hipcc failed:\n" + "\n\n".join(f"[{s}]\n{m}" for s, m in failed))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", *objs, "-o", tmp]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout.decode(errors="replace"))
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
