"""Build libcmbl.so (HIP, gfx950) in-tree with hipcc.

    python cmblensing.py_amd/build.py            # or __graft_entry__.build()
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
INCLUDE = PKG.parent / "include"
LIB = PKG / "libcmbl.so"
# (source, object, extra flags): kernels.hip is compiled twice, its FP32 and
# FP64 instantiations in separate objects, so the two compile concurrently
SOURCES = [("kernels.hip", "kernels_f32.o", ["-DCMBL_KERNELS_PART=1"]),
           ("kernels.hip", "kernels_f64.o", ["-DCMBL_KERNELS_PART=2"]),
           ("wf.hip", "wf.o", []), ("sim.hip", "sim.o", []), ("api.hip", "api.o", [])]
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 ops are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # multiply-adds fused only within one source expression, so every
    # instantiation of a kernel (launch shapes, batch sizes) rounds alike and
    # a request's result does not depend on the batch it runs in
    "-ffp-contract=on",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build libcmbl.so)")


def build(verbose: bool = False) -> Path:
    """Compile every source and link libcmbl.so; always recompiles."""
    cc = hipcc()
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    objs, procs = [], []
    # the sources compile concurrently (one hipcc process each)
    for src, oname, extra in SOURCES:
        obj = objdir / oname
        cmd = [cc, *HIPCC_FLAGS, *extra, f"-I{INCLUDE}", "-c", str(CSRC / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                                            text=True)))
        objs.append(str(obj))
    failed = []
    for src, pr in procs:
        _, err = pr.communicate()
        if pr.returncode != 0:
            failed.append(f"hipcc failed on {src}:\n{err}")
    if failed:
        raise RuntimeError("\n".join(failed))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
