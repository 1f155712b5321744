"""Builds libfem_gpu.so for gfx950 in-tree with hipcc (no JIT, no torch extension), and the C++ adapter
library libsurgsim_gpu.so (host code only, g++) on top of it."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "_lib", "libfem_gpu.so")
HOST_LIB = os.path.join(HERE, "_lib", "libsurgsim_gpu.so")
SOURCES = ["fem_kernels.hip", "fem_hex.hip", "fem_adjoint.hip", "fem_batch.hip", "fem_probe.hip", "fem_mcg.hip",
           "fem_lu.hip", "fem_api.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 VALU ops stay off (they are generated from plain float code otherwise)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def build(verbose=False):
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    objdir = os.path.join(HERE, "_lib", "obj")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    procs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src + ".o")
        lang = ["-x", "hip"]
        cmd = [HIPCC] + FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out.decode())
            raise RuntimeError("hipcc failed on " + src)
        elif verbose and out:
            sys.stderr.write(out.decode())
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    build_host(verbose)
    return LIB


def build_host(verbose=False):
    """C++ adapters (SurgSimGpu.h) linked against libfem_gpu.so, found next to it at run time."""
    tmp = HOST_LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-shared",
           os.path.join(HERE, "host", "SurgSimGpu.cpp"), os.path.join(HERE, "host", "Fem3DPly.cpp"), "-o", tmp,
           "-L" + os.path.dirname(LIB), "-lfem_gpu", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd))
    subprocess.check_call(cmd)
    os.replace(tmp, HOST_LIB)
    return HOST_LIB


if __name__ == "__main__":
    print(build(verbose=True))
