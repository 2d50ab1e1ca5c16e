"""Compile the gfx950 HIP library in-tree: dftbaby_amd/lib/libdftbaby_amd.so.

    python -m dftbaby_amd.build

Every .hip source under dftbaby_amd/csrc is compiled with hipcc for gfx950
(packed-FP32 instruction selection disabled) and linked into one shared
library with a plain C ABI (declared in include/dftbaby_amd.h).
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUTDIR = os.path.join(HERE, "lib")
LIBNAME = "libdftbaby_amd.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I", os.path.join(HERE, "..", "include")]


def _compile(src, obj):
    cmd = [HIPCC] + FLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, r.stderr))
    return obj


def build(verbose=True):
    os.makedirs(OUTDIR, exist_ok=True)
    objdir = os.path.join(OUTDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    jobs = {}
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        for f in srcs:
            obj = os.path.join(objdir, f[:-4] + ".o")
            jobs[ex.submit(_compile, os.path.join(CSRC, f), obj)] = f
        objs = []
        for fut in cf.as_completed(jobs):
            objs.append(fut.result())
    out = os.path.join(OUTDIR, LIBNAME)
    tmp = out + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + sorted(objs)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s" % r.stderr)
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    build()
    sys.exit(0)
