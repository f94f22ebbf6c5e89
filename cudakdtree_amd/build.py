"""Builds libcukd_amd.so (gfx950) in-tree with hipcc.

Usage: python -m cudakdtree_amd.build  (or cudakdtree_amd.build.build()).
Every .hip source compiles to an object in parallel; query_inst.hip is
compiled once per point dimension 1..8.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcukd_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -packed-fp32-ops: no packed FP32 instructions are generated.
# -ffp-contract=off: no fused multiply-add in distance evaluation.
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-ffp-contract=off", "-Wall", "-Wno-unused-function",
         "-I", os.path.join(HERE, "..", "include")]


def _jobs():
    jobs = []
    for name in sorted(os.listdir(CSRC)):
        if not name.endswith(".hip"):
            continue
        src = os.path.join(CSRC, name)
        if name == "query_inst.hip":
            for d in range(1, 9):
                jobs.append((src, os.path.join(OBJ, "query_d%d.o" % d),
                             ["-DCUKD_QUERY_D=%d" % d]))
        else:
            jobs.append((src, os.path.join(OBJ, name[:-4] + ".o"), []))
    return jobs


def _compile(job):
    src, obj, extra = job
    cmd = [HIPCC] + FLAGS + extra + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    jobs = _jobs()
    workers = min(len(jobs), max(1, min(16, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(workers) as ex:
        objs = list(ex.map(_compile, jobs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
