"""Build ``liblbm_hip.so`` in-tree with hipcc for gfx950."""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
# translation units (source, extra flags, object), compiled in parallel and
# linked into one library: the core, and the step kernels of each level
# capacity of lbm_levels.h
UNITS = [("lbm_core.hip", [], "lbm_core.o")] + [
    ("lbm_levels.hip", [f"-DLBM_LEVEL_CAP={cap}"], f"lbm_levels_{cap}.o") for cap in (16, 32, 48, 64)]
OUT = os.path.join(HERE, "liblbm_hip.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 arithmetic is kept off (v_pk_*_f32 not emitted)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # a*b+c is fused within a source expression only, never across
    # statements: the same expression rounds the same in every kernel it is
    # inlined into (fused and separate launch variants are bit-identical)
    "-ffp-contract=on",
]


def _quiet(stderr):
    return "\n".join(line for line in stderr.splitlines() if "packed-fp32-ops" not in line)


def build(verbose=True):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objs = [os.path.join(HERE, "csrc", obj) for _, _, obj in UNITS]
    procs = []
    for (src, extra, _), obj in zip(UNITS, objs):
        cmd = [hipcc, *HIPCC_FLAGS, *extra, "-c", "-o", obj, os.path.join(HERE, "csrc", src)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True))
    failed = []
    for proc in procs:
        _, err = proc.communicate()
        if proc.returncode != 0:
            sys.stderr.write(_quiet(err))
            failed.append(proc.returncode)
    if failed:
        raise RuntimeError(f"hipcc failed ({failed})")
    tmp = OUT + ".tmp"
    cmd = [hipcc, "--offload-arch=gfx950", "-fPIC", "-shared", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(_quiet(res.stderr))
        raise RuntimeError(f"hipcc link failed ({res.returncode})")
    os.replace(tmp, OUT)
    for obj in objs:
        os.remove(obj)
    return OUT


if __name__ == "__main__":
    build()
