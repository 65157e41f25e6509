"""Compiles libirtt.so (HIP kernels + C ABI) for gfx950 with hipcc, in-tree.

Flags:
  -ffp-contract=off                         no fused multiply-add is formed, so
                                            FP32 results follow the written
                                            operation order
  -fhip-fp32-correctly-rounded-divide-sqrt  IEEE round-to-nearest FP32 / and sqrt
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                            no packed FP32 instructions
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(HERE, "libirtt.so")

SOURCES = ["api.cpp", "compat_api.cpp", "serve.cpp", "history.cpp", "design.cpp", "lbvh.hip", "trace.hip", "operator.hip", "aca.hip",
           "solver.hip", "batch.hip", "adjoint.hip", "sparse.hip", "hmatrix.hip"]

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def build(verbose=False, out=None, extra_flags=(), objdir=None):
    """Compile each source to an object in parallel, then link libirtt.so
    (or `out`, with `extra_flags` added to every compile)."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    lib = out or LIB
    objdir = objdir or os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".", "_") + ".o")
        objs.append(obj)
        cmd = [hipcc] + [f for f in FLAGS if f != "-shared"] + list(extra_flags) + [
            "-I", os.path.join(PKG, "..", "include"), "-c",
            "-x", "hip", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                            stderr=subprocess.STDOUT)))
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            failed.append((src, text))
        elif verbose and text.strip():
            print(text)
    if failed:
        msg = "\n".join(f"--- {s} ---\n{t}" for s, t in failed)
        raise RuntimeError("hipcc failed:\n" + msg)
    tmp = lib + ".tmp"
    cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout.decode(errors="replace"))
    os.replace(tmp, lib)
    return lib


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
