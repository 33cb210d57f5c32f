"""Builds libslabpro.so (HIP, gfx950) in-tree: slab-pro_amd/lib/libslabpro.so.

Each .hip translation unit is compiled by hipcc for gfx950 only, then linked
into one shared library exporting the C ABI of include/slab_abi.h.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libslabpro.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Per-TU floating-point contraction: the FP64 form is built without FMA
# contraction so its arithmetic is the reference's operation for operation.
SOURCES = {
    "slab_kernels.hip": ["-ffp-contract=off"],
    "slab_engine32.hip": ["-ffp-contract=fast"],
    "slab_abi.hip": ["-ffp-contract=off"],
}

COMMON = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("command failed: " + " ".join(cmd))
    return out


def build(verbose=False):
    os.makedirs(LIBDIR, exist_ok=True)
    objs = []
    for src, flags in SOURCES.items():
        obj = os.path.join(LIBDIR, src.replace(".hip", ".o"))
        out = _run([HIPCC] + COMMON + flags + ["-c", os.path.join(CSRC, src), "-o", obj])
        if verbose and out:
            print(out)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
