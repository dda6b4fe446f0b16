"""Build libbolt_hip.so (gfx950) in-tree.

    python bolt.jl_amd/build.py

The library holds the HIP kernels, the C ABI (include/bolt_hip.h) and the
host-side table producers; it is written next to this file.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libbolt_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["bolt_hip.hip", "host_tables.cpp", "host_api.cpp"]


def hip_flags():
    return ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
            # keep hipcc from emitting packed FP32 instructions
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-result", "-I", os.path.join(ROOT, "include")]


def build(verbose=False, nanfill=False):
    """libbolt_hip.so; nanfill=True builds the debug variant
    libbolt_hip_nanfill.so, whose solver starts every k-mode from a workspace
    filled with NaNs (load it with BOLT_HIP_LIB=<path>)."""
    objs = []
    tmp = os.path.join(HERE, "_build_nanfill" if nanfill else "_build")
    os.makedirs(tmp, exist_ok=True)
    for src in SOURCES:
        obj = os.path.join(tmp, src + ".o")
        cmd = [HIPCC] + hip_flags() + (["-DBOLT_DEBUG_NANFILL"] if nanfill else []) + \
              ["-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd = [HIPCC, "-x", "c++"] + [f for f in hip_flags() if not f.startswith("--offload-arch")
                                           and f not in ("-Xclang", "-target-feature", "-packed-fp32-ops")] + \
                  (["-ffp-contract=off"] if src == "host_tables.cpp" else []) + \
                  ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        objs.append(obj)
    lib = LIB.replace(".so", "_nanfill.so") if nanfill else LIB
    out_tmp = lib + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", out_tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(out_tmp, lib)
    return lib


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, nanfill="--nanfill" in sys.argv))
