"""Builds patchtable_amd/libpatchtable_hip.so in-tree with hipcc for gfx950.

Every HIP source is compiled with floating-point contraction off (bit-exact parity with the
reference's unfused x86 arithmetic) and with packed-FP32 instructions disabled.
Usage: python patchtable_amd/build.py
"""
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libpatchtable_hip.so")
CLI = os.path.join(HERE, "patchtable")
CLI_SRC = os.path.join(HERE, "cli", "patchtable_main.cpp")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
          "-Wall", "-Wno-unused-function", "-Wno-unused-result"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(jobs=8):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.cpp")) + glob.glob(os.path.join(CSRC, "*.hip")))
    objs, procs = [], []
    for s in srcs:
        o = os.path.join(objdir, os.path.basename(s) + ".o")
        objs.append(o)
        if s.endswith(".hip"):
            cmd = [HIPCC, *COMMON, *DEVICE, "-x", "hip", "-c", s, "-o", o]
        else:
            cmd = [HIPCC, *COMMON, "-c", s, "-o", o]
        print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        if len(procs) >= jobs:
            if procs.pop(0).wait() != 0:
                raise RuntimeError("compilation failed")
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError("compilation failed")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", OUT + ".tmp", *objs, "-lz"])
    os.replace(OUT + ".tmp", OUT)
    # the command-line front end (host code only) links the library by its in-tree path
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I" + INCLUDE, CLI_SRC, "-o", CLI + ".tmp",
          "-L" + HERE, "-lpatchtable_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(CLI + ".tmp", CLI)
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
