#!/usr/bin/env python3
"""Build librabung.so (HIP, gfx950) in-tree next to this file.

    python rabung-atlas32_amd/build.py [--debug]

Compiles every source under csrc/ with hipcc for gfx950 and links one shared
library with a plain C ABI (include/rabung.h).  Packed FP32 instructions are
disabled for all device code.
"""
import argparse
import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
CSRC = HERE / "csrc"
OUT = HERE / "librabung.so"
WORKER = HERE / "rbg_worker"
ARCH = os.environ.get("RBG_OFFLOAD_ARCH", "gfx950")

DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(debug=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    opt = ["-O1", "-g"] if debug else ["-O3"]
    common = ["-std=c++17", "-fPIC", "-I", str(ROOT / "include"), "-I", str(CSRC)] + opt
    objs = []
    build_dir = HERE / "_build"
    build_dir.mkdir(exist_ok=True)
    for src in sorted(CSRC.glob("*.hip")):
        obj = build_dir / (src.stem + ".o")
        run([hipcc, "-x", "hip", "--offload-arch=" + ARCH, "-c", str(src), "-o", str(obj)]
            + common + DEVICE_FLAGS)
        objs.append(obj)
    for src in sorted(CSRC.glob("*.cpp")):
        obj = build_dir / (src.stem + ".o")
        run([hipcc, "-c", str(src), "-o", str(obj), "-D__HIP_PLATFORM_AMD__"] + common)
        objs.append(obj)
    tmp = OUT.with_suffix(".so.tmp")
    run([hipcc, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", str(tmp)]
        + [str(o) for o in objs] + ["-lpthread", "-lz", "-lcrypto"])
    os.replace(tmp, OUT)
    worker_src = CSRC / "worker" / "rbg_worker.cpp"
    if worker_src.exists():
        run(["g++", "-O2", "-std=c++17", "-I", str(ROOT / "include"), str(worker_src), "-o",
             str(WORKER), "-L", str(HERE), "-lrabung", "-Wl,-rpath,$ORIGIN", "-lpthread"])
    return OUT


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--debug", action="store_true")
    args = ap.parse_args()
    build(args.debug)
    return 0


if __name__ == "__main__":
    sys.exit(main())
