"""Build recipe of the native libraries, in-tree:

  pmsz_anonymous_amd/libpmsz.so      HIP kernels + C-ABI (include/pmsz.h), gfx950
  pmsz_anonymous_amd/libpmsz_io.so   host front end (include/pmsz_io.h)
  pmsz_anonymous_amd/pmsz            the command-line tool (pMSz's argv contract)

Run ``python -m pmsz_anonymous_amd.build``.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PMSZ_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-fno-gpu-flush-denormals-to-zero",
    # packed FP32 instructions stay off (the compiler would otherwise emit them)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-command-line-argument",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-Wall"]

ENGINE_SRC = ["pmsz_engine.hip", "pmsz_serve.cpp"]
ENGINE_DEPS = ["pmsz_common.h", os.path.join("..", "..", "include", "pmsz.h")]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out, flush=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")


def build_engine():
    out = os.path.join(HERE, "libpmsz.so")
    srcs = [os.path.join(CSRC, s) for s in ENGINE_SRC]
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", out + ".tmp", *srcs])
    os.replace(out + ".tmp", out)
    return out


def build_io():
    out = os.path.join(HERE, "libpmsz_io.so")
    _run(["g++", *HOST_FLAGS, "-shared", "-o", out + ".tmp",
          os.path.join(CSRC, "fieldio.cpp"), "-ldl", "-pthread"])
    os.replace(out + ".tmp", out)
    return out


def build_cli():
    src = os.path.join(CSRC, "pmsz_cli.cpp")
    if not os.path.exists(src):
        return None
    out = os.path.join(HERE, "pmsz")
    _run(["g++", *HOST_FLAGS, "-o", out + ".tmp", src,
          f"-L{HERE}", "-lpmsz", "-lpmsz_io", f"-Wl,-rpath,$ORIGIN",
          "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__", "-L/opt/rocm/lib", "-lamdhip64",
          "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(out + ".tmp", out)
    return out


def build_all():
    build_io()
    build_engine()
    build_cli()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
