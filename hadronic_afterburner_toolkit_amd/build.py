"""Build recipe of the HIP engine (gfx950) -- compiles in-tree, every call.

    python -m hadronic_afterburner_toolkit_amd.build

Products:
  hadronic_afterburner_toolkit_amd/libhbt_hip.so   C ABI of include/hbt_hip.h
  hadronic_afterburner_toolkit_amd/hbt_afterburner.e  CLI (drop-in driver)
  hadronic_afterburner_toolkit_amd/hist_atomics_bench.e  histogram-atomics
      micro-benchmark (csrc/microbench/hist_atomics.hip)
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: no FMA contraction, so every decision of the FP64
# cascade rounds like the reference's x86-64 build.  Packed FP32 VALU
# instructions are disabled (-packed-fp32-ops).
COMMON = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-ffp-contract=off",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wno-unused-result", "-Wno-unused-value", "-Wno-bitwise-instead-of-logical"]

HOST_SRCS = ["host/ParameterReader.cpp", "host/Random.cpp", "host/particleSamples.cpp",
             "host/HBT_correlation.cpp", "host/BalanceFunction.cpp"]  # the drop-in classes of the CLI


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build failed: " + " ".join(cmd))
    return out


def build_engine() -> str:
    out = os.path.join(HERE, "libhbt_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC] + COMMON + ["-shared", "-o", tmp, os.path.join(CSRC, "hbt_engine.hip"),
                             os.path.join(CSRC, "hbt_server.hip"),
                             os.path.join(CSRC, "hbt_bf.hip"),
                             os.path.join(CSRC, "hbt_noalloc.hip"),
                             os.path.join(CSRC, "host", "hbt_host.cpp"), "-lz", "-lpthread"])
    os.replace(tmp, out)
    return out


def build_cli() -> str | None:
    """The drop-in HBT_correlation class + CLI, linked against libhbt_hip.so."""
    srcs = [os.path.join(CSRC, s) for s in HOST_SRCS]
    if not all(os.path.exists(s) for s in srcs):
        return None
    out = os.path.join(HERE, "hbt_afterburner.e")
    _run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I" + os.path.join(REPO, "include"),
          "-I" + os.path.join(CSRC, "host"), "-o", out, os.path.join(CSRC, "host", "main.cpp")]
         + srcs + ["-L" + HERE, "-lhbt_hip", "-Wl,-rpath,$ORIGIN", "-lz"])
    return out


def build_microbench() -> str:
    out = os.path.join(HERE, "hist_atomics_bench.e")
    _run([HIPCC] + COMMON + ["-o", out, os.path.join(CSRC, "microbench", "hist_atomics.hip")])
    return out


def build_all() -> None:
    build_engine()
    build_cli()
    build_microbench()


if __name__ == "__main__":
    build_all()
