"""Build recipe of libxmatch_amd.so (hipcc, gfx950) -- in-tree, no JIT cache.

Compile flags:
  --offload-arch=gfx950          MI355X only
  -ffp-contract=off              the FP64 chord and the unit vectors are rounded exactly
                                 as written, on host and device alike
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                 no packed FP32 instructions are generated
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libxmatch_amd.so")
SOURCES = ["xm_capi.hip", "xm_sort.hip", "xm_catalog.hip", "xm_pair.hip", "xm_workcount.hip",
           "xm_synth.hip", "xm_server.hip", "xm_noalloc.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


CLI = os.path.join(HERE, "xmatch")
CXX = os.environ.get("CXX", "g++")
# the command-line tool is host code over the C ABI only: plain g++, linked to the
# library next to it
CLI_FLAGS = ["-O2", "-std=c++17", "-Wall", "-pthread", "-I" + os.path.join(HERE, "..", "include")]


def _run(cmd, what, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("%s failed building %s" % (cmd[0], what))


def build(verbose=False, extra=()):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    _run([HIPCC] + FLAGS + list(extra) + srcs + ["-o", tmp], LIB, verbose)
    os.replace(tmp, LIB)
    tmp = CLI + ".tmp"
    _run([CXX] + CLI_FLAGS + [os.path.join(CSRC, "xm_cli.cpp"), "-o", tmp, "-L" + HERE,
                              "-lxmatch_amd", "-Wl,-rpath,$ORIGIN"], CLI, verbose)
    os.replace(tmp, CLI)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
    print(CLI)
