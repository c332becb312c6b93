"""In-tree build of the native libraries.

  libnsx_host.so  host set-up (host/*.c), gcc, no FMA contraction
  libnsx.so       HIP kernels + C-ABI (csrc/*.hip), hipcc for gfx950

Outputs land in nsx/_build/ so they travel with the source tree.  Run as
`python build.py [host|hip|all]` or through __graft_entry__.build().
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
OUT = os.path.join(HERE, "nsx", "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("NSX_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["nsx_context.hip", "nsx_stage.hip", "nsx_bicgstab.hip"]
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
    # arithmetic order of the reference: no contraction into FMA, so the
    # element-wise operators reproduce the CPU results bit for bit
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_host():
    os.makedirs(OUT, exist_ok=True)
    _run(["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-Wall", "-Wextra",
          "-o", os.path.join(OUT, "libnsx_host.so"), os.path.join(HERE, "host", "nsx_case.c"), "-lm"])


def build_hip():
    os.makedirs(OUT, exist_ok=True)
    srcs = [os.path.join(HERE, "csrc", s) for s in HIP_SOURCES]
    tmp = os.path.join(OUT, "libnsx.so.tmp")
    _run([HIPCC, *HIP_FLAGS, "-I", os.path.join(REPO, "include"), "-I", os.path.join(HERE, "csrc"),
          "-o", tmp, *srcs])
    os.replace(tmp, os.path.join(OUT, "libnsx.so"))


def build_all():
    build_host()
    build_hip()


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    {"host": build_host, "hip": build_hip, "all": build_all}[what]()
