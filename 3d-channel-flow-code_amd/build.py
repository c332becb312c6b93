"""In-tree build of the native libraries.

  libnsx_host.so  host set-up (host/*.c), gcc, no FMA contraction
  libnsx.so       HIP kernels + C-ABI (csrc/*.hip), hipcc for gfx950
  libnsx_fortran.a + nsx_iface.mod, nsx_dropin.mod
                  the Fortran host side (fortran/*.F90, amdflang): the
                  ISO_C_BINDING interfaces and the drop-in subroutines
  tests/fortran/_build/dropin_driver, savetec_driver
                  the Fortran test drivers of the drop-in, linked to libnsx.so

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

HIP_SOURCES = ["nsx_context.hip", "nsx_stage.hip", "nsx_bicgstab.hip", "nsx_diag.hip", "nsx_stats.hip",
               "nsx_files.hip", "nsx_adjoint.hip", "nsx_modal.hip"]
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
    _run(["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-Wall", "-Wextra", "-I", os.path.join(REPO, "include"),
          "-o", os.path.join(OUT, "libnsx_host.so"), os.path.join(HERE, "host", "nsx_case.c"), "-lm"])


def build_hip():
    os.makedirs(OUT, exist_ok=True)
    srcs = [os.path.join(HERE, "csrc", s) for s in HIP_SOURCES]
    tmp = os.path.join(OUT, "libnsx.so.tmp")
    _run([HIPCC, *HIP_FLAGS, "-I", os.path.join(REPO, "include"), "-I", os.path.join(HERE, "csrc"),
          "-o", tmp, *srcs])
    os.replace(tmp, os.path.join(OUT, "libnsx.so"))


FC = os.environ.get("FC", "/opt/rocm/bin/amdflang")


def build_fortran():
    """The Fortran host side; skipped (with a message) when no Fortran compiler exists."""
    if not os.path.exists(FC):
        print(f"build: {FC} not found, the Fortran drop-in is not built", flush=True)
        return
    fdir = os.path.join(OUT, "fortran")
    os.makedirs(fdir, exist_ok=True)
    objs = []
    for src in ("nsx_iface.F90", "nsx_dropin.F90"):
        obj = os.path.join(fdir, src.replace(".F90", ".o"))
        _run([FC, "-O2", "-fPIC", "-J", fdir, "-c", os.path.join(HERE, "fortran", src), "-o", obj])
        objs.append(obj)
    lib = os.path.join(OUT, "libnsx_fortran.a")
    if os.path.exists(lib):
        os.remove(lib)
    _run(["ar", "rcs", lib, *objs])
    tdir = os.path.join(REPO, "tests", "fortran", "_build")
    os.makedirs(tdir, exist_ok=True)
    rel = os.path.relpath(OUT, tdir)
    for drv in ("dropin_driver", "savetec_driver", "routines_driver"):
        _run([FC, "-O2", "-J", fdir, "-o", os.path.join(tdir, drv),
              os.path.join(REPO, "tests", "fortran", drv + ".F90"), lib, "-L", OUT, "-lnsx",
              f"-Wl,-rpath,$ORIGIN/{rel}"])


def build_all():
    build_host()
    build_hip()
    build_fortran()


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    {"host": build_host, "hip": build_hip, "fortran": build_fortran, "all": build_all}[what]()
