"""Write tests/golden/mesh_fields.json: a crafted mesh file in the format of
the reference's data.txt (input.F90:147-194: three title records, then after a
title record each No_vp 3I8, xv yv 2F15.5, No_cp 3I8, zbv 10F8.3, nbe 20I7,
nbev 20I7) and the values the compiler of the reference build (amdflang)
reads from it with those formats.

The fields exercise Fortran's formatted input: blanks before, inside and
after a number (ignored), all-blank fields (zero), records shorter than their
fields (padded with blanks), a real without a decimal point (scaled by
10^-d), exponents with E, D, e or a bare sign, more digits than d, and signs.
A short Fortran program (no reference code) reads the file with the same
statements and prints each value (reals as their 64 bits).  Run in the build
container:
    python tests/golden/make_mesh_fields.py
"""
from __future__ import annotations

import json
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
FC = os.environ.get("FC", "/opt/rocm/bin/amdflang")
N0, NV = 4, 23
PROG = """program mf
  integer, parameter :: n0 = %d, nv = %d
  integer :: No_vp(n0,3), No_cp(n0,3), nbe(n0), nbev(nv), i, j
  real*8 :: xv(nv), yv(nv), zbv(nv)
  character*80 :: t
  open(22, file='mesh.txt', status='old')
  read(22,'(A46)') t
  read(22,'(A46)') t
  read(22,'(A46)') t
  read(22,'(A46)') t
  read(22,'(3I8)') ((No_vp(i,j),j=1,3),i=1,n0)
  read(22,'(A46)') t
  read(22,'(2F15.5)') (xv(i),yv(i),i=1,nv)
  read(22,'(A46)') t
  read(22,'(3I8)') ((No_cp(i,j),j=1,3),i=1,n0)
  read(22,'(A46)') t
  read(22,'(10F8.3)') (zbv(i),i=1,nv)
  read(22,'(A46)') t
  read(22,'(20I7)') (nbe(i),i=1,n0)
  read(22,'(A46)') t
  read(22,'(20I7)') (nbev(i),i=1,nv)
  write(*,'(12(1x,i0))') No_vp, No_cp
  write(*,'(8(1x,i0))') nbe
  write(*,'(23(1x,i0))') nbev
  do i = 1, nv
    write(*,'(3(1x,z16.16))') xv(i), yv(i), zbv(i)
  enddo
end program
""" % (N0, NV)


def fields(*f, w):
    """fixed-width fields, right-justified when shorter than w"""
    assert all(len(x) <= w for x in f), f
    return "".join(x.rjust(w) for x in f)


def crafted() -> str:
    L = [" title one", " title two", " title three", " --------NUMBERING OF CELL VERTICES----"]
    L += [fields("       1", "       2", "       3", w=8), fields("   1 2  ", "    +7  ", "-3      ", w=8),
          "       5", fields("        ", "      10", "  0     ", w=8)]
    L.append(" --------COORDINATES OF CELL VERTICES----")
    xy = [("        1.50000", "      -2.25000"), ("          12345", "         -12345"),
          ("       -2.5E+01", "        1.5D+02"), ("  3.14159265358", "             .5"),
          ("            -0.", "      1.0e-3   "), ("      12345E2  ", "          +1.  "),
          ("               ", "        0.10000"), ("    6.28319    ", "   1 234.5 6   "),
          ("        1.5e-01", "         2.5-01"), ("    1.23456E+10", "   -0.333333333"),
          ("    0.099999999", "           7.D0"), ("  0.30000000001", "       -.000001")]
    for a, b in xy:
        L.append(fields(a, b, w=15))
    L.append("        2.00000")  # short: yv reads blank
    L += [fields("%15.5f" % (0.1 * k), "%15.5f" % (-0.07 * k), w=15) for k in range(NV - len(xy) - 1)]
    L.append(" --------NUMBERING OF CELL CENTERS----")
    L += [fields("       0", "       2", "       3", w=8), fields("       1", "        ", "       4", w=8),
          fields("      -1", "       0", "       0", w=8), ""]
    L.append(" --------BOTTOM LEVELS ---------------")
    L.append(fields("   0.000", "  12.345", "   12345", "        ", "-1.5e+00", "   1.0D1", "     .25", "  -0.001",
                    "     7  ", "  1 . 5 ", w=8))
    L.append(fields(*["%8.3f" % (0.013 * k) for k in range(10)], w=8))
    L.append(fields("   1.000", "  -2.000", w=8))  # short: the third value reads blank
    L.append(" --------TYPE OF BOUNDARY CELLS ---------------")
    L.append(fields("      1", "      0", "     -1", "    2  ", w=7))
    L.append(" --------TYPE OF BOUNDARY VERTEX-----------")
    L.append(fields(*["%7d" % (k % 3) for k in range(20)], w=7))
    L.append(fields("      1", "   1 1 ", w=7))
    return "\n".join(L) + "\n"


def main():
    text = crafted()
    with tempfile.TemporaryDirectory() as td:
        with open(os.path.join(td, "mf.f90"), "w") as f:
            f.write(PROG)
        with open(os.path.join(td, "mesh.txt"), "w") as f:
            f.write(text)
        subprocess.run([FC, "-O2", "-o", os.path.join(td, "mf"), os.path.join(td, "mf.f90")], check=True,
                       capture_output=True)
        out = subprocess.run([os.path.join(td, "mf")], cwd=td, check=True, capture_output=True, text=True).stdout
    lines = out.splitlines()
    # No_vp and No_cp column by column (Fortran order), one record each
    no_vp, no_cp = [int(t) for t in lines[0].split()], [int(t) for t in lines[1].split()]
    nbe = [int(t) for t in lines[2].split()]
    nbev = [int(t) for t in lines[3].split()]
    reals = [ln.split() for ln in lines[4:4 + NV]]
    assert len(no_vp) == len(no_cp) == 3 * N0 and len(nbe) == N0 and len(nbev) == NV and len(reals) == NV
    rec = {"compiler": FC, "N_CELL0": N0, "N_VERT": NV, "text": text,
           "No_vp": no_vp, "No_cp": no_cp, "nbe": nbe, "nbev": nbev,
           "xv_bits": [r[0] for r in reals], "yv_bits": [r[1] for r in reals], "zbv_bits": [r[2] for r in reals]}
    with open(os.path.join(HERE, "mesh_fields.json"), "w") as f:
        json.dump(rec, f, indent=0)
    print("mesh_fields.json:", N0, "cells,", NV, "vertices")


if __name__ == "__main__":
    main()
