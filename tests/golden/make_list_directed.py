"""Write tests/golden/list_directed.json: REAL(8) values and how the compiler
of the reference build (amdflang) prints them list-directed (`write(*,*) x`,
the format of the "TKE =" line of glob_put_plane_mm, stats.F90:362) and with
the e12.5 edit descriptor (its data lines, stats.F90:258).

A five-line Fortran program (no reference code) is compiled with amdflang in
a temporary directory and fed seeded values spanning 1e-8 .. 1e18, both signs,
ties and the extremes.  Run in the build container:
    python tests/golden/make_list_directed.py
"""
from __future__ import annotations

import json
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
FC = os.environ.get("FC", "/opt/rocm/bin/amdflang")
PROG = """program ld
  real*8 :: x
  integer :: ios
  open(10, file='vals.txt')
  do
    read(10, *, iostat=ios) x
    if (ios /= 0) exit
    write(*,*) x
    write(*,'(e12.5)') x
  enddo
end program
"""


def main():
    rng = np.random.default_rng(5)
    vals = list(rng.uniform(0, 1, 150)) + list(10 ** rng.uniform(-8, 18, 150)) + list(-10 ** rng.uniform(-3, 3, 60))
    vals += [0.3, 2 / 3, 0.2 / 3, 0.1 + 0.2, 1e23, 5e-324, 1.7976931348623157e308, 9.5, 0.95, 0.095, 0.0999, 0.09,
             99.99999999999999, 0.0, 1.0, 1e15, 1e16, 0.5, 0.25, 0.125, 1.0 / 3.0, 123456789.0, 1e-5,
             0.21775297606029484]
    vals = [float(v) for v in vals]
    with tempfile.TemporaryDirectory() as td:
        with open(os.path.join(td, "ld.f90"), "w") as f:
            f.write(PROG)
        with open(os.path.join(td, "vals.txt"), "w") as f:
            f.write("".join(repr(v) + "\n" for v in vals))
        subprocess.run([FC, "-O2", "-o", os.path.join(td, "ld"), os.path.join(td, "ld.f90")], check=True,
                       capture_output=True)
        out = subprocess.run([os.path.join(td, "ld")], cwd=td, check=True, capture_output=True, text=True).stdout
    lines = out.splitlines()
    assert len(lines) == 2 * len(vals)
    listed = [ln.strip() for ln in lines[0::2]]
    e125 = lines[1::2]
    with open(os.path.join(HERE, "list_directed.json"), "w") as f:
        json.dump({"compiler": FC, "pairs": [[repr(v), p] for v, p in zip(vals, listed)],
                   "e12.5": [[repr(v), p] for v, p in zip(vals, e125)]}, f, indent=0)
    print("list_directed.json:", len(vals), "values")


if __name__ == "__main__":
    main()
