"""Run the compiled Fortran reference (oracle/_ref) on an nsx case.

Test infrastructure only: writes the binary case file read by
nsx_ref_driver.F90, runs the driver, and parses its dump into numpy arrays
shaped like the Fortran arrays.  The reference itself is built from
/root/reference by build_ref.sh and exists only where that tree does.
"""
from __future__ import annotations

import os
import re
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_OUT = os.path.join(os.path.dirname(HERE), "_ref")
MAGIC = 1314084913


# the channel's boundary conditions (TopBC, BotBC, XPB, YPB, XBC, YBC) of the
# default builds; build_ref.sh gives any other set a directory of its own
CHANNEL_BC = (0, 0, 1, 1, 0, 1)
DEFAULT_BOX = ("6.283185307179586", "3.141592653589793")


def driver_path(nn: int, box: tuple | None = None, bc: tuple | None = None) -> str:
    sub = f"nn{nn}" if box is None else f"nn{nn}_box{box[0]}x{box[1]}"
    if bc is not None and tuple(bc) != CHANNEL_BC:
        sub += "_bc" + "".join(str(int(v)) for v in bc)
    return os.path.join(REF_OUT, sub, "nsx_ref_driver")


def ensure_built(nn: int, box: tuple | None = None, bc: tuple | None = None) -> str:
    """The reference driver for nn layers; box = (Lx, Ly) as the strings the
    build writes into common.mpf (default 2*pi x pi); bc = (TopBC, BotBC,
    XPB, YPB, XBC, YBC) (default the channel's)."""
    exe = driver_path(nn, box, bc)
    src = os.path.join(HERE, "nsx_ref_driver.F90")
    if not os.path.exists(exe) or os.path.getmtime(src) > os.path.getmtime(exe):
        extra = [] if box is None else [str(box[0]), str(box[1])]
        if bc is not None and tuple(bc) != CHANNEL_BC:
            extra = (extra or list(DEFAULT_BOX)) + ["2.0"] + [str(int(v)) for v in bc]
        subprocess.run(["bash", os.path.join(HERE, "build_ref.sh"), str(nn), *extra], check=True,
                       stdout=subprocess.DEVNULL)
    return exe


def write_case(path: str, case, *, mode: int, nsteps: int = 1, dump_every: int = 1,
               p0=None, mesh_from_file: bool = False) -> None:
    p = case.params
    hdr = np.zeros(16, dtype=np.int32)
    hdr[:11] = [MAGIC, case.N_CELL0, case.N_VERT, case.nn, nsteps, p.FinRK, p.MaxIters, p.Iturbu,
                mode, dump_every, 1 if mesh_from_file else 0]
    rhdr = np.zeros(16, dtype=np.float64)
    rhdr[:6] = [p.dt, p.Re, p.bdx, p.bdy, p.bdz, p.eps]
    n0 = case.N_CELL0
    pressure = case.p0 if p0 is None else p0
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(rhdr.tobytes())
        f.write(np.asfortranarray(case.No_vp).astype(np.int32).tobytes(order="F"))
        f.write(np.asfortranarray(case.No_cp[:n0]).astype(np.int32).tobytes(order="F"))
        f.write(case.nbe.astype(np.int32).tobytes())
        f.write(case.nbev.astype(np.int32).tobytes())
        for a in (case.xv, case.yv, case.zbv, case.sig, case.dsig, case.sigv, case.dsigv):
            f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
        for a in (case.u0, case.v0, case.w0, pressure):
            f.write(np.asfortranarray(a[:n0]).tobytes(order="F"))


FLAT = {"ic1tec", "ic2tec", "ic3tec", "Numictec"}


def read_dump(path: str, case) -> dict:
    out = {}
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    while pos < len(data):
        name = data[pos:pos + 32].decode().strip()
        typ = int(np.frombuffer(data, np.int32, 1, pos + 32)[0])
        n = int(np.frombuffer(data, np.int64, 1, pos + 36)[0])
        pos += 44
        dt = np.float64 if typ == 1 else np.int32
        arr = np.frombuffer(data, dt, n, pos).copy()
        pos += n * arr.itemsize
        # the Tecplot connectivity lists stay flat whatever their length
        out[name] = arr if name in FLAT else _shape(arr, case)
    return out


def _shape(a, c):
    n = a.size
    table = [(c.N_CELL * c.NZ, (c.N_CELL, c.NZ)), (c.N_CELL0 * c.NZ, (c.N_CELL0, c.NZ)),
             (c.N_VERT * (c.NZ - 1), (c.N_VERT, c.NZ - 1)), (3 * c.N_CELL, (c.N_CELL, 3)),
             (3 * c.N_CELL0, (c.N_CELL0, 3)), (10 * c.N_VERT, (c.N_VERT, 10))]
    for size, shape in table:
        if n == size:
            return a.reshape(shape, order="F")
    return a


_ITER = re.compile(r"iter=\s*(\d+)\s*e1norm=\s*(\S+)\s*e2norm=\s*(\S+)")


def parse_solves(log: str) -> list:
    """Final (iter, e1norm, e2norm) of every BiCGSTAB solve printed by the reference."""
    solves = []
    cur = None
    for line in log.splitlines():
        if "SOLVE FOR P FIELD" in line:
            if cur is not None:
                solves.append(cur)
            cur = (0, None, None)
        m = _ITER.search(line)
        if m and cur is not None:
            cur = (int(m.group(1)), float(m.group(2).replace("D", "E")), float(m.group(3).replace("D", "E")))
    if cur is not None:
        solves.append(cur)
    return solves


def run(case, *, mode: int, nsteps: int = 1, dump_every: int = 1, p0=None, workdir: str | None = None,
        mesh_file: str | None = None, box: tuple | None = None, bc: tuple | None = None):
    """Run the reference; returns (dump dict, stdout text).  mesh_file: a
    data.txt the reference's input.F90 reads instead of the case's mesh;
    box: (Lx, Ly) of a build other than the default 2*pi x pi; bc: the
    boundary conditions (TopBC, BotBC, XPB, YPB, XBC, YBC) of the build."""
    exe = ensure_built(case.nn, box, bc)
    with tempfile.TemporaryDirectory(dir=workdir) as td:
        # the reference writes its statistics to ../output/Stats of its working directory
        wd = os.path.join(td, "run")
        stats_dir = os.path.join(td, "output", "Stats")
        serial_dir = os.path.join(td, "output", "Serial")
        os.makedirs(wd)
        os.makedirs(stats_dir)
        os.makedirs(serial_dir)
        cf = os.path.join(td, "case.bin")
        of = os.path.join(td, "out.bin")
        if mesh_file:
            import shutil
            shutil.copy(mesh_file, os.path.join(wd, "data.txt"))
        write_case(cf, case, mode=mode, nsteps=nsteps, dump_every=dump_every, p0=p0,
                   mesh_from_file=bool(mesh_file))
        r = subprocess.run([exe, cf, of], cwd=wd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"reference driver failed ({r.returncode}):\n{r.stdout[-3000:]}\n{r.stderr[-3000:]}")
        out = r.stdout
        mu = os.path.join(wd, "meanuv.dat")
        if os.path.exists(mu):
            out += "".join("MEANUV " + ln for ln in open(mu))
        d = read_dump(of, case)
        d["stats_files"] = {fn: open(os.path.join(stats_dir, fn)).read() for fn in sorted(os.listdir(stats_dir))}
        d["serial_files"] = {fn: open(os.path.join(serial_dir, fn)).read() for fn in sorted(os.listdir(serial_dir))}
        return d, out


CFL_KEYS = {"MAX U": "umax", "MAX V": "vmax", "MAX W": "wmax", "MAX CFL": "cfl", "Total  flux": "flux_total",
            "Maxim  flux": "flux_max", "BC     flux": "flux_bc", "BOT    flux": "flux_bot", "TOP    flux": "flux_top",
            "Side   flux": "flux_side"}


def parse_cfl(stdout: str) -> dict:
    """The values glob_cfl printed after the NSX_CFL marker (and the means
    it appended to meanuv.dat, f12.5)"""
    res = {}
    tail = stdout.split("NSX_CFL", 1)[1] if "NSX_CFL" in stdout else ""
    for line in tail.splitlines():
        line = line.strip()
        if line.startswith("MEANUV "):
            t, u, v, w = (float(x) for x in line.split()[1:5])
            res.update(aveu=u, avev=v, avew=w)
            continue
        for k, name in CFL_KEYS.items():
            if line.startswith(k) and "=" in line:
                res[name] = float(line.split("=", 1)[1])
    return res
