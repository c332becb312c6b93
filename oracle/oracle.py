"""Python handle on the CPU restatement (nsx_oracle.c).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker or the timed CPU baseline.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(BUILD, "liboracle.so")
# the same source built with -fopenmp: the solver's loops split over threads
# (sums in another order), for the all-cores CPU baseline only
LIB_OMP = os.path.join(BUILD, "liboracle_omp.so")
SRC = os.path.join(HERE, "nsx_oracle.c")

_libs = {}


def build() -> str:
    os.makedirs(BUILD, exist_ok=True)
    # -ffp-contract=off and no -march: plain SSE2 double arithmetic in the
    # reference's operation order, the same as the Fortran build
    for out, extra in ((LIB, []), (LIB_OMP, ["-fopenmp"])):
        tmp = out + ".tmp"
        subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared", "-Wall", *extra,
                        "-o", tmp, SRC, "-lm"], check=True)
        os.replace(tmp, out)
    return LIB


def lib(parallel: bool = False) -> ctypes.CDLL:
    path = LIB_OMP if parallel else LIB
    if path not in _libs:
        if not os.path.exists(path) or os.path.getmtime(path) < os.path.getmtime(SRC):
            build()
        L = ctypes.CDLL(path)
        vp, ci, cd, cs = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_char_p
        L.orc_create.restype = vp
        L.orc_create.argtypes = [ci] * 5
        L.orc_free.argtypes = [vp]
        L.orc_set_array.argtypes = [vp, cs, vp]
        L.orc_state.restype = ctypes.POINTER(cd)
        L.orc_state.argtypes = [vp, cs]
        L.orc_set_params.argtypes = [vp, cd, cd, cd, cd, cd, cd, ci, ci, ci, ci, ci, ci, ci, cd, cd, cd]
        L.orc_init_state.argtypes = [vp]
        L.orc_step.argtypes = [vp]
        L.orc_hydro.argtypes = [vp]
        L.orc_solver_record.argtypes = [vp, ctypes.POINTER(ci), ctypes.POINTER(cd)]
        L.orc_timers.argtypes = [vp, ctypes.POINTER(cd)]
        for name, n in [("orc_bcglobalvc", 6), ("orc_bcglobalp", 2), ("orc_grad_lsm", 4), ("orc_calcul_face", 7),
                        ("orc_advdiff", 7), ("orc_sgles", 3), ("orc_source3d", 4), ("orc_update_rhsp", 3),
                        ("orc_bicgstab", 3), ("orc_newu3d", 2), ("orc_interpolate_u", 3),
                        ("orc_bcpcenter", 1)]:
            f = getattr(L, name)
            f.argtypes = [vp] + [vp] * n + ([ci] if name in ("orc_bcglobalvc", "orc_bcglobalp") else [])
        L.orc_glob_cfl.argtypes = [vp, vp, vp, vp, vp]
        L.orc_bicgstab.restype = ci
        L.orc_hydro.restype = ci
        L.orc_step.restype = ci
        L.orc_threads.restype = ci
        L.orc_sm_init.argtypes = [vp]
        L.orc_glob_sample.argtypes = [vp, vp, vp, vp, vp]
        L.orc_sm_planes.argtypes = [vp, vp]
        L.orc_sm_planes.restype = ci
        _libs[path] = L
    return _libs[path]


_BORROW = ["No_cp", "No_vp", "nbe", "nbev", "areaCell", "dlVV", "xnn", "ynn", "dxCC", "dyCC", "dlCC", "doe",
           "dle1", "dle2", "sum_xc2", "sum_yc2", "sum_xcyc", "surrounding", "Dimsurrounding", "vertex_pair",
           "vertex_edge", "weight", "dlVsum", "bcuc", "bcvc", "bcwc", "bcpc", "bcuv", "bcvv", "bcwv", "bcpv",
           "AmC0", "AmC1", "AmC2", "AmC3", "AmCT", "AmCB", "sig", "dsig", "sigv", "dsigv", "walld"]


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


class Oracle:
    """The reference's step on the CPU, holding its own state (reference layout)."""

    def __init__(self, case, *, p0=None, parallel: bool = False):
        self.L = L = lib(parallel)
        self.case = case
        c = case
        self.h = L.orc_create(c.N_CELL0, c.N_CELL, c.N_VERT, c.NZ, c.N_CELLexact)
        self._keep = {}
        for name in _BORROW:
            a = np.asfortranarray(c.arrays[name])
            if a.dtype.kind == "i":
                a = a.astype(np.int32, order="F")
            self._keep[name] = a
            if L.orc_set_array(self.h, name.encode(), _ptr(a)) != 0:
                raise KeyError(name)
        p = c.params
        L.orc_set_params(self.h, p.dt, p.Re, p.bdx, p.bdy, p.bdz, p.eps, p.MaxIters, p.FinRK, p.Iturbu,
                         p.TopBC, p.BotBC, c.bctop, c.bcbot, c.TotalVolume, p.sigIni, p.sigFin)
        for nm, src in (("ufn", c.u0), ("vfn", c.v0), ("wfn", c.w0), ("pfn", c.p0 if p0 is None else p0)):
            self.state(nm)[:] = src
        L.orc_init_state(self.h)

    def __del__(self):
        if getattr(self, "h", None) and getattr(self, "L", None) is not None:
            self.L.orc_free(self.h)
            self.h = None

    def threads(self) -> int:
        return self.L.orc_threads()

    def state(self, name: str) -> np.ndarray:
        """Live view of a state array, shaped like the Fortran array."""
        c = self.case
        ptr = self.L.orc_state(self.h, name.encode())
        if not ptr:
            raise KeyError(name)
        if name in ("ufv", "vfv", "wfv", "pfv"):
            shape = (c.N_VERT, c.NZ - 1)
        elif name in ("U1", "U2", "U3", "UT", "UB"):
            shape = (c.N_CELL0, c.NZ)
        else:
            shape = (c.N_CELL, c.NZ)
        n = shape[0] * shape[1]
        return np.ctypeslib.as_array(ptr, shape=(n,)).reshape(shape, order="F")

    def step(self) -> int:
        return self.L.orc_step(self.h)

    def hydro(self) -> int:
        return self.L.orc_hydro(self.h)

    def solver_record(self) -> dict:
        iv = (ctypes.c_int * 4)()
        dv = (ctypes.c_double * 2)()
        self.L.orc_solver_record(self.h, iv, dv)
        return dict(iter=iv[0], n_solves=iv[1], iter_total=iv[2], exit=iv[3], e1=dv[0], e2=dv[1])

    def timers(self) -> dict:
        """wall seconds spent in bicgstab and in hydro (solver included)"""
        dv = (ctypes.c_double * 2)()
        self.L.orc_timers(self.h, dv)
        return dict(solve=dv[0], hydro=dv[1])

    def set_max_iters(self, n: int) -> None:
        c = self.case
        p = c.params
        self.L.orc_set_params(self.h, p.dt, p.Re, p.bdx, p.bdy, p.bdz, p.eps, n, p.FinRK, p.Iturbu, p.TopBC,
                             p.BotBC, c.bctop, c.bcbot, c.TotalVolume, p.sigIni, p.sigFin)

    # -- single routines on caller arrays (reference layout, modified in place)
    def call(self, name: str, *arrays, tagdir=None):
        f = getattr(self.L, "orc_" + name)
        want = len(f.argtypes) - 1 - (1 if name in ("bcglobalvc", "bcglobalp") else 0)
        if len(arrays) != want:
            # the C routine would read (and write) the wrong buffers
            raise TypeError(f"orc_{name} takes {want} arrays, got {len(arrays)}")
        args = [self.h] + [_ptr(a) for a in arrays]
        if tagdir is not None:
            args.append(tagdir)
        return getattr(self.L, "orc_" + name)(*args)

    CFL_KEYS = ("umax", "vmax", "wmax", "cfl", "aveu", "avev", "avew", "flux_total", "flux_max", "flux_bc",
                "flux_bot", "flux_top", "flux_side")

    def glob_cfl(self, u=None, v=None, w=None) -> dict:
        """cfl.F90:84-200 on (u, v, w) (default: the state at time level n)
        and the resident face velocities."""
        u = self.state("ufn") if u is None else np.asfortranarray(u)
        v = self.state("vfn") if v is None else np.asfortranarray(v)
        w = self.state("wfn") if w is None else np.asfortranarray(w)
        out = np.zeros(13)
        self.L.orc_glob_cfl(self.h, _ptr(u), _ptr(v), _ptr(w), _ptr(out))
        return dict(zip(self.CFL_KEYS, out.tolist()))

    # -- channel statistics (stats.F90)
    def sm_init(self) -> None:
        self.L.orc_sm_init(self.h)

    def glob_sample(self, u=None, v=None, w=None, p=None) -> None:
        """glob_sample on (u, v, w, p) (default: the state at time level n)"""
        arrs = [self.state(n) if a is None else np.asfortranarray(a)
                for n, a in (("ufn", u), ("vfn", v), ("wfn", w), ("pfn", p))]
        self.L.orc_glob_sample(self.h, *[_ptr(a) for a in arrs])

    def sm_planes(self):
        """(planes (9, NZ), ns): the plane sums of glob_put_plane_mm"""
        out = np.zeros(9 * self.case.NZ)
        ns = self.L.orc_sm_planes(self.h, _ptr(out))
        return out.reshape(9, self.case.NZ), ns

    def cell(self):
        return self.case.cell_field()

    def vert(self):
        return self.case.vert_field()

    def face(self):
        c = self.case
        return np.zeros((c.N_CELL0, c.NZ), order="F")
