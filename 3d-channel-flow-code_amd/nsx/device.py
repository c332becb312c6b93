"""Host-side mirror of the reference's hot-path interface over the HIP library.

`Context` owns one device context (include/nsx.h).  Its methods carry the
names and argument meaning of the reference's Fortran routines
(src/Main/hydro.F90, bicgstab.F90, src/BoundaryConditions/BC.F90, ...): they
take and return numpy arrays in the reference layout ((N_CELL,NZ),
(N_VERT,NZ-1), (N_CELL0,NZ) Fortran order).  Field state otherwise stays
resident on the device between calls (`step`, `hydro`).  Every failure of the
library raises NsxError: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .abi import CELL_SCRATCH, F, FACE_SCRATCH, PREC_FP64, PREC_MIXED, VERT_SCRATCH, CflInfo, SolveInfo  # noqa: F401


class NsxError(RuntimeError):
    pass


def _p(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


def _f64(a) -> np.ndarray:
    return np.asfortranarray(a, dtype=np.float64)


def _i32(a) -> np.ndarray:
    return np.asfortranarray(a, dtype=np.int32)


class Context:
    """A device context holding one case (mesh, geometry, matrix, state)."""

    def __init__(self, case):
        self.lib = _lib.device()
        self.case = c = case
        h = ctypes.c_void_p()
        rc = self.lib.nsx_create(c.N_CELL0, c.N_CELL, c.N_VERT, c.NZ, ctypes.byref(h))
        if rc != 0:
            raise NsxError(f"nsx_create failed ({rc}): no usable HIP device")
        self.h = h
        A = c.arrays
        keep = []

        def k(a, conv=_f64):
            a = conv(a)
            keep.append(a)
            return _p(a)

        self._check(self.lib.nsx_set_mesh(h, k(A["No_cp"], _i32), k(A["No_vp"], _i32), k(A["nbe"], _i32),
                                          k(A["nbev"], _i32)))
        self._check(self.lib.nsx_set_geometry(h, *[k(A[n]) for n in ("areaCell", "dlVV", "dlCC", "dxCC", "dyCC",
                                                                      "xnn", "ynn", "dle1", "dle2", "doe",
                                                                      "sum_xc2", "sum_yc2", "sum_xcyc")],
                                              c.TotalVolume))
        p = c.params
        self._check(self.lib.nsx_set_sigma(h, k(A["sig"]), k(A["dsig"]), k(A["sigv"]), k(A["dsigv"]), p.sigIni,
                                           p.sigFin))
        self._check(self.lib.nsx_set_vertex_weights(h, k(A["surrounding"], _i32), k(A["weight"]),
                                                    k(A["Dimsurrounding"], _i32), k(A["dlVsum"]),
                                                    k(A["vertex_pair"], _i32), k(A["vertex_edge"], _i32)))
        self._check(self.lib.nsx_set_bc(h, *[k(A[n], _i32) for n in ("bcuc", "bcvc", "bcwc", "bcpc", "bcuv", "bcvv",
                                                                      "bcwv", "bcpv")],
                                        c.bctop, c.bcbot, p.TopBC, p.BotBC))
        self._check(self.lib.nsx_set_ppe_matrix(h, *[k(A[n]) for n in ("AmC0", "AmC1", "AmC2", "AmC3", "AmCT",
                                                                        "AmCB")]))
        self._check(self.lib.nsx_set_les(h, k(A["walld"])))
        self._check(self.lib.nsx_set_params(h, p.dt, p.Re, p.bdx, p.bdy, p.bdz, p.eps, p.MaxIters, p.FinRK,
                                            p.Iturbu))
        self._check(self.lib.nsx_commit(h))

    # ------------------------------------------------------------ basics ---
    def _check(self, rc):
        if rc < 0:
            raise NsxError(self.lib.nsx_last_error(self.h).decode(errors="replace"))
        return rc

    def close(self):
        if getattr(self, "h", None):
            self.lib.nsx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _shape(self, fid: int):
        c = self.case
        if 32 <= fid < 48 or VERT_SCRATCH <= fid < VERT_SCRATCH + 8:
            return (c.N_VERT, c.NZ - 1)
        if 48 <= fid < 64 or FACE_SCRATCH <= fid < FACE_SCRATCH + 10:
            return (c.N_CELL0, c.NZ)
        return (c.N_CELL, c.NZ)

    def put(self, field, a) -> None:
        fid = F[field] if isinstance(field, str) else field
        a = _f64(a)
        if a.shape != self._shape(fid):
            raise ValueError(f"field {field}: shape {a.shape}, expected {self._shape(fid)}")
        self._check(self.lib.nsx_put_field(self.h, fid, _p(a)))

    def get(self, field, fill_ghost_rows: bool = True) -> np.ndarray:
        fid = F[field] if isinstance(field, str) else field
        out = np.zeros(self._shape(fid), order="F")
        self._check(self.lib.nsx_get_field(self.h, fid, _p(out), int(fill_ghost_rows)))
        return out

    def device_memory_used(self) -> int:
        """Bytes of device memory in use on the context's GPU (hipMemGetInfo)."""
        import os
        # the ROCm runtime the library links (a torch wheel may carry another
        # copy of it under the same name)
        hip = ctypes.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
        free, total = ctypes.c_size_t(), ctypes.c_size_t()
        if hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) != 0:
            raise NsxError("hipMemGetInfo failed")
        return total.value - free.value

    def solve_info(self) -> dict:
        si = SolveInfo()
        self._check(self.lib.nsx_solve_info_get(self.h, ctypes.byref(si)))
        return si.as_dict()

    # ---------------------------------------------------------- stepping ---
    def init_state(self, u0=None, v0=None, w0=None, p0=None) -> None:
        """user_initial.F90:229-253 + LoopTimeInitial with the case's initial field."""
        c = self.case
        for name, a in (("UFN", c.u0 if u0 is None else u0), ("VFN", c.v0 if v0 is None else v0),
                        ("WFN", c.w0 if w0 is None else w0), ("PFN", c.p0 if p0 is None else p0)):
            self.put(name, a)
        self._check(self.lib.nsx_init_state(self.h))

    def step(self, nsteps: int = 1, precision: int = PREC_FP64) -> dict:
        """nsteps time steps (nsmp3D.F90:404-564) on the resident state."""
        si = SolveInfo()
        self._check(self.lib.nsx_step(self.h, nsteps, precision, ctypes.byref(si)))
        return si.as_dict()

    def hydro(self, flag_ab: int = 0) -> dict:
        """One stage from (UF..PF) into (UFNP..PFNP) (hydro.F90:8-501)."""
        si = SolveInfo()
        self._check(self.lib.nsx_hydro(self.h, flag_ab, ctypes.byref(si)))
        return si.as_dict()

    def copy(self, dst, src) -> None:
        self._check(self.lib.nsx_copy_field(self.h, F[dst] if isinstance(dst, str) else dst,
                                            F[src] if isinstance(src, str) else src))

    # ---------------------------------------- single routines (reference names)
    def _cells(self, *arrays):
        ids = []
        for n, a in enumerate(arrays):
            fid = CELL_SCRATCH + n
            self.put(fid, a)
            ids.append(fid)
        return ids

    def BCglobalVC(self, u, v, w, tagdir: int):
        """BC.F90:202-295: returns (u, v, w) with ghosts filled and (uv, vv, wv)."""
        cu, cv, cw = self._cells(u, v, w)
        self._check(self.lib.nsx_bcglobalvc(self.h, cu, cv, cw, VERT_SCRATCH, VERT_SCRATCH + 1, VERT_SCRATCH + 2,
                                            tagdir))
        return (self.get(cu), self.get(cv), self.get(cw), self.get(VERT_SCRATCH), self.get(VERT_SCRATCH + 1),
                self.get(VERT_SCRATCH + 2))

    def BCglobalP(self, p, tagdir: int):
        """BC.F90:299-363: returns (p with ghosts, pv)."""
        (cp,) = self._cells(p)
        self._check(self.lib.nsx_bcglobalp(self.h, cp, VERT_SCRATCH, tagdir))
        return self.get(cp), self.get(VERT_SCRATCH)

    def grandientLSM(self, f):
        """gradientLSM.F90:8-211: returns (dfdx, dfdy, dfdsig)."""
        (cf,) = self._cells(f)
        o = CELL_SCRATCH + 4
        self._check(self.lib.nsx_grad_lsm(self.h, o, o + 1, o + 2, cf))
        return self.get(o), self.get(o + 1), self.get(o + 2)

    def calcul_face(self, f, fv):
        """interpolationFace.F90:7-238: the five face values of f."""
        (cf,) = self._cells(f)
        self.put(VERT_SCRATCH, fv)
        o = FACE_SCRATCH
        self._check(self.lib.nsx_calcul_face(self.h, o, o + 1, o + 2, o + 3, o + 4, cf, VERT_SCRATCH))
        return tuple(self.get(o + q) for q in range(5))

    def sgles(self, u, v, w):
        """sgm_les.F90:24-275: eddy viscosity of (u,v,w)."""
        cu, cv, cw = self._cells(u, v, w)
        self._check(self.lib.nsx_sgles(self.h, cu, cv, cw))
        return self.get("EDDY")

    def AdvDiffNS(self, rhsu, rhsv, rhsw, u, v, w, eddy):
        """AdvDiffNS.F90:8-140 with the resident face velocities; returns the three rhs."""
        ids = self._cells(rhsu, rhsv, rhsw, u, v, w, eddy)
        self._check(self.lib.nsx_advdiff(self.h, *ids))
        return self.get(ids[0]), self.get(ids[1]), self.get(ids[2])

    def Source3D(self, u, v, w):
        """source3D.F90:8-174: returns (src, U1FACE..UBFACE); the faces stay resident."""
        cs, cu, cv, cw = self._cells(np.zeros_like(u), u, v, w)
        self._check(self.lib.nsx_source3d(self.h, cs, cu, cv, cw))
        return (self.get(cs, fill_ghost_rows=False),) + tuple(self.get(n) for n in ("U1FACE", "U2FACE", "U3FACE",
                                                                                      "UTFACE", "UBFACE"))

    def update_rhsp(self, rhs, p):
        """source3D.F90:176-335: returns (rhs, pv)."""
        cr, cp = self._cells(rhs, p)
        self._check(self.lib.nsx_update_rhsp(self.h, cr, cp, VERT_SCRATCH))
        return self.get(cr, fill_ghost_rows=False), self.get(VERT_SCRATCH)

    def bicgstab(self, phi, rhs):
        """bicgstab.F90:8-168: returns (phi, phiv, solve info)."""
        cphi, crhs = self._cells(phi, rhs)
        si = SolveInfo()
        self._check(self.lib.nsx_bicgstab(self.h, cphi, VERT_SCRATCH, crhs, ctypes.byref(si)))
        return self.get(cphi), self.get(VERT_SCRATCH), si.as_dict()

    def newu3D(self, p):
        """newu3D.F90:8-257 on the resident faces: returns (faces, pv)."""
        (cp,) = self._cells(p)
        self._check(self.lib.nsx_newu3d(self.h, cp, VERT_SCRATCH))
        return tuple(self.get(n) for n in ("U1FACE", "U2FACE", "U3FACE", "UTFACE", "UBFACE")), self.get(VERT_SCRATCH)

    # -------------------------------------------------- differentiation ---
    def set_body_force(self, force=None) -> None:
        """The per-prism streamwise body force in place of bdx (an (N_CELL,NZ)
        array, or interior (N_CELL0, NZ-2)), or None to go back to bdx."""
        if force is None:
            self._check(self.lib.nsx_set_body_force(self.h, 0))
            return
        self.put("FORCE", self._full_cell(force))
        self._check(self.lib.nsx_set_body_force(self.h, 1))

    def _full_cell(self, a) -> np.ndarray:
        c = self.case
        a = np.asarray(a, dtype=np.float64)
        if a.shape == (c.N_CELL0, c.NZ - 2):
            full = c.cell_field()
            full[:c.N_CELL0, 1:-1] = a
            return full
        return a

    def wall_shear(self, nsteps: int, every: int = 5, grad: bool = False):
        """nsteps RK2 steps from the resident state and the time-averaged
        bottom-wall shear J; with grad, also dJ/d(body force) by the adjoint
        (interior prisms, (N_CELL0, NZ-2)).  Returns J or (J, grad)."""
        J = ctypes.c_double()
        g = CELL_SCRATCH + 15 if grad else -1
        rc = self._check(self.lib.nsx_wall_shear_run(self.h, int(nsteps), int(every), ctypes.byref(J), g))
        if rc:
            raise NsxError("nsx_wall_shear_run: a pressure solve reached MaxIters")
        if not grad:
            return J.value
        return J.value, self.get(g, fill_ghost_rows=False)[:self.case.N_CELL0, 1:-1].copy()

    def wall_shear_tangent(self, nsteps: int, direction):
        """The same steps with the tangent-linear model along a body-force
        direction: (J, dJ)."""
        d = CELL_SCRATCH + 14
        self.put(d, self._full_cell(direction))
        J, dJ = ctypes.c_double(), ctypes.c_double()
        rc = self._check(self.lib.nsx_wall_shear_tangent(self.h, int(nsteps), d, ctypes.byref(J), ctypes.byref(dJ)))
        if rc:
            raise NsxError("nsx_wall_shear_tangent: a pressure solve reached MaxIters")
        return J.value, dJ.value

    def _state_ids(self, base_cell: int, base_face: int):
        ids = (ctypes.c_int * 9)(*([base_cell + q for q in range(4)] + [base_face + q for q in range(5)]))
        return ids

    def step_tangent(self, tin, dforce=None):
        """One RK2 step's tangent-linear map at the resident state (left
        unchanged): tin = 9 arrays (u, v, w, p as (N_CELL,NZ), the five faces
        as (N_CELL0,NZ)); returns the 9 tangents of the new level."""
        ti, to = self._state_ids(CELL_SCRATCH, FACE_SCRATCH), self._state_ids(CELL_SCRATCH + 4, FACE_SCRATCH + 5)
        for q in range(9):
            self.put(ti[q], tin[q])
        df = -1
        if dforce is not None:
            df = CELL_SCRATCH + 8
            self.put(df, self._full_cell(dforce))
        self._check(self.lib.nsx_step_tangent(self.h, ti, to, df))
        return [self.get(to[q], fill_ghost_rows=False) for q in range(9)]

    def step_adjoint(self, bin_):
        """The transpose of step_tangent: the bars of the new level (9 arrays)
        give those of level n (9 arrays) and the body-force bar."""
        bi, bo = self._state_ids(CELL_SCRATCH, FACE_SCRATCH), self._state_ids(CELL_SCRATCH + 4, FACE_SCRATCH + 5)
        for q in range(9):
            self.put(bi[q], bin_[q])
        gf = CELL_SCRATCH + 8
        self._check(self.lib.nsx_step_adjoint(self.h, bi, bo, gf))
        return [self.get(bo[q], fill_ghost_rows=False) for q in range(9)], self.get(gf, fill_ghost_rows=False)

    # ------------------------------------------------------------- timing ---
    def bicgstab_adjoint(self, g):
        """Adjoint of bicgstab() with respect to its right-hand side: returns
        (lam, solve info) with <g, d phi> = <lam, d rhs>."""
        (cg,) = self._cells(g)
        lam = CELL_SCRATCH + 1
        si = SolveInfo()
        self._check(self.lib.nsx_bicgstab_adjoint(self.h, lam, cg, ctypes.byref(si)))
        return self.get(lam), si.as_dict()

    def glob_cfl(self, u=None, v=None, w=None) -> dict:
        """glob_cfl (cfl.F90:7-269) on device: resident UFN, VFN, WFN by
        default, or the given (N_CELL,NZ) arrays, with the resident faces."""
        if u is None:
            ids = (F["UFN"], F["VFN"], F["WFN"])
        else:
            ids = self._cells(u, v, w)
        out = CflInfo()
        self._check(self.lib.nsx_glob_cfl(self.h, *ids, ctypes.byref(out)))
        return out.as_dict()

    def glob_cfl_noalloc(self, u=None, v=None, w=None) -> dict:
        """glob_cfl through nsx_glob_cfl_noalloc: the same values, with the
        partial sums in the library's static per-device scratch."""
        if u is None:
            ids = (F["UFN"], F["VFN"], F["WFN"])
        else:
            ids = self._cells(u, v, w)
        out = CflInfo()
        self._check(self.lib.nsx_glob_cfl_noalloc(self.h, *ids, ctypes.byref(out)))
        return out.as_dict()

    # -- channel statistics (stats.F90)
    def sm_init(self) -> None:
        """sm_init (stats.F90:34-84) on the device accumulators."""
        self._check(self.lib.nsx_sm_init(self.h))

    def glob_sample(self, u=None, v=None, w=None, p=None) -> None:
        """glob_sample (stats.F90:91-197): resident UFN, VFN, WFN, PFN by
        default, or the given (N_CELL,NZ) arrays; the eddy viscosity term is
        the resident NSX_F_EDDY."""
        if u is None:
            ids = (F["UFN"], F["VFN"], F["WFN"], F["PFN"])
        else:
            ids = self._cells(u, v, w, p)
        self._check(self.lib.nsx_glob_sample(self.h, *ids))

    def sm_count(self) -> int:
        n = ctypes.c_int()
        self._check(self.lib.nsx_sm_count(self.h, ctypes.byref(n)))
        return n.value

    def sm_get(self, name: str) -> np.ndarray:
        """One accumulator (sm_uf .. sm_ef) as (N_CELL,NZ); rows past N_CELL0 are 0."""
        from .abi import SM_NAMES
        out = np.zeros((self.case.N_CELL, self.case.NZ), order="F")
        self._check(self.lib.nsx_sm_get(self.h, SM_NAMES.index(name), out.ctypes.data))
        return out

    def sm_planes(self):
        """(planes (9, NZ), ns): glob_put_plane_mm's plane sums (stats.F90:266-279)."""
        out = np.zeros(9 * self.case.NZ)
        ns = ctypes.c_int()
        self._check(self.lib.nsx_sm_planes(self.h, out.ctypes.data, ctypes.byref(ns)))
        return out.reshape(9, self.case.NZ), ns.value

    def glob_put_plane_mm(self, directory: str) -> None:
        """glob_put_plane_mm (stats.F90:204-391): the three statistics files in directory."""
        self._check(self.lib.nsx_glob_put_plane_mm(self.h, directory.encode()))

    # -- Tecplot snapshots (src/Save)
    def _coords(self, xvt, yvt, zvt):
        shape = (self.case.N_VERT, self.case.NZ - 1)
        out = []
        for a in (xvt, yvt, zvt):
            a = np.asfortranarray(a, dtype=np.float64)
            if a.shape != shape:
                raise ValueError(f"vertex coordinates must be {shape}, got {a.shape}")
            out.append(a)
        return out

    def savetec_vertex(self, directory: str, counter: int, time: float, xvt, yvt, zvt) -> None:
        """SavetecVertex (SavetecVertex.F90:185-235): directory/V-NNNN.tec of
        the resident vertex fields, with the vertex coordinates xvt, yvt, zvt
        (N_VERT, NZ-1)."""
        x, y, z = self._coords(xvt, yvt, zvt)
        self._check(self.lib.nsx_savetec_vertex(self.h, directory.encode(), int(counter), float(time),
                                                x.ctypes.data, y.ctypes.data, z.ctypes.data))

    def savetec_center(self, directory: str, counter: int, time: float, xvt, yvt, zvt,
                       fields=("UFNP", "VFNP", "WFNP", "PFNP")) -> None:
        """SavetecCenter (SavetecCenter.F90:185-235): directory/C-NNNNN.tec of
        the resident cell fields named in fields (u, v, w, p)."""
        x, y, z = self._coords(xvt, yvt, zvt)
        ids = [F[n] if isinstance(n, str) else int(n) for n in fields]
        self._check(self.lib.nsx_savetec_center(self.h, directory.encode(), int(counter), float(time),
                                                x.ctypes.data, y.ctypes.data, z.ctypes.data, *ids))

    def save_paraview_vertex(self, directory: str, counter: int, xvt, yvt, zvt) -> None:
        """SaveParaviewVertex (SaveParaviewVertex.F90:185-278): directory/V-NNNN.vtk
        of the resident vertex velocities."""
        x, y, z = self._coords(xvt, yvt, zvt)
        self._check(self.lib.nsx_save_paraview_vertex(self.h, directory.encode(), int(counter), x.ctypes.data,
                                                      y.ctypes.data, z.ctypes.data))

    def savetec_vc(self, directory: str, counter: int, time: float, xvt, yvt, zvt, xct, yct, zct,
                   fields=("UFNP", "VFNP", "WFNP", "PFNP")) -> None:
        """SavetecVC (SavetecVertexCenter.F90:212-305): directory/VC-NNNNN.tec of
        the resident vertex fields and the cell fields named in fields, with
        the cell-centre coordinates xct, yct, zct (N_CELL, NZ)."""
        x, y, z = self._coords(xvt, yvt, zvt)
        cs = []
        for a in (xct, yct, zct):
            a = np.asfortranarray(a, dtype=np.float64)
            if a.shape != (self.case.N_CELL, self.case.NZ):
                raise ValueError(f"cell coordinates must be {(self.case.N_CELL, self.case.NZ)}, got {a.shape}")
            cs.append(a)
        ids = [F[n] if isinstance(n, str) else int(n) for n in fields]
        self._check(self.lib.nsx_savetec_vc(self.h, directory.encode(), int(counter), float(time), x.ctypes.data,
                                            y.ctypes.data, z.ctypes.data, *[a.ctypes.data for a in cs], *ids))

    def set_max_iters(self, n: int) -> None:
        """MaxIters of the pressure solve (the case's other parameters kept)."""
        self.set_params(max_iters=n)

    def set_params(self, eps: float = None, max_iters: int = None, dt: float = None) -> None:
        """The tolerance eps and MaxIters of the pressure solve and the time
        step dt (nsx_set_params); the parameters not given keep the values
        last set (the case's to begin with)."""
        p = self.case.params
        cur = getattr(self, "_params", {"eps": p.eps, "max_iters": p.MaxIters, "dt": p.dt})
        if eps is not None:
            cur["eps"] = float(eps)
        if max_iters is not None:
            cur["max_iters"] = int(max_iters)
        if dt is not None:
            cur["dt"] = float(dt)
        self._params = cur
        self._check(self.lib.nsx_set_params(self.h, cur["dt"], p.Re, p.bdx, p.bdy, p.bdz, cur["eps"],
                                            cur["max_iters"], p.FinRK, p.Iturbu))

    def set_precision(self, precision: int, refine: int = 5, tol: float = 0.1, max_passes: int = 30) -> None:
        """Arithmetic of the pressure solve (PREC_FP64, or PREC_MIXED: FP32 CG
        with reliable updates against the FP64 solution: refine 5 =
        single-reduction CG with the 4-layer block-Jacobi preconditioner and
        updates at threshold tol (the default), 6 = the same with the line
        (whole column) preconditioner, 4 = the same with point Jacobi, 3 = the
        point-Jacobi CG in two kernels per iteration); step() takes its own
        precision and uses this refinement."""
        self._check(self.lib.nsx_set_precision(self.h, precision, refine, tol, max_passes))

    def set_initial_guess(self, mode: int) -> None:
        """Initial guess of step()'s pressure solves in the precision forms:
        GUESS_PREVIOUS (the stage's input pressure, the reference's) or
        GUESS_EXTRAPOLATE (2 p_k - p_{k-1} from the same RK stage's last two
        solutions)."""
        self._check(self.lib.nsx_set_initial_guess(self.h, mode))

    def timing(self, on: bool) -> None:
        self._check(self.lib.nsx_timing_enable(self.h, int(on)))

    def timing_reset(self) -> None:
        self._check(self.lib.nsx_timing_reset(self.h))

    def timer(self, name: str):
        ms = ctypes.c_double()
        n = ctypes.c_longlong()
        self._check(self.lib.nsx_timing_get(self.h, name.encode(), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value
