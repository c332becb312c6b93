"""ctypes declarations of the C ABI in include/nsx.h."""
from __future__ import annotations

import ctypes

c_int = ctypes.c_int
c_dbl = ctypes.c_double
c_vp = ctypes.c_void_p
c_cp = ctypes.c_char_p

# field ids (enum nsx_field)
F = dict(UFN=0, VFN=1, WFN=2, PFN=3, UF=4, VF=5, WF=6, PF=7, UFNP=8, VFNP=9, WFNP=10, PFNP=11,
         RHSUF=12, RHSVF=13, RHSWF=14, EDDY=15, FORCE=16, UFV=32, VFV=33, WFV=34, PFV=35,
         U1FACE=48, U2FACE=49, U3FACE=50, UTFACE=51, UBFACE=52)
CELL_SCRATCH = 64
VERT_SCRATCH = 96
FACE_SCRATCH = 112
N_CELL_SCRATCH = 16
N_VERT_SCRATCH = 8
N_FACE_SCRATCH = 10

PREC_FP64 = 0
PREC_MIXED = 1
PREC_FP64_CG = 2
PREC_MIXED_FLUX32 = 3
# initial guess of nsx_step's pressure solves in the precision forms
GUESS_PREVIOUS = 0
GUESS_EXTRAPOLATE = 1
REFINE_CG = 3
REFINE_CGSR = 4
REFINE_CGSR_BLOCK = 5
REFINE_CGSR_LINE = 6
REFINE_MODAL = 7

SOLVE_EXIT = {0: "converged", 1: "no solve", 2: "breakdown", 3: "max iterations"}

# symbols the header declares (checked by the CPU tests)
EXPORTS = [
    "nsx_create", "nsx_destroy", "nsx_last_error", "nsx_device", "nsx_set_mesh", "nsx_set_geometry",
    "nsx_set_sigma", "nsx_set_vertex_weights", "nsx_set_bc", "nsx_set_ppe_matrix", "nsx_set_les",
    "nsx_set_params", "nsx_commit", "nsx_put_field", "nsx_get_field", "nsx_copy_field", "nsx_init_state",
    "nsx_hydro", "nsx_step", "nsx_solve_info_get", "nsx_bcglobalvc", "nsx_bcglobalp", "nsx_grad_lsm",
    "nsx_calcul_face", "nsx_sgles", "nsx_advdiff", "nsx_advdiff_gam", "nsx_interpolate_u", "nsx_source3d",
    "nsx_update_rhsp", "nsx_bicgstab", "nsx_bicgstab_solve",
    "nsx_newu3d", "nsx_bicgstab_adjoint", "nsx_set_body_force", "nsx_wall_shear_run", "nsx_wall_shear_tangent",
    "nsx_step_tangent", "nsx_step_adjoint", "nsx_glob_cfl", "nsx_glob_cfl_noalloc", "nsx_set_precision", "nsx_set_initial_guess", "nsx_timing_enable", "nsx_timing_get", "nsx_timing_reset",
    "nsx_sm_init", "nsx_glob_sample", "nsx_sm_count", "nsx_sm_get", "nsx_sm_planes", "nsx_glob_put_plane_mm",
    "nsx_write_plane_stats", "nsx_format_real", "nsx_savetec_vertex", "nsx_savetec_center", "nsx_write_tec_vertex",
    "nsx_write_tec_center", "nsx_save_paraview_vertex", "nsx_savetec_vc", "nsx_write_vtk_vertex", "nsx_write_tec_vc",
]
# the accumulators of module stats (nsx_sm_get)
SM_NAMES = ["sm_uf", "sm_vf", "sm_wf", "sm_u2f", "sm_v2f", "sm_w2f", "sm_uvf", "sm_vwf", "sm_wuf", "sm_pf", "sm_p2f",
            "sm_ef"]


class SolveInfo(ctypes.Structure):
    _fields_ = [("iters", c_int), ("exit_code", c_int), ("e1norm", c_dbl), ("e2norm", c_dbl),
                ("n_solves", c_int), ("iters_total", c_int), ("outer_passes", c_int),
                ("iters_fp32", c_int)]

    def as_dict(self):
        return dict(iters=self.iters, exit=SOLVE_EXIT.get(self.exit_code, str(self.exit_code)), e1=self.e1norm,
                    e2=self.e2norm, n_solves=self.n_solves, iters_total=self.iters_total,
                    outer_passes=self.outer_passes, iters_fp32=self.iters_fp32)


class CflInfo(ctypes.Structure):
    _fields_ = [(n, c_dbl) for n in ("umax", "vmax", "wmax", "cfl", "aveu", "avev", "avew", "flux_total",
                                      "flux_max", "flux_bc", "flux_bot", "flux_top", "flux_side")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


def declare(lib: ctypes.CDLL) -> None:
    P = ctypes.POINTER
    sig = {
        "nsx_create": [c_int, c_int, c_int, c_int, P(c_vp)],
        "nsx_destroy": [c_vp],
        "nsx_device": [c_vp],
        "nsx_set_mesh": [c_vp, c_vp, c_vp, c_vp, c_vp],
        "nsx_set_geometry": [c_vp] + [c_vp] * 13 + [c_dbl],
        "nsx_set_sigma": [c_vp, c_vp, c_vp, c_vp, c_vp, c_dbl, c_dbl],
        "nsx_set_vertex_weights": [c_vp] * 7,
        "nsx_set_bc": [c_vp] + [c_vp] * 8 + [c_int] * 4,
        "nsx_set_ppe_matrix": [c_vp] * 7,
        "nsx_set_les": [c_vp, c_vp],
        "nsx_set_params": [c_vp, c_dbl, c_dbl, c_dbl, c_dbl, c_dbl, c_dbl, c_int, c_int, c_int],
        "nsx_commit": [c_vp],
        "nsx_put_field": [c_vp, c_int, c_vp],
        "nsx_get_field": [c_vp, c_int, c_vp, c_int],
        "nsx_copy_field": [c_vp, c_int, c_int],
        "nsx_init_state": [c_vp],
        "nsx_hydro": [c_vp, c_int, P(SolveInfo)],
        "nsx_step": [c_vp, c_int, c_int, P(SolveInfo)],
        "nsx_solve_info_get": [c_vp, P(SolveInfo)],
        "nsx_bcglobalvc": [c_vp] + [c_int] * 7,
        "nsx_bcglobalp": [c_vp, c_int, c_int, c_int],
        "nsx_grad_lsm": [c_vp] + [c_int] * 4,
        "nsx_calcul_face": [c_vp] + [c_int] * 7,
        "nsx_sgles": [c_vp] + [c_int] * 3,
        "nsx_advdiff": [c_vp] + [c_int] * 7,
        "nsx_advdiff_gam": [c_vp] + [c_int] * 9,
        "nsx_interpolate_u": [c_vp] + [c_int] * 8,
        "nsx_bicgstab_solve": [c_vp, c_int, c_int, P(SolveInfo)],
        "nsx_source3d": [c_vp] + [c_int] * 4,
        "nsx_update_rhsp": [c_vp] + [c_int] * 3,
        "nsx_bicgstab": [c_vp, c_int, c_int, c_int, P(SolveInfo)],
        "nsx_newu3d": [c_vp, c_int, c_int],
        "nsx_set_precision": [c_vp, c_int, c_int, c_dbl, c_int],
        "nsx_set_initial_guess": [c_vp, c_int],
        "nsx_glob_cfl": [c_vp, c_int, c_int, c_int, P(CflInfo)],
        "nsx_glob_cfl_noalloc": [c_vp, c_int, c_int, c_int, P(CflInfo)],
        "nsx_bicgstab_adjoint": [c_vp, c_int, c_int, P(SolveInfo)],
        "nsx_set_body_force": [c_vp, c_int],
        "nsx_wall_shear_run": [c_vp, c_int, c_int, P(c_dbl), c_int],
        "nsx_wall_shear_tangent": [c_vp, c_int, c_int, P(c_dbl), P(c_dbl)],
        "nsx_step_tangent": [c_vp, P(c_int), P(c_int), c_int],
        "nsx_step_adjoint": [c_vp, P(c_int), P(c_int), c_int],
        "nsx_timing_enable": [c_vp, c_int],
        "nsx_timing_get": [c_vp, c_cp, P(c_dbl), P(ctypes.c_longlong)],
        "nsx_timing_reset": [c_vp],
        "nsx_sm_init": [c_vp],
        "nsx_glob_sample": [c_vp, c_int, c_int, c_int, c_int],
        "nsx_sm_count": [c_vp, P(c_int)],
        "nsx_sm_get": [c_vp, c_int, c_vp],
        "nsx_sm_planes": [c_vp, c_vp, P(c_int)],
        "nsx_glob_put_plane_mm": [c_vp, c_cp],
        "nsx_write_plane_stats": [c_vp, c_int, c_int, c_int, c_dbl, c_cp],
        "nsx_format_real": [c_dbl, c_int, c_int, c_int, c_cp, c_int],
        "nsx_savetec_vertex": [c_vp, c_cp, c_int, c_dbl, c_vp, c_vp, c_vp],
        "nsx_savetec_center": [c_vp, c_cp, c_int, c_dbl, c_vp, c_vp, c_vp, c_int, c_int, c_int, c_int],
        "nsx_write_tec_vertex": [c_cp, c_int, c_dbl, c_int, c_int, c_int] + [c_vp] * 8,
        "nsx_write_tec_center": [c_cp, c_int, c_dbl, c_int, c_int, c_int, c_int] + [c_vp] * 8,
        "nsx_save_paraview_vertex": [c_vp, c_cp, c_int, c_vp, c_vp, c_vp],
        "nsx_savetec_vc": [c_vp, c_cp, c_int, c_dbl] + [c_vp] * 6 + [c_int] * 4,
        "nsx_write_vtk_vertex": [c_cp, c_int, c_int, c_int, c_int] + [c_vp] * 10,
        "nsx_write_tec_vc": [c_cp, c_int, c_dbl, c_int, c_int, c_int, c_int] + [c_vp] * 15,
    }
    for name, args in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = c_int
    lib.nsx_last_error.argtypes = [c_vp]
    lib.nsx_last_error.restype = c_cp
