/*
 * nsx.h -- C ABI of the MI355X drop-in for the per-time-step hot path of the
 * NSMP3D prismatic sigma-grid solver (reference: techi28/3D-channel-flow-code).
 *
 * The reference's hot-path routines are Fortran external subroutines with
 * implicit interfaces, explicit-shape REAL*8 arrays laid out (N_CELL,NZ)
 * column-major with the triangle index fastest, and module state (USE
 * geometry / USE les).  This ABI replaces them with entry points over one
 * device context that keeps every field resident in HBM:
 *
 *   - set-up calls take the reference's own module arrays (host pointers,
 *     reference layout, 1-based connectivity) once;
 *   - nsx_put_field / nsx_get_field move a field between a host array in the
 *     reference layout and the device copy (explicit sync points: I/O,
 *     diagnostics);
 *   - hot-path calls take field ids (NSX_F_*) of device-resident fields.
 *
 * Every call returns 0 on success, > 0 for a soft solver outcome (see
 * NSX_SOLVE_*), < 0 for an error; nsx_last_error() gives the message.  One
 * host thread per context; calls are synchronous unless stated otherwise.
 * Each entry point cites the reference routine it replaces.  The Fortran
 * binding (ISO_C_BINDING) is 3d-channel-flow-code_amd/fortran/nsx_iface.F90.
 */
#ifndef NSX_H
#define NSX_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nsx_ctx nsx_ctx;

/* ---- field ids ------------------------------------------------------- */
enum nsx_field {
  /* cell fields (N_CELL,NZ) in the reference layout (module variables) */
  NSX_F_UFN = 0, NSX_F_VFN, NSX_F_WFN, NSX_F_PFN, /* time level n */
  NSX_F_UF, NSX_F_VF, NSX_F_WF, NSX_F_PF,         /* RK stage input */
  NSX_F_UFNP, NSX_F_VFNP, NSX_F_WFNP, NSX_F_PFNP, /* stage output / n+1 */
  NSX_F_RHSUF, NSX_F_RHSVF, NSX_F_RHSWF,          /* AB2 history (module geometry) */
  NSX_F_EDDY,                                     /* eddy viscosity (module les) */
  NSX_F_FORCE,                                    /* per-prism streamwise body force (nsx_set_body_force) */
  /* vertex fields (N_VERT,NZ-1) */
  NSX_F_UFV = 32, NSX_F_VFV, NSX_F_WFV, NSX_F_PFV,
  /* face-normal velocities (N_CELL0,NZ): U1FACE..UBFACE (module geometry) */
  NSX_F_U1FACE = 48, NSX_F_U2FACE, NSX_F_U3FACE, NSX_F_UTFACE, NSX_F_UBFACE,
  /* scratch fields for single-routine calls */
  NSX_F_CELL_SCRATCH = 64, /* 64..79: cell fields */
  NSX_F_VERT_SCRATCH = 96, /* 96..103: vertex fields */
  NSX_F_FACE_SCRATCH = 112 /* 112..121: face fields (N_CELL0,NZ) */
};
#define NSX_N_CELL_SCRATCH 16
#define NSX_N_VERT_SCRATCH 8
#define NSX_N_FACE_SCRATCH 10

/* ---- solver outcome (bicgstab.F90:462-575) --------------------------- */
enum nsx_solve_exit {
  NSX_SOLVE_CONVERGED = 0,   /* min(e1,e2) < eps after iterations */
  NSX_SOLVE_NO_SOLVE = 1,    /* "Pressure converged without solve" */
  NSX_SOLVE_BREAKDOWN = 2,   /* |rho|, |omega| or |<r0,v>| < 1e-30 */
  NSX_SOLVE_MAX_ITERS = 3    /* iter > MaxIters */
};

typedef struct nsx_solve_info {
  int iters;       /* BiCGSTAB iterations of the last solve */
  int exit_code;   /* enum nsx_solve_exit */
  double e1norm;   /* sum|r| / sum|b| */
  double e2norm;   /* ||r||_2 / ||b||_2 */
  int n_solves;    /* solves since nsx_init_state */
  int iters_total; /* iterations since nsx_init_state */
  int outer_passes; /* NSX_PREC_MIXED, NSX_PREC_FP64_CG: reliable updates of the last solve */
  int iters_fp32;   /* NSX_PREC_MIXED: FP32 iterations of the last solve (included in iters) */
} nsx_solve_info;

/* ---- precision of the step --------------------------------------------- */
enum nsx_precision {
  NSX_PREC_FP64 = 0,  /* the reference's arithmetic */
  NSX_PREC_MIXED = 1, /* the precision form: FP32 conjugate gradients (SpMV, preconditioner,
                          updates; by default the single-reduction form with the 4-layer
                          block-Jacobi preconditioner, or with NSX_REFINE_MODAL the
                          vertical-mode preconditioner) with reliable updates against the
                          FP64 solution; true residuals, inner products and norms in FP64 */
  NSX_PREC_FP64_CG = 2, /* the precision form's solver with FP64 vectors: single-reduction CG
                          with the 4-layer block-Jacobi preconditioner and reliable updates,
                          all in FP64 (the refinement set by nsx_set_precision is not used) */
  NSX_PREC_MIXED_FLUX32 = 3 /* NSX_PREC_MIXED with the momentum fluxes and the Smagorinsky eddy
                          viscosity in FP32 arithmetic too (stencil differences formed in FP64,
                          rounded once; fields stay FP64; the predictor adds the FP32 rate to
                          the FP64 velocity) */
};

/* ---- lifecycle ----------------------------------------------------------- */
/* Sizes of the reference's COMMON /coo/ (common.mpf:323-325). */
int nsx_create(int n_cell0, int n_cell, int n_vert, int nz, nsx_ctx **ctx);
int nsx_destroy(nsx_ctx *ctx);
const char *nsx_last_error(const nsx_ctx *ctx);
/* device id the context runs on (hipGetDevice at nsx_create) */
int nsx_device(const nsx_ctx *ctx);

/* ---- set-up: the reference's module arrays, host pointers -------------- */
/* module variables No_cp (N_CELL,3) incl. ghost rows, No_vp (N_CELL0,3),
 * nbe (N_CELL0), nbev (N_VERT); read by calcul_geometry.F90 / input.F90 */
int nsx_set_mesh(nsx_ctx *ctx, const int *No_cp, const int *No_vp, const int *nbe, const int *nbev);
/* module geometry (module_geometry.F90:92-140): areaCell (N_CELL), the
 * (N_CELL0,3) edge arrays, the LSM sums (N_CELL0) and TotalVolume */
int nsx_set_geometry(nsx_ctx *ctx, const double *areaCell, const double *dlVV, const double *dlCC,
                     const double *dxCC, const double *dyCC, const double *xnn, const double *ynn,
                     const double *dle1, const double *dle2, const double *doe, const double *sum_xc2,
                     const double *sum_yc2, const double *sum_xcyc, double TotalVolume);
/* sig, dsig (NZ), sigv, dsigv (NZ-1) and sigIni/sigFin (common.mpf) */
int nsx_set_sigma(nsx_ctx *ctx, const double *sig, const double *dsig, const double *sigv, const double *dsigv,
                  double sigIni, double sigFin);
/* inverse-distance weights surrounding/weight (N_VERT,10), Dimsurrounding,
 * dlVsum (N_VERT) and periodic vertex links vertex_pair/vertex_edge (N_VERT) */
int nsx_set_vertex_weights(nsx_ctx *ctx, const int *surrounding, const double *weight, const int *Dimsurrounding,
                           const double *dlVsum, const int *vertex_pair, const int *vertex_edge);
/* boundary flags of update_bc.F90 (bc?c (N_CELL), bc?v (N_VERT)), bctop/bcbot
 * and TopBC/BotBC (common.mpf).  A horizontal ghost cell is periodic (flag 3
 * for u, v, w and p) or a wall (u, v, w 0 no slip, 1 copy or 2 inflow with
 * the reference's constant inflow() of Functions.F90:109-121, p 1: the flags
 * update_bc assigns to walled sides, BC.F90:540-581, 945-964); parallel (-1)
 * ghosts are refused at nsx_commit.  Sigma walls: TopBC 0
 * no slip, 1 free slip, 2 the lid (u = 1 at the top, BC.F90:512-521,
 * 771-777); BotBC 0 or 1. */
int nsx_set_bc(nsx_ctx *ctx, const int *bcuc, const int *bcvc, const int *bcwc, const int *bcpc, const int *bcuv,
               const int *bcvv, const int *bcwv, const int *bcpv, int bctop, int bcbot, int TopBC, int BotBC);
/* PPE matrix AmC0..AmCB (N_CELL0,NZ) of PDE_center (PDE.F90:69-121).  The
 * device regenerates it from the geometry; nsx_commit checks that the
 * regenerated coefficients equal these bit for bit. */
int nsx_set_ppe_matrix(nsx_ctx *ctx, const double *AmC0, const double *AmC1, const double *AmC2,
                       const double *AmC3, const double *AmCT, const double *AmCB);
/* channel wall distance walld(:,k) = |sig(k)| (sgm_les.F90:856-862), per level (NZ) */
int nsx_set_les(nsx_ctx *ctx, const double *walld_level);
/* run parameters of dataParameters.txt (input_parameters.F90:255-436) */
int nsx_set_params(nsx_ctx *ctx, double dt, double Re, double bdx, double bdy, double bdz, double eps, int MaxIters,
                   int FinRK, int Iturbu);
/* builds the device data; hot-path calls are valid afterwards */
int nsx_commit(nsx_ctx *ctx);

/* ---- fields -------------------------------------------------------------- */
/* host array in the reference layout; periodic ghost rows are ignored (the
 * device reads the partner), the wall ghost rows of a cell field are taken */
int nsx_put_field(nsx_ctx *ctx, int field, const double *host);
/* host array in the reference layout; when fill_ghost_rows != 0, periodic
 * ghost rows are filled from their partners (as the reference's BC fills do)
 * and wall ghost rows from the device's wall ghost columns */
int nsx_get_field(nsx_ctx *ctx, int field, double *host, int fill_ghost_rows);
int nsx_copy_field(nsx_ctx *ctx, int dst, int src);

/* ---- time stepping ------------------------------------------------------- */
/* user_initial.F90:229-253 and LoopTimeInitial: after ufn..pfn are put, fill
 * their ghosts, vertex values (tagdir 3), face velocities; zero the AB2 history */
int nsx_init_state(nsx_ctx *ctx);
/* one projection stage, hydro.F90:8-501: (UF,VF,WF,PF) -> (UFNP,..,PFNP),
 * vertex values, faces, AB2 history, eddy.  flag_ab selects AB2 (hydro.F90:225-231). */
int nsx_hydro(nsx_ctx *ctx, int flag_ab, nsx_solve_info *info);
/* nsteps time steps of nsmp3D.F90:404-564 on the resident state (FinRK
 * stages, RK2 average of u,v,w, n <- n+1) */
int nsx_step(nsx_ctx *ctx, int nsteps, int precision, nsx_solve_info *info);
int nsx_solve_info_get(nsx_ctx *ctx, nsx_solve_info *info);

/* ---- single routines on resident fields (unit parity, piecemeal use) ---- */
/* BCglobalVC (BC.F90:202-295): ghost fill of u,v,w in place, vertex values */
int nsx_bcglobalvc(nsx_ctx *ctx, int u, int v, int w, int uv, int vv, int wv, int tagdir);
/* BCglobalP (BC.F90:299-363) */
int nsx_bcglobalp(nsx_ctx *ctx, int p, int pv, int tagdir);
/* grandientLSM (gradientLSM.F90:8-211); gradientP Option 2 (gradientP.F90:8-150) */
int nsx_grad_lsm(nsx_ctx *ctx, int dfdx, int dfdy, int dfds, int f);
/* calcul_face (interpolationFace.F90:7-238), Kim-Choi; f1..fB are face fields */
int nsx_calcul_face(nsx_ctx *ctx, int f1, int f2, int f3, int fT, int fB, int fun, int funv);
/* sgles (sgm_les.F90:24-275): eddy viscosity into NSX_F_EDDY */
int nsx_sgles(nsx_ctx *ctx, int u, int v, int w);
/* AdvDiffNS (AdvDiffNS.F90:8-140): rhs? += diffusion - advection, Gam = eddy + 1/Re,
 * advecting with the resident U1FACE..UBFACE */
int nsx_advdiff(nsx_ctx *ctx, int rhsu, int rhsv, int rhsw, int u, int v, int w, int eddy);
/* Source3D (source3D.F90:8-174): faces into U1FACE..UBFACE, divergence into src */
int nsx_source3d(nsx_ctx *ctx, int src, int u, int v, int w);
/* AdvDiffNS (AdvDiffNS.F90:8-140) with the reference's diffusivity fields
 * Gamx, Gamy, Gamz (cell fields; diffusion3D.F90:187-189 multiplies each
 * component's face-flux sum by its own): rhs? += Gam? diffusion - advection */
int nsx_advdiff_gam(nsx_ctx *ctx, int rhsu, int rhsv, int rhsw, int u, int v, int w, int gamx, int gamy, int gamz);
/* interpolateU (interpolationU.F90:5-190): ghost layers of u, v, w, then
 * their face-normal velocities into the face fields f1..fB */
int nsx_interpolate_u(nsx_ctx *ctx, int f1, int f2, int f3, int fT, int fB, int u, int v, int w);
/* update_rhsp (source3D.F90:176-335): rhs -= lagged Kim-Choi flux of p */
int nsx_update_rhsp(nsx_ctx *ctx, int rhs, int p, int pv);
/* bicgstab (bicgstab.F90:8-168): mean removal, BiCGSTAB_SOLVE, mean removal,
 * BCglobalP(tagdir 3); solves in place into phi, phiv */
int nsx_bicgstab(nsx_ctx *ctx, int phi, int phiv, int rhs, nsx_solve_info *info);
/* BICGSTAB_SOLVE (bicgstab.F90:303-591) alone: A phi = rhs with the PPE
 * matrix of the set-up (the matrix bicgstab passes it), from the phi given,
 * with the reference's stopping rule, in FP64; no mean removal */
int nsx_bicgstab_solve(nsx_ctx *ctx, int phi, int rhs, nsx_solve_info *info);
/* newu3D (newu3D.F90:8-257): U1FACE..UBFACE -= dt grad_n p */
int nsx_newu3d(nsx_ctx *ctx, int p, int pv);

/* ---- differentiation (no counterpart in the reference) ----------------------- */
/* Adjoint of nsx_bicgstab with respect to its right-hand side: for the map
 * phi = Glob_Norm0(A^+ Glob_Norm1(rhs)) it returns lam = Glob_Norm0((A^T)^+
 * Glob_Norm1(g)), so that <g, dphi> = <lam, drhs> (interior cells).  A^T is
 * the PPE operator with each lateral coupling taken from the neighbour's row;
 * the solve is the FP64 BiCGSTAB from lam = 0 with the reference's stopping
 * rule.  lam and g are cell fields, lam != g. */
int nsx_bicgstab_adjoint(nsx_ctx *ctx, int lam, int g, nsx_solve_info *info);

/* The differentiation form.  The control is a per-prism streamwise body force
 * in the field NSX_F_FORCE; nsx_set_body_force(ctx, 1) makes the u momentum
 * of nsx_hydro / nsx_step use it in place of bdx (0 goes back to bdx).  The
 * objective is the bottom-wall shear of the channel averaged over nsteps
 * RK2 steps (FinRK = 2):
 *   J = (1/nsteps) sum_{n=1..nsteps} sum_i area_i nu (u_n(i,2) - u_n(i,1)) / dsig(1) / sum_i area_i,
 * with u(i,1) the bottom ghost layer (-u(i,2) at a no-slip wall).  Solves run
 * in FP64 BiCGSTAB to eps (nsx_set_params); the adjoint solves use A^T, by
 * the FP64 single-reduction CG at eps/100 (at least 1e-16). */
int nsx_set_body_force(nsx_ctx *ctx, int on);
/* nsteps RK2 steps from the resident state (which ends at the new level, as
 * after nsx_step) and J; when grad >= 0 (a cell field other than
 * NSX_F_FORCE), dJ/d(force) per prism into grad by the discrete adjoint:
 * checkpoints of the state every `every` steps, each segment recomputed
 * forward with its stage inputs stored, then stepped back with the
 * transposed kernels (reversed RK2 stages and the adjoint pressure solves).
 * Returns NSX_SOLVE_MAX_ITERS (> 0) when a forward solve reached MaxIters
 * (J and the gradient are then not those of converged steps); the same for
 * nsx_wall_shear_tangent. */
int nsx_wall_shear_run(nsx_ctx *ctx, int nsteps, int every, double *J, int grad);
/* the same nsteps steps with the tangent-linear model along the force
 * direction dir (a cell field): J and dJ = sum over prisms of dJ/d(force) dir */
int nsx_wall_shear_tangent(nsx_ctx *ctx, int nsteps, int dir, double *J, double *dJ);
/* One RK2 step's tangent-linear map and its transpose, at the resident state
 * (which is left unchanged).  A state argument is nine field ids: u, v, w, p
 * (cell fields) and the five face velocities (face fields).  tangent: the
 * tangent tin of the level-n state and dforce of the body force (cell field,
 * or -1 for none) give tout at level n+1; adjoint: the bars bin of the new
 * level give bout of level n and gforce, the bar of the body force. */
int nsx_step_tangent(nsx_ctx *ctx, const int *tin, const int *tout, int dforce);
int nsx_step_adjoint(nsx_ctx *ctx, const int *bin, const int *bout, int gforce);

/* ---- diagnostics ------------------------------------------------------------ */
/* glob_cfl (cfl.F90:7-269, sequential branch) on the cell fields u, v, w and
 * the resident face velocities U1FACE..UBFACE; the values it prints and
 * writes to meanuv.dat, returned instead */
typedef struct nsx_cfl_info {
  double umax, vmax, wmax; /* signed value of largest magnitude (first in k-then-i order) */
  double cfl;              /* max over cells of 0.5 dt sum|face flux| / (area dsig(k)) */
  double aveu, avev, avew; /* volume means, sum phi area dsig(k) / (TotalVolume (sigFin-sigIni)) */
  double flux_total;       /* sum over cells of the net face flux (lateral faces with dsig(k-1)) */
  double flux_max;         /* signed net cell flux of largest magnitude */
  double flux_bc;          /* flux_bot + flux_top + flux_side */
  double flux_bot, flux_top, flux_side; /* through the bottom/top walls and the side (non-interior) faces */
} nsx_cfl_info;
int nsx_glob_cfl(nsx_ctx *ctx, int u, int v, int w, nsx_cfl_info *out);
/* No counterpart in the reference: nsx_glob_cfl for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid allocation).  It allocates nothing: its partial sums and outputs
 * live in a static per-device scratch of the library (262,272 bytes: 16 rows
 * of at most 2048 blocks, and the outputs); a context whose grid exceeds 2048
 * blocks is refused.  The scratch is shared by every call on a GPU and keeps
 * the last call's contents, so calls that overlap on one GPU race. */
int nsx_glob_cfl_noalloc(nsx_ctx *ctx, int u, int v, int w, nsx_cfl_info *out);

/* ---- channel statistics (stats.F90) ---------------------------------------- */
/* The twelve accumulators of module stats (module_stats.F90), resident on the
 * device as cell fields. */
enum nsx_sm {
  NSX_SM_UF = 0, NSX_SM_VF, NSX_SM_WF, NSX_SM_U2F, NSX_SM_V2F, NSX_SM_W2F,
  NSX_SM_UVF, NSX_SM_VWF, NSX_SM_WUF, NSX_SM_PF, NSX_SM_P2F, NSX_SM_EF
};
/* sm_init (stats.F90:34-84): allocate (first call) and zero the accumulators,
 * sample count 0 */
int nsx_sm_init(nsx_ctx *ctx);
/* glob_sample (stats.F90:91-197): add u, v, w, p, their products and the eddy
 * viscosity (NSX_F_EDDY) of every column i <= N_CELL0 and every layer k = 1..NZ
 * (cell field ids); the sample count grows by one */
int nsx_glob_sample(nsx_ctx *ctx, int u, int v, int w, int p);
/* sm_n of module stats */
int nsx_sm_count(nsx_ctx *ctx, int *n);
/* one accumulator NSX_SM_* into host(N_CELL,NZ) (rows 1..N_CELL0 written) */
int nsx_sm_get(nsx_ctx *ctx, int which, double *host);
/* the plane sums of glob_put_plane_mm (stats.F90:266-279), in its order of
 * summation: planes[q*NZ + k-1] = sum_{i=1..N_CELL0} sm_q(i,k)/sm_n for the
 * nine velocity moments q = NSX_SM_UF..NSX_SM_WUF and k = 2..NZ-1 (0 at k = 1
 * and NZ); *ns = sm_n */
int nsx_sm_planes(nsx_ctx *ctx, double *planes, int *ns);
/* glob_put_plane_mm (stats.F90:204-391): the plane sums, then the files
 * <dir>/sm-NNNNN.dat (means and the TKE line), srms-NNNNN.dat (rms) and
 * stau-NNNNN.dat (stresses), NNNNN = sm_n, in the reference's formats */
int nsx_glob_put_plane_mm(nsx_ctx *ctx, const char *dir);
/* the file writing of glob_put_plane_mm alone (stats.F90:301-363), on plane
 * sums laid out as nsx_sm_planes returns them; needs no device */
int nsx_write_plane_stats(const double *planes, int NZ, int n_cell0, int ns, double Re, const char *dir);
/* a REAL(8) formatted as the reference's output does: kind 0 = Ew.d edit
 * descriptor, kind 1 = list-directed (the compiler of the reference build) */
int nsx_format_real(double x, int kind, int w, int d, char *out, int outlen);

/* ---- Tecplot snapshots (src/Save) ------------------------------------------ */
/* SavetecVertex (SavetecVertex.F90:185-235): <dir>/V-NNNN.tec (NNNN = counter,
 * the reference's SaveCounter) with the vertex coordinates xvt, yvt, zvt
 * (N_VERT,NZ-1) of the caller and the resident vertex fields NSX_F_UFV..PFV */
int nsx_savetec_vertex(nsx_ctx *ctx, const char *dir, int counter, double time, const double *xvt,
                       const double *yvt, const double *zvt);
/* SavetecCenter (SavetecCenter.F90:185-235): <dir>/C-NNNNN.tec with the vertex
 * coordinates and the resident cell fields u, v, w, p (field ids) */
int nsx_savetec_center(nsx_ctx *ctx, const char *dir, int counter, double time, const double *xvt,
                       const double *yvt, const double *zvt, int u, int v, int w, int p);
/* the file writing of the two alone, on host arrays in the reference layout
 * (No_vp (N_CELL0,3) 1-based; vertex fields (N_VERT,NZ-1); cell fields
 * (N_CELL,NZ)); needs no device */
int nsx_write_tec_vertex(const char *dir, int counter, double time, int N_CELL0, int N_VERT, int NZ,
                         const int *No_vp, const double *xvt, const double *yvt, const double *zvt,
                         const double *ufv, const double *vfv, const double *wfv, const double *pfv);
int nsx_write_tec_center(const char *dir, int counter, double time, int N_CELL0, int N_CELL, int N_VERT, int NZ,
                         const int *No_vp, const double *xvt, const double *yvt, const double *zvt,
                         const double *u, const double *v, const double *w, const double *p);
/* SaveParaviewVertex (SaveParaviewVertex.F90:185-278): <dir>/V-NNNN.vtk with
 * the resident vertex velocities (the solid-phase ones are zero) */
int nsx_save_paraview_vertex(nsx_ctx *ctx, const char *dir, int counter, const double *xvt, const double *yvt,
                             const double *zvt);
/* SavetecVC (SavetecVertexCenter.F90:212-305): <dir>/VC-NNNNN.tec, vertices
 * and cell centres as nodes: the caller's coordinates xvt, yvt, zvt
 * (N_VERT,NZ-1) and xct, yct, zct (N_CELL,NZ), the resident vertex fields
 * NSX_F_UFV..PFV and the cell fields u, v, w, p (field ids) */
int nsx_savetec_vc(nsx_ctx *ctx, const char *dir, int counter, double time, const double *xvt, const double *yvt,
                   const double *zvt, const double *xct, const double *yct, const double *zct, int u, int v, int w,
                   int p);
/* their file writing alone, on host arrays (usv, vsv, wsv may be NULL) */
int nsx_write_vtk_vertex(const char *dir, int counter, int N_CELL0, int N_VERT, int NZ, const int *No_vp,
                         const double *xvt, const double *yvt, const double *zvt, const double *ufv, const double *vfv,
                         const double *wfv, const double *usv, const double *vsv, const double *wsv);
int nsx_write_tec_vc(const char *dir, int counter, double time, int N_CELL0, int N_CELL, int N_VERT, int NZ,
                     const int *No_vp, const double *xvt, const double *yvt, const double *zvt, const double *xct,
                     const double *yct, const double *zct, const double *ufv, const double *vfv, const double *wfv,
                     const double *pfv, const double *u, const double *v, const double *w, const double *p);

/* ---- precision form (no counterpart in the reference) ------------------------ */
/* Arithmetic of the pressure solve in nsx_hydro / nsx_bicgstab (nsx_step takes
 * its own precision and uses the refinement set here).  NSX_PREC_MIXED runs
 * the Krylov vectors, SpMV and preconditioner in FP32 and keeps the solution,
 * the true residual, inner products and norms in FP64; the solve stops when
 * the true FP64 residual meets the reference's rule with eps
 * (nsx_set_params).  The FP32 solver is conjugate gradients on the
 * correction e; whenever its recursive residual has fallen by tol (delta,
 * e.g. 0.1) since the last true residual, or meets the rule, the true FP64
 * residual of x + e replaces it and e is folded into x (reliable updates).
 * The operator is symmetric negative semi-definite, and CG's smooth residual
 * keeps FP32 from stalling.  A CG that reaches its cap (2 MaxIters + 100
 * iterations) or breaks down leaves x at its last update and the FP64
 * BiCGSTAB finishes.  Refinement (default NSX_REFINE_CGSR_BLOCK, delta 0.1):
 *   NSX_REFINE_CG          point-Jacobi CG, two kernels per iteration;
 *   NSX_REFINE_CGSR        the same CG in the single-reduction form of
 *                          Chronopoulos and Gear: one kernel and one global
 *                          reduction per iteration; tol at least 0.05 (its
 *                          FP32 recurrences drift faster);
 *   NSX_REFINE_CGSR_BLOCK  the same with a block-Jacobi preconditioner: A
 *                          restricted to each group of four consecutive
 *                          layers of a column (a 4x4 tridiagonal solve in
 *                          the kernel, no extra memory traffic);
 *   NSX_REFINE_CGSR_LINE   the same with a line preconditioner: A restricted
 *                          to the interior layers of each column (a
 *                          tridiagonal solve over the column in the kernel's
 *                          shared memory; columns of at most 1024 layers).
 *   NSX_REFINE_MODAL       CG preconditioned by the vertical modes of the
 *                          operator: per column, the generalized eigenvectors
 *                          of the vertical couplings split A into one 2D
 *                          problem per mode; the 16 lowest get a V-cycle of
 *                          aggregation multigrid on the column graph, the
 *                          others their diagonal.  The FP32 CG runs in
 *                          cycles that start from the true FP64 residual and
 *                          end when <r, M^-1 r> has fallen by 1e-10 (tol
 *                          unused).  Its tables are built at the first solve.
 * Values 0..2 (FP32 BiCGSTAB forms) are refused: FP32 BiCGSTAB breaks down on
 * this operator.  max_passes is unused (at least 1). */
enum nsx_refine {
  NSX_REFINE_CG = 3,
  NSX_REFINE_CGSR = 4,
  NSX_REFINE_CGSR_BLOCK = 5,
  NSX_REFINE_CGSR_LINE = 6,
  NSX_REFINE_MODAL = 7
};
int nsx_set_precision(nsx_ctx *ctx, int precision, int refine, double tol, int max_passes);

/* Initial guess of nsx_step's pressure solves in the precision forms
 * (NSX_PREC_MIXED, NSX_PREC_MIXED_FLUX32, NSX_PREC_FP64_CG; no counterpart in
 * the reference).  NSX_GUESS_PREVIOUS (the default): the stage's input
 * pressure, as bicgstab starts from pfnp in the reference.
 * NSX_GUESS_EXTRAPOLATE: 2 p_k - p_{k-1} from the last two solutions of the
 * same RK stage (the stage's input pressure until two are stored); about 5 %
 * fewer iterations at 6.3M prisms and eps 1e-13.  The solution still meets
 * the same stopping rule.  It holds five more cell fields (allocated here:
 * after nsx_commit), applies to nsx_step only (not to nsx_hydro, nsx_bicgstab
 * or the differentiation form, whose checkpointed recomputation repeats the
 * forward bit for bit), and starts again after nsx_init_state, nsx_put_field
 * or a step in FP64. */
enum nsx_guess { NSX_GUESS_PREVIOUS = 0, NSX_GUESS_EXTRAPOLATE = 1 };
int nsx_set_initial_guess(nsx_ctx *ctx, int mode);

/* ---- timing ---------------------------------------------------------------- */
/* per-kernel device time accumulated with HIP events since the last reset;
 * names: "bicg_a", "bicg_b", "bicg_c" (the three BiCGSTAB kernels in FP64,
 * iterations that ran only), "cg32_w", "cg32_x" (the two kernels of an FP32
 * CG iteration), "cg32_sr" (the kernel of the single-reduction FP32 CG, every
 * launch that ran), "cg64_sr" (its FP64 instance) and "stage" (all non-solver
 * kernels of a projection stage) */
int nsx_timing_enable(nsx_ctx *ctx, int on);
int nsx_timing_get(nsx_ctx *ctx, const char *name, double *total_ms, long long *launches);
int nsx_timing_reset(nsx_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif /* NSX_H */
