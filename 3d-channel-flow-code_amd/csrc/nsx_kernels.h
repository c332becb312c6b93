// nsx_kernels.h -- launchers of the stage kernels (nsx_stage.hip) and of the
// pressure solver (nsx_bicgstab.hip), called by the context (nsx_context.hip).
#pragma once
#include "nsx_dev.h"
#include "nsx_modal.h"

// Per-block partial sums of one reduction site: [rows][grid]
struct Partials {
  double *src;   // Glob_Norm(rhs,1) sum (k_source / k_gnorm_sum), 1 row
  double *init;  // initial residual norms, 5 rows
  double *a;     // <rhat, v>, 1 row
  double *b;     // <t,s>, <t,t>, 2 rows
  double *c;     // |r|^2, |r|_1, <rhat,r>, sum x Vol, 4 rows
  double *u;     // FP32 CG: true residual after a reliable update, 4 rows
};

// BiCGSTAB scalars, device resident (bicgstab.F90:355-368 locals).  Each
// field has one writer kernel (block 0 of it) and is read only by later
// kernels, so no block reads a value that its own kernel may be rewriting.
// scalars of the single-reduction FP32 CG passed from launch to launch
// (launch j reads SolverState::cgs[(j-1)&1], block 0 writes cgs[j&1])
struct CgsCarry {
  double gamma, alpha;  // gamma_i and alpha_i of the last iteration
  double beta;          // beta_i of the last iteration (0 if it was the first)
  double maxr;          // largest |r| since the last true residual
  int mode;             // this launch's mode: 0 product (M), 1 iteration (N), 2 true residual (T)
  int iter;             // iterations (launches in mode N) so far
  int set;              // buffer set holding the current rt, om, sg
  int done;
  int pend;             // p lags: it holds p_{i-1} of the last iteration i, which stored no p
                        // (u_i is in rt of the other set; no p at all if i was the first)
  int pend_e;           // e lacks alpha_i p_i of that iteration (cleared by the fold into x)
  int first;            // the last iteration was the solve's first
  int pad;
};

struct SolverState {
  double a_rho[2];   // rho_k = <rhat, r_k> by parity of k        (phase A)
  double b_alpha;    // alpha_k                                    (phase B; init: 1)
  double c_omega;    // omega_k                                    (phase C; init: 1)
  double b1, b2;     // |b|_1, |b|_2 (0 -> 1)                      (phase A, k = 0)
  double e1, e2;     // relative residual norms of x_k             (phase A)
  double xvol;       // sum x_k Vol for Glob_Norm(phi,0)            (phase A)
  double avec;       // Glob_Norm(rhs,1) average                   (init)
  double eps;        //                                            (init)
  int c_iter;        // iterations completed before phase A runs   (phase C; init: 0)
  int a_iter;        // k of the last phase A                      (phase A)
  int done, exit_code;                                        //   (phase A or B)
  int max_iters, pad;                                         //   (init)
  // reliable updates of the CG forms
  int mp_pass;       // true residuals formed after the start (reliable updates)
  int pad2;
  double delta;      // update when |r| < delta x the largest |r| since the last one  (init)
  double r0N;        // |r| of the initial residual                    (phase A, k = 0)
  double maxrr[2];   // largest |r| since the last update, by parity
  double a_beta;     // beta_k of y = rt + beta q                      (phase A)
  int ureq[2];       // FP32 CG: W_k asks U_{k+1} for a true residual, by parity of k (phase W)
  CgsCarry cgs[2];   // single-reduction FP32 CG, by parity of the launch
  int cgs_runs;      // launches of the single-reduction CG that ran (block 0 of each)
  int pad3;
  // the CG with the vertical-mode preconditioner (NSX_REFINE_MODAL)
  double m_rho;       // rho_i = <r_i, M^-1 r_i> of the iteration running     (product kernel)
  double m_rho_prev;  // rho of the last completed iteration                   (update kernel)
  double m_rho0;      // rho of the cycle's first iteration                    (product kernel)
  int m_stop;         // the cycle has met its drop: its remaining launches return (product kernel; fold: 0)
  int m_iters;        // iterations completed                                   (update kernel)
};

// the vectors of the single-reduction CG: e and p in place, the scaled
// residual rt = r/A0, om = (A rt)/A0 and sg = s/A0 in two sets (FP32 in the
// precision form, FP64 in its FP64 counterpart)
template <typename T>
struct CgsVecsT {
  T *e, *p, *rt[2], *om[2], *sg[2];
};
using CgsVecs = CgsVecsT<float>;

int nsx_grid_for(uint32_t n);
int solver_blocks_per_cu(const Geo &g);
int cg_blocks_per_cu(const Geo &g);
int cgs_blocks_per_cu(const Geo &g);
int cgs64_blocks_per_cu(const Geo &g);

// ---- stage kernels (stream-ordered launches; return hipError_t as int)
int launch_vertex(const Geo &g, hipStream_t s, int nf, const double *const f[3], double *const fv[3],
                  const signed char *const act[3], const int zero_bot[3], const int zero_top[3], int tagdir,
                  const int *lid_top = nullptr, const double *inflow = nullptr);
int launch_wall_ghosts(const Geo &g, hipStream_t s, double *f, int kind);
// (Cs Delta)^2 with Mason's damping per prism, once at commit (Geo::nutg)
int launch_les_length(const Geo &g, hipStream_t s, double *nutg);
int launch_les(const Geo &g, hipStream_t s, const double *u, const double *v, const double *w, const double *uv,
               const double *vv, const double *wv, double *eddy);
int launch_momentum(const Geo &g, hipStream_t s, const double *const uvw[3], const double *const vtx[3],
                    const double *const face[5], const double *eddy, const double *const rhs_in[3],
                    double *const hist[3], double *const rhs_out[3], double *const ustar[3], double c1, double c2,
                    const double *force = nullptr, const double *const gam[3] = nullptr);
// the same two in FP32 arithmetic on the FP64 fields (NSX_PREC_MIXED_FLUX32)
int launch_les32(const Geo &g, hipStream_t s, const double *u, const double *v, const double *w, const double *uv,
                 const double *vv, const double *wv, double *eddy);
int launch_momentum32(const Geo &g, hipStream_t s, const double *const uvw[3], const double *const vtx[3],
                      const double *const face[5], const double *eddy, double *const hist[3],
                      double *const ustar[3], double c1, double c2, const double *force);
int launch_source(const Geo &g, hipStream_t s, const double *const uvw[3], double *const face[5], double *src_out,
                  double *rhsp_out, const double *pv, Partials part);
int launch_update_rhsp(const Geo &g, hipStream_t s, double *rhs, const double *pv);
int launch_grad_lsm(const Geo &g, hipStream_t s, const double *f, double *dx, double *dy, double *ds);
int launch_calcul_face(const Geo &g, hipStream_t s, const double *f, const double *fv, double *const out[5]);
int launch_corrector(const Geo &g, hipStream_t s, const double *p, const double *const us[3],
                     double *const unp[3]);
int launch_newu(const Geo &g, hipStream_t s, const double *p, const double *pv, double *const face[5]);
int launch_rk2avg(const Geo &g, hipStream_t s, const double *const un[3], double *const unp[3]);
// o = 2 a - b over the cell fields (the extrapolated initial guess)
int launch_extrapolate(const Geo &g, hipStream_t s, const double *a, const double *b, double *o);
int launch_ghost_layers(const Geo &g, hipStream_t s, double *f, int kind /*0 p, 1 u, 2 v, 3 w*/);
int launch_ppe_diag(const Geo &g, hipStream_t s, double *A0);
int launch_copy(hipStream_t s, size_t n, const double *src, double *dst);
// glob_cfl (nsx_diag.hip): partial rows per block, and the 13 outputs into out
// channel statistics (stats.F90): the twelve accumulators of glob_sample,
// in the order of module stats; the first nine are glob_put_plane_mm's
enum { SM_UF = 0, SM_VF, SM_WF, SM_U2F, SM_V2F, SM_W2F, SM_UVF, SM_VWF, SM_WUF, SM_PF, SM_P2F, SM_EF, SM_N };
enum { SM_NPLANE = 9 };
struct StatsBufs {
  double *a[SM_N];
};
int launch_glob_sample(const Geo &g, hipStream_t s, const double *u, const double *v, const double *w,
                       const double *p, const double *eddy /* nullable */, StatsBufs sm);
int launch_plane_sums(const Geo &g, hipStream_t s, StatsBufs sm, int ns, double *out /* SM_NPLANE x NZ */);

int glob_cfl_rows();
int launch_glob_cfl(const Geo &g, hipStream_t s, const double *u, const double *v, const double *w,
                    const double *const face[5], double *part, double *out);
// the same into the library's static per-device scratch (nsx_diag.hip);
// glob_cfl_scratch_init(dev) must have run once for dev; returns -1 when the
// grid exceeds the scratch, -2 when it was not set up, else a hipError_t
int glob_cfl_scratch_init(int dev);
int launch_glob_cfl_scratch(const Geo &g, hipStream_t s, int dev, const double *u, const double *v, const double *w,
                            const double *const face[5], double **out_slot);

// ---- differentiation form (nsx_adjoint.hip): tangent-linear kernels of the
// nonlinear stage steps, transposed kernels of every stage step, the
// bottom-wall-shear objective
int launch_tlm_les(const Geo &g, hipStream_t s, const double *const uvw[3], const double *const vtx[3],
                   const double *const duvw[3], const double *const dvtx[3], double *deddy);
int launch_adj_les(const Geo &g, hipStream_t s, const double *const uvw[3], const double *const vtx[3],
                   const double *eddyb, double *const uvwb[3], double *const vtxb[3]);
int launch_tlm_momentum(const Geo &g, hipStream_t s, const double *const uvw[3], const double *const vtx[3],
                        const double *const face[5], const double *eddy, const double *const duvw[3],
                        const double *const dvtx[3], const double *const dface[5], const double *deddy,
                        const double *dforce, double *const dustar[3]);
int launch_adj_momentum(const Geo &g, hipStream_t s, const double *const uvw[3], const double *const vtx[3],
                        const double *const face[5], const double *eddy, const double *const usb[3],
                        double *const uvwb[3], double *const vtxb[3], double *const faceb[5], double *eddyb,
                        double *forceb);
int launch_adj_source(const Geo &g, hipStream_t s, const double *const faceb[5], const double *rhspb,
                      double *const uvwb[3], double *pvb);
int launch_adj_corrector(const Geo &g, hipStream_t s, const double *const ub[3], double *const usb[3], double *pb);
int launch_adj_newu(const Geo &g, hipStream_t s, const double *const faceb[5], double *pb, double *pvb);
// valb: a vertex-field work buffer
int launch_adj_vertex(const Geo &g, hipStream_t s, const double *fvb, const signed char *act, double *valb,
                      double *fb);
int launch_adj_ghost(const Geo &g, hipStream_t s, double *fb, int kind /*0 p, 1 u, 2 v, 3 w*/);
// J += scale * bottom-wall shear of u (part: 64 doubles of partial sums)
int launch_wall_shear(const Geo &g, hipStream_t s, const double *u, double scale, double *part, double *J);
int launch_wall_shear_seed(const Geo &g, hipStream_t s, double *ub, double scale);
// y = a y + b x (x may be null: y = a y)
int launch_axpby(hipStream_t s, size_t n, double a, double *y, double b, const double *x);

// ---- solver
int launch_bicg_init(const Geo &g, hipStream_t s, const double *rhs, const double *phi_in, double *x, double *r,
                     double *rhat, Partials part, SolverState *st, double eps, int max_iters);
// one iteration, three kernels (A, B, C) on the scaled vectors rt = r/A0,
// q = (p - omega v)/A0, vt = v/A0, and y, t; x is the solution
int launch_bicg_iter(const Geo &g, hipStream_t s, double *x, double *rt, const double *rhat, double *q, double *vt,
                     double *tv, Partials part, SolverState *st,
                     hipEvent_t *ev /* nullable: 6 events, start/stop of A, B, C */);
// mixed precision, FP32 CG with reliable updates: start, then W, X per iteration;
// v = e, rt, p, w, e_b, rt_b, p_b (the second set takes the true-residual pass)
int launch_cg_init(const Geo &g, hipStream_t s, const double *rhs, const double *phi_in, double *x, float *rt,
                   float *scratch, float *e, Partials part, SolverState *st, double eps, double delta, int max_iters);
int launch_cg_iter(const Geo &g, hipStream_t s, const double *rhs, double *x, float *const v[7], Partials part,
                   SolverState *st, hipEvent_t *ev /* nullable: 4 events, start/stop of W, X */);
// one launch of the single-reduction FP32 CG (launch j of the solve; its mode
// is decided on the device); ev: one timed kernel; pc: the preconditioner,
// 0 point Jacobi, 1 4-layer blocks, 2 lines (whole columns)
int launch_cgs(const Geo &g, hipStream_t s, int j, const double *rhs, double *x, const CgsVecs &v, Partials part,
               SolverState *st, hipEvent_t *ev, int pc);
// the same CG with FP64 vectors and the 4-layer block preconditioner
int launch_cg_init64(const Geo &g, hipStream_t s, const double *rhs, const double *phi_in, double *x, double *rt,
                     double *scratch, double *e, Partials part, SolverState *st, double eps, double delta,
                     int max_iters);
int launch_cgs64(const Geo &g, hipStream_t s, int j, const double *rhs, double *x, const CgsVecsT<double> &v,
                 Partials part, SolverState *st, hipEvent_t *ev);
int launch_gnorm_apply(const Geo &g, hipStream_t s, double *x, const SolverState *st);
int launch_gnorm_sum(const Geo &g, hipStream_t s, const double *x, Partials part);

// ---- the CG with the vertical-mode preconditioner (NSX_REFINE_MODAL)
// u = M^-1 r (modal transform, the low modes' V-cycle, the others' diagonal,
// transform back; u is stored at every slot, zero outside the interior) and
// the partial sums of <r, u> into P.a
// (ev: null, or six events: the forward transform's start and stop (recorded
// by its dispatch), the V-cycle's start and stop (recorded on the stream
// around its launches), the backward transform's start and stop)
int launch_modal_apply(const Geo &g, hipStream_t s, const ModalDev &M, const float *r, float *u, Partials part,
                       const SolverState *st, hipEvent_t *ev = nullptr);
// p_out = u + beta p_in (beta = 0 when first: the cycle's first iteration) at
// the seven points, w = A p_out, partials of <p, w> into P.b; rho from P.a
// (M.agrid partials) kept in the state; a rho below NSX_MODAL_DROP times the
// cycle's first ends the cycle (st->m_stop)
int launch_modal_spmv(const Geo &g, hipStream_t s, const ModalDev &M, const float *u, const float *pin, float *pout,
                      float *w, Partials part, SolverState *st, int first, hipEvent_t *ev = nullptr);
// alpha = rho / <p, w>; e += alpha p, r -= alpha w
int launch_modal_update(const Geo &g, hipStream_t s, const float *p, const float *w, float *e, float *r,
                        Partials part, SolverState *st, hipEvent_t *ev = nullptr);
// the true FP64 residual of x + e into r and its partials (P.c); then x += e,
// e = 0, a new cycle (m_stop = 0) and the reference's stopping rule (st: e1,
// e2, xvol, done, exit code; the iteration cap after st->m_iters iterations)
int launch_modal_true(const Geo &g, hipStream_t s, const double *rhs, double *x, float *e, float *r, Partials part,
                      SolverState *st, int cap);
