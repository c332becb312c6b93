// nsx_bicgstab.hip -- Jacobi-preconditioned BiCGSTAB on the 7-point prism
// stencil (bicgstab.F90:303-591), three fused phases per iteration, one per
// global synchronisation point of the algorithm:
//
//   A  p = r + beta (p - omega v)   (p = r at iter 0)     update_pk   :726-773
//      v = A (p / A0)               (Jacobi + SpMV)       diag_precdt :779-925, get_Apk :931-1001
//      <rhat, v>  -> alpha = rho / <rhat, v>                               :509-520
//   B  s = r - alpha v ; t = A (s / A0) ; <t,s>, <t,t> -> omega          :523-537
//   C  x += alpha p/A0 + omega s/A0 ; r = s - omega t                    update_xkrk :1047-1111
//      ||r||_2^2, ||r||_1, <rhat, r> (next rho), sum x Vol (Glob_Norm)
//
// The stencil input of A and B (p and s) is recomputed at the 7 points from
// the previous iterates, so p and s never make a separate pass through HBM;
// p and v are double-buffered across iterations.  The matrix is PDE_center's,
// regenerated per row from the edge coefficient ppec(i,j), the vertical
// coefficient W(i,k) and dsigv (bit-identical to AmC0..AmCB), and the Jacobi
// scaling multiplies by invD = 1/AmC0.
//
// Algorithmic HBM bytes per interior prism and iteration (FP64):
//   A: read r, p, v, rhat, W, invD; write p', v'          8 x 8 B
//   B: read r, v', W, invD; write s, t                      6 x 8 B
//   C: read x, p', s, t, invD, rhat; write x, r            8 x 8 B
// (neighbour re-reads and the per-column ppec/area are cache hits).
//
// Each kernel exits at once when the device flag st->done is set, so the host
// can queue many iterations ahead and poll rarely.
#include "nsx_kernels.h"

#define FOR_EACH_CELL(g, c)                                                         \
  uint32_t beg_, end_;                                                              \
  block_range((g).ncell, beg_, end_);                                               \
  for (uint32_t base_ = beg_; base_ < end_; base_ += NSX_BLOCK)                     \
    for (uint32_t c = base_ + threadIdx.x, once_ = 1; once_ && c < end_; once_ = 0)

#define TINY ((double)1.0e-30f) /* REAL(4) literal of bicgstab.F90:479-517 */

enum { EXIT_CONVERGED = 0, EXIT_NO_SOLVE = 1, EXIT_BREAKDOWN = 2, EXIT_MAX = 3 };

// Row of the PPE matrix at (i,k): diagonal and the 5 off-diagonals
struct Row {
  double A0, a1, a2, a3, AT, AB;
};

__device__ inline Row ppe_row(const Geo &g, const PPEOp &op, int i, int k, size_t s) {
  Row r;
  const double sv = g.dsigv[k - 1];
  r.AT = op.W[s];
  r.AB = op.W[s - 1];
  r.a1 = g.ppec[i * 3] * sv;
  r.a2 = g.ppec[i * 3 + 1] * sv;
  r.a3 = g.ppec[i * 3 + 2] * sv;
  r.A0 = -1.0 * (r.AT + r.AB + r.a1 + r.a2 + r.a3);
  return r;
}

// y = A (f * invD) at (i,k) where f(x) is evaluated at the 7 points; the
// vertical ghost of a pressure-like vector is the Neumann copy (BCpcenter)
template <class F>
__device__ inline double apply_row(const Geo &g, const PPEOp &op, int i, int k, size_t s, F f, double &fs) {
  const Row A = ppe_row(g, op, i, k, s);
  fs = f(s);
  const double ys = fs * op.invD[s];
  const double yu = (k + 1 <= g.NZ - 2) ? f(s + 1) * op.invD[s + 1] : ys;
  const double yd = (k - 1 >= 1) ? f(s - 1) * op.invD[s - 1] : ys;
  const int e = i * 3;
  const size_t x0 = (size_t)g.nbr[e] * g.NZ + k, x1 = (size_t)g.nbr[e + 1] * g.NZ + k,
               x2 = (size_t)g.nbr[e + 2] * g.NZ + k;
  const double y0 = f(x0) * op.invD[x0], y1 = f(x1) * op.invD[x1], y2 = f(x2) * op.invD[x2];
  return A.A0 * ys + A.a1 * y0 + A.a2 * y1 + A.a3 * y2 + A.AT * yu + A.AB * yd;
}

// ------------------------------------------------------------ initial -----
// bicgstab.F90:420-469 after Glob_Norm(rhs,1): r = rhat = b - A x0, norms
__global__ __launch_bounds__(NSX_BLOCK) void k_bicg_init(Geo g, PPEOp op, const double *rhs, const double *phi,
                                                          double *x, double *r, double *rhat, double *partials,
                                                          unsigned int *ticket, SolverState *st, double eps,
                                                          int max_iters) {
  const double avec = st->avec_rhs;
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  FOR_EACH_CELL(g, c) {
    int i, k;
    cell_of(g, c, i, k);
    const size_t s = (size_t)i * g.NZ + k;
    const double area = g.area[i];
    const double b = rhs[s] - avec * area * g.dsig[k];
    const Row A = ppe_row(g, op, i, k, s);
    const int e = i * 3;
    const double ph = phi[s];
    const double Ax = A.A0 * ph + A.a1 * phi[(size_t)g.nbr[e] * g.NZ + k] + A.a2 * phi[(size_t)g.nbr[e + 1] * g.NZ + k] +
                      A.a3 * phi[(size_t)g.nbr[e + 2] * g.NZ + k] + A.AT * phi[s + 1] + A.AB * phi[s - 1];
    const double res = b - Ax;
    r[s] = res;
    rhat[s] = res;
    x[s] = ph;
    acc[0] += b * b;
    acc[1] += res * res;
    acc[2] += fabs(b);
    acc[3] += fabs(res);
    acc[4] += ph * (area * g.dsig[k]);
  }
  if (block_partials_and_ticket<5>(acc, partials, ticket)) {
    double tot[5];
    final_sums<5>(tot, partials, ticket);
    if (threadIdx.x == 0) {
      double b2 = sqrt(tot[0]), r2 = sqrt(tot[1]), b1 = tot[2], r1 = tot[3];
      if (b2 == 0) b2 = 1.0;
      if (b1 == 0) b1 = 1.0;
      st->b1 = b1;
      st->b2 = b2;
      st->e2 = r2 / b2;
      st->e1 = r1 / b1;
      st->rho = tot[1];  // <rhat, r> with rhat = r
      st->rhokm = 1.0;
      st->alpha = 1.0;
      st->omega = 1.0;
      st->iter = 0;
      st->eps = eps;
      st->max_iters = max_iters;
      st->xvol = tot[4];
      const bool conv = (st->e1 < eps) || (st->e2 < eps);
      st->done = conv ? 1 : 0;
      st->exit_code = conv ? EXIT_NO_SOLVE : EXIT_CONVERGED;
    }
  }
}

// ------------------------------------------------------------- phase A ----
__global__ __launch_bounds__(NSX_BLOCK) void k_bicg_a(Geo g, PPEOp op, const double *r, const double *rhat,
                                                       const double *pold, const double *vold, double *pnew,
                                                       double *vnew, double *partials, unsigned int *ticket,
                                                       SolverState *st) {
  if (st->done) return;
  const bool first = st->iter == 0;
  const double beta = first ? 0.0 : (st->rho / st->rhokm) * (st->alpha / st->omega);
  const double om = st->omega;
  auto pn = [&](size_t x) -> double { return first ? r[x] : r[x] + beta * (pold[x] - om * vold[x]); };
  double acc[1] = {0.0};
  FOR_EACH_CELL(g, c) {
    int i, k;
    cell_of(g, c, i, k);
    const size_t s = (size_t)i * g.NZ + k;
    double ps;
    const double vs = apply_row(g, op, i, k, s, pn, ps);
    pnew[s] = ps;
    vnew[s] = vs;
    acc[0] += rhat[s] * vs;
  }
  if (block_partials_and_ticket<1>(acc, partials, ticket)) {
    double tot[1];
    final_sums<1>(tot, partials, ticket);
    if (threadIdx.x == 0) {
      if (fabs(tot[0]) < TINY) {
        st->done = 1;
        st->exit_code = EXIT_BREAKDOWN;
      } else {
        st->alpha = st->rho / tot[0];
      }
    }
  }
}

// ------------------------------------------------------------- phase B ----
__global__ __launch_bounds__(NSX_BLOCK) void k_bicg_b(Geo g, PPEOp op, const double *r, const double *v,
                                                       double *sout, double *tout, double *partials,
                                                       unsigned int *ticket, SolverState *st) {
  if (st->done) return;
  const double al = st->alpha;
  auto sn = [&](size_t x) -> double { return r[x] - al * v[x]; };
  double acc[2] = {0.0, 0.0};
  FOR_EACH_CELL(g, c) {
    int i, k;
    cell_of(g, c, i, k);
    const size_t s = (size_t)i * g.NZ + k;
    double ss;
    const double ts = apply_row(g, op, i, k, s, sn, ss);
    sout[s] = ss;
    tout[s] = ts;
    acc[0] += ts * ss;
    acc[1] += ts * ts;
  }
  if (block_partials_and_ticket<2>(acc, partials, ticket)) {
    double tot[2];
    final_sums<2>(tot, partials, ticket);
    if (threadIdx.x == 0) st->omega = tot[0] / tot[1];
  }
}

// ------------------------------------------------------------- phase C ----
__global__ __launch_bounds__(NSX_BLOCK) void k_bicg_c(Geo g, PPEOp op, double *x, double *r, const double *rhat,
                                                       const double *p, const double *sv, const double *tv,
                                                       double *partials, unsigned int *ticket, SolverState *st) {
  if (st->done) return;
  const double al = st->alpha, om = st->omega;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  FOR_EACH_CELL(g, c) {
    int i, k;
    cell_of(g, c, i, k);
    const size_t s = (size_t)i * g.NZ + k;
    const double d = op.invD[s];
    const double y = p[s] * d, z = sv[s] * d;
    const double xn = x[s] + al * y + om * z;
    const double rr = sv[s] - om * tv[s];
    x[s] = xn;
    r[s] = rr;
    acc[0] += rr * rr;
    acc[1] += fabs(rr);
    acc[2] += rhat[s] * rr;
    acc[3] += xn * (g.area[i] * g.dsig[k]);
  }
  if (block_partials_and_ticket<4>(acc, partials, ticket)) {
    double tot[4];
    final_sums<4>(tot, partials, ticket);
    if (threadIdx.x == 0) {
      const double r2 = sqrt(tot[0]), r1 = tot[1];
      st->e2 = r2 / st->b2;
      st->e1 = r1 / st->b1;
      st->rhokm = st->rho;
      st->iter = st->iter + 1;
      st->xvol = tot[3];
      double enorm = (double)1.0e10f;
      if (enorm > st->e1) enorm = st->e1;
      if (enorm > st->e2) enorm = st->e2;
      if (enorm < st->eps) {
        st->done = 1;
        st->exit_code = EXIT_CONVERGED;
      } else if (st->iter > st->max_iters) {
        st->done = 1;
        st->exit_code = EXIT_MAX;
      } else if (fabs(st->rhokm) < TINY || fabs(st->omega) < TINY) {
        st->done = 1;
        st->exit_code = EXIT_BREAKDOWN;
      }
      st->rho = tot[2];
    }
  }
}

// ---------------------------------------------------------- Glob_Norm -----
// bicgstab.F90:598-666.  mode 1: sum f (rhs, weights area*dsig(k) applied on
// use); mode 0: sum f*area*dsig(k).
__global__ __launch_bounds__(NSX_BLOCK) void k_gnorm_sum(Geo g, const double *f, double *partials,
                                                          unsigned int *ticket, SolverState *st, int mode) {
  double acc[1] = {0.0};
  FOR_EACH_CELL(g, c) {
    int i, k;
    cell_of(g, c, i, k);
    const size_t s = (size_t)i * g.NZ + k;
    acc[0] += mode == 1 ? f[s] : f[s] * (g.area[i] * g.dsig[k]);
  }
  if (block_partials_and_ticket<1>(acc, partials, ticket)) {
    double tot[1];
    final_sums<1>(tot, partials, ticket);
    if (threadIdx.x == 0) {
      if (mode == 1) st->avec_rhs = tot[0] / (g.TotalVolume * (g.sigFin - g.sigIni));
      else st->xvol = tot[0];
    }
  }
}

// phi -= avec (Glob_Norm(phi,0)) and the Neumann ghost layers of BCpcenter
__global__ __launch_bounds__(NSX_BLOCK) void k_gnorm_apply(Geo g, double *x, const SolverState *st) {
  const double avec = st->xvol / (g.TotalVolume * (g.sigFin - g.sigIni));
  FOR_EACH_CELL(g, c) {
    int i, k;
    cell_of(g, c, i, k);
    const size_t s = (size_t)i * g.NZ + k;
    const double v = x[s] - avec;
    x[s] = v;
    if (k == 1 && g.bcbot == 1) x[s - 1] = v;
    if (k == g.NZ - 2 && g.bctop == 1) x[s + 1] = v;
  }
}

// ------------------------------------------------------------ launchers ---
int launch_bicg_init(const Geo &g, hipStream_t s, PPEOp op, const double *rhs, const double *phi_in, double *x,
                     double *r, double *rhat, Partials part, SolverState *st, double eps, int max_iters) {
  hipLaunchKernelGGL(k_bicg_init, dim3(nsx_grid_for(g.ncell)), dim3(NSX_BLOCK), 0, s, g, op, rhs, phi_in, x, r,
                     rhat, part.buf, part.ticket, st, eps, max_iters);
  return (int)hipGetLastError();
}

int launch_bicg_iter(const Geo &g, hipStream_t s, PPEOp op, int parity, double *x, double *r, const double *rhat,
                     double *p0, double *v0, double *p1, double *v1, double *sv, double *tv, Partials part,
                     SolverState *st, hipEvent_t *ev) {
  const dim3 grid(nsx_grid_for(g.ncell)), blk(NSX_BLOCK);
  double *pold = parity ? p1 : p0, *vold = parity ? v1 : v0;
  double *pnew = parity ? p0 : p1, *vnew = parity ? v0 : v1;
  if (ev) (void)hipEventRecord(ev[0], s);
  hipLaunchKernelGGL(k_bicg_a, grid, blk, 0, s, g, op, r, rhat, pold, vold, pnew, vnew, part.buf, part.ticket, st);
  if (ev) {
    (void)hipEventRecord(ev[1], s);
    (void)hipEventRecord(ev[2], s);
  }
  hipLaunchKernelGGL(k_bicg_b, grid, blk, 0, s, g, op, r, vnew, sv, tv, part.buf, part.ticket + 1, st);
  if (ev) {
    (void)hipEventRecord(ev[3], s);
    (void)hipEventRecord(ev[4], s);
  }
  hipLaunchKernelGGL(k_bicg_c, grid, blk, 0, s, g, op, x, r, rhat, pnew, sv, tv, part.buf, part.ticket + 2, st);
  if (ev) (void)hipEventRecord(ev[5], s);
  return (int)hipGetLastError();
}

int launch_gnorm_sum(const Geo &g, hipStream_t s, const double *x, Partials part, SolverState *st) {
  hipLaunchKernelGGL(k_gnorm_sum, dim3(nsx_grid_for(g.ncell)), dim3(NSX_BLOCK), 0, s, g, x, part.buf, part.ticket,
                     st, 1);
  return (int)hipGetLastError();
}

int launch_gnorm_apply(const Geo &g, hipStream_t s, double *x, const SolverState *st) {
  hipLaunchKernelGGL(k_gnorm_apply, dim3(nsx_grid_for(g.ncell)), dim3(NSX_BLOCK), 0, s, g, x, st);
  return (int)hipGetLastError();
}
