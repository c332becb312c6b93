// nsx_bicgstab.hip -- Jacobi-preconditioned BiCGSTAB on the 7-point prism
// stencil (bicgstab.F90:303-591).  One iteration is five kernels, split at
// the algorithm's three global synchronisation points and at the two
// stencil applications:
//
//   A1  stop test of k-1, p = r + beta (p - omega v), y = p / A0   update_pk :726-773, diag_precdt :831-836
//   A2  v = A y, partial <rhat, v>                                 get_Apk :931-1001
//   B1  alpha = rho / <rhat,v>, s = r - alpha v, z = s / A0        update_sk :1007-1041
//   B2  t = A z, partial <t,s>, <t,t>
//   C   omega = <t,s>/<t,t>, x += alpha y + omega z, r = s - omega t,
//       partial |r|^2, |r|_1, <rhat, r>, sum x Vol                 update_xkrk :1047-1111
//
// Only y and z are gathered at the 7 points, so each stencil reads one
// neighbour array.  The matrix is PDE_center's, regenerated per row from the
// per-edge coefficient ppec, the cell area and the sigma spacings (bit-
// identical to AmC0..AmCB, checked at set-up), so no coefficient array is
// read; y = p / A0 and z = s / A0 divide exactly as diag_precdt does.
//
// Algorithmic HBM bytes per interior prism and iteration (FP64, 8 B each):
//   A1 read r, p, v; write p, y        5      A2 read y, rhat; write v    3
//   B1 read r, v; write s, z           4      B2 read z, s; write t       3
//   C  read x, y, z, s, t, rhat; write x, r                              8
//   total 23 x 8 B = 184 B (neighbour re-reads of y, z are cache hits).
// Host-side timing brackets each of the five kernels with HIP events.
//
// The scalars of a phase come from the partial sums of the phase before it:
// B1 sums A2's partials and forms alpha itself, C forms omega from B2's, and
// A1 of iteration k+1 forms the norms, the stopping test and rho_{k+1} from
// C's (see reduce_partials).  Block 0 records the scalars in SolverState for
// the host and for later phases.  Each kernel exits at once when st->done is
// set, so the host can queue many iterations ahead and poll rarely.
#include "nsx_kernels.h"

#define FOR_EACH_CELL(g, c)                                                         \
  uint32_t beg_, end_;                                                              \
  block_range((g).ncell, beg_, end_);                                               \
  for (uint32_t base_ = beg_; base_ < end_; base_ += NSX_BLOCK)                     \
    for (uint32_t c = base_ + threadIdx.x, once_ = 1; once_ && c < end_; once_ = 0)

#define TINY ((double)1.0e-30f) /* REAL(4) literal of bicgstab.F90:479-517 */

enum { EXIT_CONVERGED = 0, EXIT_NO_SOLVE = 1, EXIT_BREAKDOWN = 2, EXIT_MAX = 3 };

// Row of the PPE matrix at (i,k), regenerated from the geometry exactly as
// PDE_center forms it (PDE.F90:76-109): AA_j = (dle1/dlCC)_j dlVV_j dsigv(k-1)
// (ppec holds the first two factors), AAB = area/dsig(k-1), AAT = area/dsig(k),
// zero at the sigma walls, AA0 = -(AAT+AAB+AA1+AA2+AA3).
struct Row {
  double A0, a1, a2, a3, AT, AB;
};

__device__ inline Row ppe_row(const Geo &g, int i, int k) {
  Row r;
  const double sv = g.dsigv[k - 1], area = g.area[i];
  r.AT = (k == g.NZ - 2 && g.bctop == 1) ? 0.0 : area / g.dsig[k];
  r.AB = (k == 1 && g.bcbot == 1) ? 0.0 : area / g.dsig[k - 1];
  r.a1 = g.ppec[i * 3] * sv;
  r.a2 = g.ppec[i * 3 + 1] * sv;
  r.a3 = g.ppec[i * 3 + 2] * sv;
  r.A0 = -1.0 * (r.AT + r.AB + r.a1 + r.a2 + r.a3);
  return r;
}

// A y at (i,k) (get_Apk, bicgstab.F90:976-998): the vertical ghost of y is
// the Neumann copy of BCpcenter, the periodic ghosts are the partner columns
__device__ inline double spmv_row(const Geo &g, const Row &A, int i, int k, size_t s, const double *y) {
  const double ys = y[s];
  const double yu = (k + 1 <= g.NZ - 2) ? y[s + 1] : ys;
  const double yd = (k - 1 >= 1) ? y[s - 1] : ys;
  const int e = i * 3;
  const double y0 = y[(size_t)g.nbr[e] * g.NZ + k], y1 = y[(size_t)g.nbr[e + 1] * g.NZ + k],
               y2 = y[(size_t)g.nbr[e + 2] * g.NZ + k];
  return A.A0 * ys + A.a1 * y0 + A.a2 * y1 + A.a3 * y2 + A.AT * yu + A.AB * yd;
}

// ------------------------------------------------------------ initial -----
// bicgstab.F90:420-451 after Glob_Norm(rhs,1) (bicgstab.F90:635-663): the
// rhs average comes from the partial sums of k_source / k_gnorm_sum; r = rhat
// = b - A x0 and the partial sums of the norms.
__global__ __launch_bounds__(NSX_BLOCK) void k_bicg_init(Geo g, const double *rhs, const double *phi,
                                                          double *x, double *r, double *rhat, Partials P,
                                                          SolverState *st, double eps, int max_iters) {
  double tsrc[1];
  reduce_partials<1>(P.src, gridDim.x, tsrc);
  const double avec = tsrc[0] / (g.TotalVolume * (g.sigFin - g.sigIni));
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  FOR_EACH_CELL(g, c) {
    int i, k;
    cell_of(g, c, i, k);
    const size_t s = (size_t)i * g.NZ + k;
    const double area = g.area[i];
    const double b = rhs[s] - avec * area * g.dsig[k];
    const Row A = ppe_row(g, i, k);
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
  write_partials<5>(acc, P.init);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->avec = avec;
    st->eps = eps;
    st->max_iters = max_iters;
    st->b_alpha = 1.0;
    st->c_omega = 1.0;
    st->a_rho[0] = st->a_rho[1] = 1.0;
    st->c_iter = 0;
    st->a_iter = 0;
    st->done = 0;
    st->exit_code = EXIT_CONVERGED;
  }
}

// ------------------------------------------------------------ phase A1 ----
// Closes iteration k-1 (norms and the stopping test of bicgstab.F90:544-575,
// the breakdown guards of :479-490; at k = 0 the norms of :453-469), then
// update_pk (:726-773) p = r + beta (p - omega v) (p = r at k = 0) and
// diag_precdt Opt 1 (:831-836) y = p / AmC0.  Point-wise, in place.
__global__ __launch_bounds__(NSX_BLOCK) void k_bicg_a1(Geo g, const double *r, double *p, const double *v,
                                                        double *y, Partials P, SolverState *st) {
  if (st->done) return;
  const int k_it = st->c_iter;
  const bool first = k_it == 0;
  double rho, e1, e2, xvol, b1, b2;
  int stop = 0, code = EXIT_CONVERGED;
  if (first) {
    double t[5];
    reduce_partials<5>(P.init, gridDim.x, t);
    b2 = sqrt(t[0]);
    const double r2 = sqrt(t[1]);
    b1 = t[2];
    const double r1 = t[3];
    if (b2 == 0) b2 = 1.0;
    if (b1 == 0) b1 = 1.0;
    e2 = r2 / b2;
    e1 = r1 / b1;
    rho = t[1];  // <rhat, r0> with rhat = r0
    xvol = t[4];
    if (e1 < st->eps || e2 < st->eps) {
      stop = 1;
      code = EXIT_NO_SOLVE;
    }
  } else {
    double t[4];
    reduce_partials<4>(P.c, gridDim.x, t);
    b1 = st->b1;
    b2 = st->b2;
    const double r2 = sqrt(t[0]), r1 = t[1];
    e2 = r2 / b2;
    e1 = r1 / b1;
    rho = t[2];
    xvol = t[3];
    double enorm = (double)1.0e10f;
    if (enorm > e1) enorm = e1;
    if (enorm > e2) enorm = e2;
    const double rhokm = st->a_rho[(k_it - 1) & 1];
    if (enorm < st->eps) {
      stop = 1;
      code = EXIT_CONVERGED;
    } else if (k_it > st->max_iters) {
      stop = 1;
      code = EXIT_MAX;
    } else if (fabs(rhokm) < TINY || fabs(st->c_omega) < TINY) {
      stop = 1;
      code = EXIT_BREAKDOWN;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->a_iter = k_it;
    st->a_rho[k_it & 1] = rho;
    st->e1 = e1;
    st->e2 = e2;
    st->xvol = xvol;
    if (first) {
      st->b1 = b1;
      st->b2 = b2;
    }
    if (stop) {
      st->exit_code = code;
      st->done = 1;
    }
  }
  if (stop) return;
  const double om = st->c_omega;
  const double beta = first ? 0.0 : (rho / st->a_rho[(k_it - 1) & 1]) * (st->b_alpha / om);
  FOR_EACH_CELL(g, c) {
    int i, k;
    cell_of(g, c, i, k);
    const size_t s = (size_t)i * g.NZ + k;
    const double pn = first ? r[s] : r[s] + beta * (p[s] - om * v[s]);
    p[s] = pn;
    y[s] = pn / ppe_row(g, i, k).A0;
  }
}

// ------------------------------------------------------------ phase A2 ----
// get_Apk (:931-1001) v = A y, and the partial sums of <rhat, v> (:509)
__global__ __launch_bounds__(NSX_BLOCK) void k_bicg_a2(Geo g, const double *y, const double *rhat, double *v,
                                                        Partials P, const SolverState *st) {
  if (st->done) return;
  double acc[1] = {0.0};
  FOR_EACH_CELL(g, c) {
    int i, k;
    cell_of(g, c, i, k);
    const size_t s = (size_t)i * g.NZ + k;
    const double vs = spmv_row(g, ppe_row(g, i, k), i, k, s, y);
    v[s] = vs;
    acc[0] += rhat[s] * vs;
  }
  write_partials<1>(acc, P.a);
}

// ------------------------------------------------------------ phase B1 ----
// alpha = rho / <rhat, v> (breakdown guard of :512-517), update_sk
// (:1007-1041) s = r - alpha v, diag_precdt z = s / AmC0.  Point-wise.
__global__ __launch_bounds__(NSX_BLOCK) void k_bicg_b1(Geo g, const double *r, const double *v, double *sv,
                                                        double *z, Partials P, SolverState *st) {
  if (st->done) return;
  double t[1];
  reduce_partials<1>(P.a, gridDim.x, t);
  const int k_it = st->a_iter;
  if (fabs(t[0]) < TINY) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->exit_code = EXIT_BREAKDOWN;
      st->done = 1;
    }
    return;
  }
  const double al = st->a_rho[k_it & 1] / t[0];
  if (blockIdx.x == 0 && threadIdx.x == 0) st->b_alpha = al;
  FOR_EACH_CELL(g, c) {
    int i, k;
    cell_of(g, c, i, k);
    const size_t s = (size_t)i * g.NZ + k;
    const double ss = r[s] - al * v[s];
    sv[s] = ss;
    z[s] = ss / ppe_row(g, i, k).A0;
  }
}

// ------------------------------------------------------------ phase B2 ----
// get_Apk t = A z and the partial sums of <t,s>, <t,t> (:535-536)
__global__ __launch_bounds__(NSX_BLOCK) void k_bicg_b2(Geo g, const double *z, const double *sv, double *tv,
                                                        Partials P, const SolverState *st) {
  if (st->done) return;
  double acc[2] = {0.0, 0.0};
  FOR_EACH_CELL(g, c) {
    int i, k;
    cell_of(g, c, i, k);
    const size_t s = (size_t)i * g.NZ + k;
    const double ts = spmv_row(g, ppe_row(g, i, k), i, k, s, z);
    tv[s] = ts;
    acc[0] += ts * sv[s];
    acc[1] += ts * ts;
  }
  write_partials<2>(acc, P.b);
}

// ------------------------------------------------------------- phase C ----
// omega = <t,s>/<t,t>, update_xkrk (:1047-1111) x += alpha y + omega z,
// r = s - omega t, partial |r|^2, |r|_1, <rhat, r>, sum x Vol
__global__ __launch_bounds__(NSX_BLOCK) void k_bicg_c(Geo g, double *x, double *r, const double *rhat,
                                                       const double *y, const double *z, const double *sv,
                                                       const double *tv, Partials P, SolverState *st) {
  if (st->done) return;
  double t[2];
  reduce_partials<2>(P.b, gridDim.x, t);
  const double al = st->b_alpha, om = t[0] / t[1];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->c_omega = om;
    st->c_iter = st->a_iter + 1;
  }
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  FOR_EACH_CELL(g, c) {
    int i, k;
    cell_of(g, c, i, k);
    const size_t s = (size_t)i * g.NZ + k;
    const double xn = x[s] + al * y[s] + om * z[s];
    const double rr = sv[s] - om * tv[s];
    x[s] = xn;
    r[s] = rr;
    acc[0] += rr * rr;
    acc[1] += fabs(rr);
    acc[2] += rhat[s] * rr;
    acc[3] += xn * (g.area[i] * g.dsig[k]);
  }
  write_partials<4>(acc, P.c);
}

// ---------------------------------------------------------- Glob_Norm -----
// bicgstab.F90:598-666: partial sums of f (the rhs, tags = 1)
__global__ __launch_bounds__(NSX_BLOCK) void k_gnorm_sum(Geo g, const double *f, double *P) {
  double acc[1] = {0.0};
  FOR_EACH_CELL(g, c) {
    int i, k;
    cell_of(g, c, i, k);
    acc[0] += f[(size_t)i * g.NZ + k];
  }
  write_partials<1>(acc, P);
}

// phi -= sum(phi Vol)/(TotalVolume (sigFin-sigIni)) (Glob_Norm(phi,0)) and
// the Neumann ghost layers of BCpcenter
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
int launch_bicg_init(const Geo &g, hipStream_t s, const double *rhs, const double *phi_in, double *x, double *r,
                     double *rhat, Partials part, SolverState *st, double eps, int max_iters) {
  hipLaunchKernelGGL(k_bicg_init, dim3(nsx_grid_for(g.ncell)), dim3(NSX_BLOCK), 0, s, g, rhs, phi_in, x, r,
                     rhat, part, st, eps, max_iters);
  return (int)hipGetLastError();
}

int launch_bicg_iter(const Geo &g, hipStream_t s, double *x, double *r, const double *rhat, double *p, double *v,
                     double *y, double *z, double *sv, double *tv, Partials part, SolverState *st, hipEvent_t *ev) {
  const dim3 grid(nsx_grid_for(g.ncell)), blk(NSX_BLOCK);
  if (ev) (void)hipEventRecord(ev[0], s);
  hipLaunchKernelGGL(k_bicg_a1, grid, blk, 0, s, g, r, p, v, y, part, st);
  if (ev) (void)hipEventRecord(ev[1], s);
  hipLaunchKernelGGL(k_bicg_a2, grid, blk, 0, s, g, y, rhat, v, part, st);
  if (ev) (void)hipEventRecord(ev[2], s);
  hipLaunchKernelGGL(k_bicg_b1, grid, blk, 0, s, g, r, v, sv, z, part, st);
  if (ev) (void)hipEventRecord(ev[3], s);
  hipLaunchKernelGGL(k_bicg_b2, grid, blk, 0, s, g, z, sv, tv, part, st);
  if (ev) (void)hipEventRecord(ev[4], s);
  hipLaunchKernelGGL(k_bicg_c, grid, blk, 0, s, g, x, r, rhat, y, z, sv, tv, part, st);
  if (ev) (void)hipEventRecord(ev[5], s);
  return (int)hipGetLastError();
}

// the regenerated diagonal AmC0 for the set-up check against PDE_center's
__global__ __launch_bounds__(NSX_BLOCK) void k_ppe_diag(Geo g, double *A0) {
  FOR_EACH_CELL(g, c) {
    int i, k;
    cell_of(g, c, i, k);
    A0[(size_t)i * g.NZ + k] = ppe_row(g, i, k).A0;
  }
}

int launch_ppe_diag(const Geo &g, hipStream_t s, double *A0) {
  hipLaunchKernelGGL(k_ppe_diag, dim3(nsx_grid_for(g.ncell)), dim3(NSX_BLOCK), 0, s, g, A0);
  return (int)hipGetLastError();
}

int launch_gnorm_sum(const Geo &g, hipStream_t s, const double *x, Partials part) {
  hipLaunchKernelGGL(k_gnorm_sum, dim3(nsx_grid_for(g.ncell)), dim3(NSX_BLOCK), 0, s, g, x, part.src);
  return (int)hipGetLastError();
}

int launch_gnorm_apply(const Geo &g, hipStream_t s, double *x, const SolverState *st) {
  hipLaunchKernelGGL(k_gnorm_apply, dim3(nsx_grid_for(g.ncell)), dim3(NSX_BLOCK), 0, s, g, x, st);
  return (int)hipGetLastError();
}
