// nsx_bicgstab.hip -- Jacobi-preconditioned BiCGSTAB on the 7-point prism
// stencil (bicgstab.F90:303-591), in FP64 (the reference's arithmetic) and
// as the FP32 inner solver of an FP64 iterative refinement.
//
// One iteration is five kernels, split at the algorithm's three global
// synchronisation points and at the two stencil applications:
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
// Work decomposition: a thread owns four consecutive layers k0..k0+3 of one
// column (k0 a multiple of 4, the column stride a multiple of 4), so every
// field access is a 16-byte-aligned vector load (2x double2 or 1x float4),
// the column data (area, ppec, neighbours) is fetched once per four cells,
// and the five divisions area/dsig(k0-1..k0+3) give AB and AT of all four
// rows.  The sigma spacings of the block live in LDS.  Slots outside the
// interior layers (ghost layers, padding) are written as zero in the work
// vectors and left unchanged in x.  Cell-wise arithmetic keeps the
// reference's operation order, so in FP64 every element is what the
// reference computes from the same scalars.
//
// Algorithmic HBM bytes per interior prism and iteration (b = 8 FP64, 4 FP32):
//   A1 read r, p, v; write p, y        5b     A2 read y, rhat; write v    3b
//   B1 read r, v; write s, z           4b     B2 read z, s; write t       3b
//   C  read x, y, z, s, t, rhat; write x, r                              8b
//   total 23b (neighbour re-reads of y, z are cache hits).
//
// The scalars of a phase come from the partial sums of the phase before it:
// B1 sums A2's partials and forms alpha itself, C forms omega from B2's, and
// A1 of iteration k+1 forms the norms, the stopping test and rho_{k+1} from
// C's (see reduce_partials).  Block 0 records the scalars in SolverState for
// the host and for later phases.  Each kernel exits at once when st->done is
// set, so the host can queue many iterations ahead and poll rarely.  Inner
// products and norms are accumulated in FP64 in both precisions.
#include <hip/hip_ext.h>

#include "nsx_kernels.h"

// With timing on, a kernel is launched with its own start/stop events
// (recorded by the dispatch itself, so the elapsed time is the kernel's)
#define LAUNCH_TIMED(ev, q, kern, grid, blk, lds, s, ...)                                         \
  do {                                                                                          \
    if (ev)                                                                                     \
      hipExtLaunchKernelGGL(kern, grid, blk, lds, s, (ev)[2 * (q)], (ev)[2 * (q) + 1], 0, __VA_ARGS__); \
    else                                                                                        \
      hipLaunchKernelGGL(kern, grid, blk, lds, s, __VA_ARGS__);                                 \
  } while (0)

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

// ------------------------------------------------- four-layer helpers -----
template <typename T>
struct Prec;
template <>
struct Prec<double> {
  __device__ static const double *area(const Geo &g) { return g.area; }
  __device__ static const double *ppec(const Geo &g) { return g.ppec; }
  __device__ static const double *dsig(const Geo &g) { return g.dsig; }
  __device__ static const double *dsigv(const Geo &g) { return g.dsigv; }
};
template <>
struct Prec<float> {
  __device__ static const float *area(const Geo &g) { return g.area_f; }
  __device__ static const float *ppec(const Geo &g) { return g.ppec_f; }
  __device__ static const float *dsig(const Geo &g) { return g.dsig_f; }
  __device__ static const float *dsigv(const Geo &g) { return g.dsigv_f; }
};

__device__ inline void ld4(const double *p, double (&o)[4]) {
  const double2 a = *(const double2 *)p, b = *(const double2 *)(p + 2);
  o[0] = a.x;
  o[1] = a.y;
  o[2] = b.x;
  o[3] = b.y;
}
__device__ inline void ld4(const float *p, float (&o)[4]) {
  const float4 a = *(const float4 *)p;
  o[0] = a.x;
  o[1] = a.y;
  o[2] = a.z;
  o[3] = a.w;
}
__device__ inline void st4(double *p, const double (&o)[4]) {
  *(double2 *)p = make_double2(o[0], o[1]);
  *(double2 *)(p + 2) = make_double2(o[2], o[3]);
}
__device__ inline void st4(float *p, const float (&o)[4]) { *(float4 *)p = make_float4(o[0], o[1], o[2], o[3]); }

// sigma spacings of the block in LDS: lv[0..NS) = dsig, lv[NS..2NS) = dsigv
// (padding slots hold 1 so that no masked-out row divides by zero)
template <typename T>
__device__ inline T *load_levels(const Geo &g) {
  extern __shared__ __align__(16) unsigned char nsx_dyn_lds[];
  T *lv = reinterpret_cast<T *>(nsx_dyn_lds);
  const T *ds = Prec<T>::dsig(g), *dv = Prec<T>::dsigv(g);
  for (int k = threadIdx.x; k < g.NS; k += NSX_BLOCK) {
    lv[k] = k < g.NZ ? ds[k] : T(1);
    lv[g.NS + k] = k < g.NZ - 1 ? dv[k] : T(1);
  }
  __syncthreads();
  return lv;
}

// the rows of layers k0..k0+3 of column i; rows outside the interior are zero
template <typename T>
struct Rows4 {
  T A0[4], a1[4], a2[4], a3[4], AT[4], AB[4];
  bool in[4];
};

template <typename T>
__device__ inline void rows4(const Geo &g, const T *lv, int i, int k0, Rows4<T> &R) {
  const T area = Prec<T>::area(g)[i];
  const T *pc = Prec<T>::ppec(g) + i * 3;
  const T c1 = pc[0], c2 = pc[1], c3 = pc[2];
  T q[5];  // q[u] = area / dsig(k0-1+u)
#pragma unroll
  for (int u = 0; u < 5; ++u) {
    const int k = k0 - 1 + u;
    q[u] = area / lv[k < 0 ? 0 : k];
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int k = k0 + u;
    R.in[u] = k >= 1 && k <= g.NZ - 2;
    const T sv = lv[g.NS + (k >= 1 ? k - 1 : 0)];
    R.AT[u] = (k == g.NZ - 2 && g.bctop == 1) ? T(0) : q[u + 1];
    R.AB[u] = (k == 1 && g.bcbot == 1) ? T(0) : q[u];
    R.a1[u] = c1 * sv;
    R.a2[u] = c2 * sv;
    R.a3[u] = c3 * sv;
    R.A0[u] = T(-1) * (R.AT[u] + R.AB[u] + R.a1[u] + R.a2[u] + R.a3[u]);
  }
}

// out = A y on layers k0..k0+3 of column i (get_Apk, bicgstab.F90:976-998):
// the vertical ghost of y is the Neumann copy of BCpcenter, the periodic
// ghosts are the partner columns
template <typename T>
__device__ inline void spmv4(const Geo &g, const Rows4<T> &R, int i, int k0, size_t s0, const T *y, T (&out)[4]) {
  T yc[4], y0[4], y1[4], y2[4];
  ld4(y + s0, yc);
  const T ym = k0 >= 1 ? y[s0 - 1] : T(0);
  const T yp = (k0 + 4 <= g.NZ - 2) ? y[s0 + 4] : T(0);
  const int e = i * 3;
  ld4(y + (size_t)g.nbr[e] * g.NS + k0, y0);
  ld4(y + (size_t)g.nbr[e + 1] * g.NS + k0, y1);
  ld4(y + (size_t)g.nbr[e + 2] * g.NS + k0, y2);
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int k = k0 + u;
    const T ys = yc[u];
    const T yu = (k + 1 <= g.NZ - 2) ? (u < 3 ? yc[u < 3 ? u + 1 : 3] : yp) : ys;
    const T yd = (k - 1 >= 1) ? (u > 0 ? yc[u > 0 ? u - 1 : 0] : ym) : ys;
    out[u] = R.in[u] ? R.A0[u] * ys + R.a1[u] * y0[u] + R.a2[u] * y1[u] + R.a3[u] * y2[u] + R.AT[u] * yu +
                           R.AB[u] * yd
                     : T(0);
  }
}

template <class F>
__device__ inline void for_each_quad(const Geo &g, F &&f) {
  uint32_t beg, end;
  block_range(g.nquad, beg, end);
  for (uint32_t base = beg; base < end; base += NSX_BLOCK) {
    const uint32_t t = base + threadIdx.x;
    if (t < end) {
      const uint32_t i = fdiv(t, g.divNQ);
      const int k0 = 4 * (int)(t - i * (uint32_t)g.NQ);
      f((int)i, k0, (size_t)i * g.NS + k0);
    }
  }
}

// ------------------------------------------------------------ initial -----
// bicgstab.F90:420-451 after Glob_Norm(rhs,1) (bicgstab.F90:635-663): the
// rhs average comes from the partial sums of k_source / k_gnorm_sum; r = b -
// A x0 in FP64 with the partial sums |b|^2, |r|^2, |b|_1, |r|_1, sum x0 Vol
// into `outer`.  FP64 solve (TO = double): r, rhat = r, x = x0.  Refinement
// pass of the mixed solve (TO = float): r32 = rhat32 = (float) r, the inner
// solution e = 0, and the inner norms |r32|^2, |r32|_1 into `inner` laid out
// as the initial partials of an FP32 solve with zero start.
template <typename TO>
__global__ __launch_bounds__(NSX_BLOCK) void k_resid(Geo g, const double *rhs, const double *phi, double *x,
                                                      TO *r, TO *rhat, TO *e, Partials P, double *outer,
                                                      double *inner, int copy_x, SolverState *st) {
  double tsrc[1];
  reduce_partials<1>(P.src, gridDim.x, tsrc);
  const double avec = tsrc[0] / (g.TotalVolume * (g.sigFin - g.sigIni));
  if (blockIdx.x == 0 && threadIdx.x == 0) st->avec = avec;
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  double ai[2] = {0.0, 0.0};
  FOR_EACH_CELL(g, c) {
    int i, k;
    cell_of(g, c, i, k);
    const size_t s = (size_t)i * g.NS + k;
    const double area = g.area[i];
    const double b = rhs[s] - avec * area * g.dsig[k];
    const Row A = ppe_row(g, i, k);
    const int ee = i * 3;
    const double ph = phi[s];
    const double Ax = A.A0 * ph + A.a1 * phi[(size_t)g.nbr[ee] * g.NS + k] +
                      A.a2 * phi[(size_t)g.nbr[ee + 1] * g.NS + k] + A.a3 * phi[(size_t)g.nbr[ee + 2] * g.NS + k] +
                      A.AT * phi[s + 1] + A.AB * phi[s - 1];
    const double res = b - Ax;
    const TO ro = (TO)res;
    r[s] = ro;
    rhat[s] = ro;
    if (e) e[s] = TO(0);
    if (copy_x) x[s] = ph;
    acc[0] += b * b;
    acc[1] += res * res;
    acc[2] += fabs(b);
    acc[3] += fabs(res);
    acc[4] += ph * (area * g.dsig[k]);
    if (inner) {
      const double rd = (double)ro;
      ai[0] += rd * rd;
      ai[1] += fabs(rd);
    }
  }
  write_partials<5>(acc, outer);
  __syncthreads();  // write_partials' LDS is reused below
  if (inner) {
    double in5[5] = {ai[0], ai[0], ai[1], ai[1], 0.0};
    // the block sums of equal inputs are equal, so rows 0/1 and 2/3 agree
    write_partials<5>(in5, inner);
  }
}

// resets the iteration scalars of an FP64 solve (block 0 of the first kernel)
__device__ inline void reset_state(SolverState *st, double eps, int max_iters) {
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

__global__ __launch_bounds__(NSX_BLOCK) void k_bicg_init_state(SolverState *st, double eps, int max_iters,
                                                                double delta) {
  if (threadIdx.x == 0) {
    reset_state(st, eps, max_iters);
    st->mp_pass = 0;
    st->mp_done = 0;
    st->delta = delta;
    st->upd = 0;
  }
}

// ------------------------------------------------- reliable update (U) ----
// Mixed precision with the Krylov space kept: the FP32 iteration runs on,
// and after iteration k-1 every block sums C's partials and takes the same
// decision (Sleijpen & van der Vorst's reliable updates, in the form of
// Clark et al. 2010): when |r| has fallen below delta times the largest
// |r| since the last update, or the recursive residual meets the stopping
// rule, the true residual of the FP64 solution x plus the FP32 correction e
// is formed in FP64, r = b - A (x + e), and replaces the recursive one
// (stored in FP32), with its partial sums into P.u.  A1 of the same
// iteration then folds e into x and tests the true residual; C restarts e
// from zero.  x and e are only read here, so the stencil sees one state.
__global__ __launch_bounds__(NSX_BLOCK) void k_rel_u(Geo g, const double *rhs, const double *x, const float *e,
                                                      float *r, const float *rhat, Partials P, SolverState *st) {
  if (st->done) return;
  const int k_it = st->c_iter;
  if (k_it == 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) st->upd = 0;
    return;
  }
  double t[4];
  reduce_partials<4>(P.c, gridDim.x, t);
  const double rN = sqrt(t[0]), r1 = t[1];
  const double e1 = r1 / st->b1, e2 = rN / st->b2;
  const double mrr = fmax(st->maxrr[(k_it - 1) & 1], rN);
  const double r0N = st->r0N;
  const bool conv = e1 < st->eps || e2 < st->eps;
  const bool upd = conv || (rN < st->delta * mrr && r0N <= mrr);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->upd = upd ? 1 : 0;
    st->maxrr[k_it & 1] = mrr;
  }
  if (!upd) return;
  const double avec = st->avec;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  FOR_EACH_CELL(g, c) {
    int i, k;
    cell_of(g, c, i, k);
    const size_t s = (size_t)i * g.NS + k;
    const double area = g.area[i];
    const double b = rhs[s] - avec * area * g.dsig[k];
    const Row A = ppe_row(g, i, k);
    const int ee = i * 3;
    const size_t n0 = (size_t)g.nbr[ee] * g.NS + k, n1 = (size_t)g.nbr[ee + 1] * g.NS + k,
                 n2 = (size_t)g.nbr[ee + 2] * g.NS + k;
    const double xs = x[s] + (double)e[s];
    const double Ax = A.A0 * xs + A.a1 * (x[n0] + (double)e[n0]) + A.a2 * (x[n1] + (double)e[n1]) +
                      A.a3 * (x[n2] + (double)e[n2]) + A.AT * (x[s + 1] + (double)e[s + 1]) +
                      A.AB * (x[s - 1] + (double)e[s - 1]);
    const double res = b - Ax;
    const float ro = (float)res;
    r[s] = ro;
    acc[0] += res * res;
    acc[1] += fabs(res);
    acc[2] += (double)rhat[s] * (double)ro;
    acc[3] += xs * (area * g.dsig[k]);
  }
  write_partials<4>(acc, P.u);
}

// ------------------------------------------------------------ phase A1 ----
// Closes iteration k-1 (norms and the stopping test of bicgstab.F90:544-575,
// the breakdown guards of :479-490; at k = 0 the norms of :453-469), then
// update_pk (:726-773) p = r + beta (p - omega v) (p = r at k = 0) and
// diag_precdt Opt 1 (:831-836) y = p / AmC0.  Point-wise, in place.
template <typename T>
__global__ __launch_bounds__(NSX_BLOCK) void k_bicg_a1(Geo g, const T *r, T *p, const T *v, T *y, Partials P,
                                                        SolverState *st, double *xfold, const T *efold) {
  if (st->done) return;
  const int k_it = st->c_iter;
  const bool first = k_it == 0;
  const bool upd = xfold && !first && st->upd;
  double rho, e1, e2, xvol, b1, b2;
  int stop = 0, code = EXIT_CONVERGED;
  if (upd) {
    // reliable update: the true residual of x + e (k_rel_u); fold e into x
    double t[4];
    reduce_partials<4>(P.u, gridDim.x, t);
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
    const bool finite = isfinite(t[0]) && isfinite(t[1]) && isfinite(t[2]) && isfinite(t[3]);
    if (!finite) {
      // the FP32 correction has overflowed: leave x as it is
      stop = 1;
      code = EXIT_BREAKDOWN;
    } else if (enorm < st->eps) {
      stop = 1;
      code = EXIT_CONVERGED;
    } else if (k_it > st->max_iters) {
      stop = 1;
      code = EXIT_MAX;
    } else if (fabs(rhokm) < TINY || fabs(st->c_omega) < TINY) {
      stop = 1;
      code = EXIT_BREAKDOWN;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->r0N = r2;
      st->maxrr[k_it & 1] = r2;
      st->mp_pass += 1;
    }
    if (finite) {
      FOR_EACH_CELL(g, c) {
        int i, k;
        cell_of(g, c, i, k);
        const size_t s = (size_t)i * g.NS + k;
        xfold[s] = xfold[s] + (double)efold[s];
      }
    }
  } else if (first) {
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
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->r0N = r2;
      st->maxrr[0] = r2;
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
    } else if (!isfinite(enorm) || !isfinite(rho) || !isfinite(st->c_omega)) {
      // no counterpart in the reference (whose FP64 iterates do not
      // overflow): an overflowed FP32 iteration stops as a breakdown
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
  const T om = (T)st->c_omega;
  const T beta = first ? T(0) : (T)((rho / st->a_rho[(k_it - 1) & 1]) * (st->b_alpha / st->c_omega));
  const T *lv = load_levels<T>(g);
  for_each_quad(g, [&](int i, int k0, size_t s0) {
    Rows4<T> R;
    rows4<T>(g, lv, i, k0, R);
    T rr[4], pp[4], vv[4], yy[4];
    ld4(r + s0, rr);
    ld4(p + s0, pp);
    ld4(v + s0, vv);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const T pn = first ? rr[u] : rr[u] + beta * (pp[u] - om * vv[u]);
      pp[u] = R.in[u] ? pn : T(0);
      yy[u] = R.in[u] ? pn / R.A0[u] : T(0);
    }
    st4(p + s0, pp);
    st4(y + s0, yy);
  });
}

// ------------------------------------------------------------ phase A2 ----
// get_Apk (:931-1001) v = A y, and the partial sums of <rhat, v> (:509)
template <typename T>
__global__ __launch_bounds__(NSX_BLOCK) void k_bicg_a2(Geo g, const T *y, const T *rhat, T *v, Partials P,
                                                        const SolverState *st) {
  if (st->done) return;
  const T *lv = load_levels<T>(g);
  double acc[1] = {0.0};
  for_each_quad(g, [&](int i, int k0, size_t s0) {
    Rows4<T> R;
    rows4<T>(g, lv, i, k0, R);
    T vs[4], rh[4];
    spmv4<T>(g, R, i, k0, s0, y, vs);
    ld4(rhat + s0, rh);
    st4(v + s0, vs);
#pragma unroll
    for (int u = 0; u < 4; ++u) acc[0] += (double)rh[u] * (double)vs[u];
  });
  write_partials<1>(acc, P.a);
}

// ------------------------------------------------------------ phase B1 ----
// alpha = rho / <rhat, v> (breakdown guard of :512-517), update_sk
// (:1007-1041) s = r - alpha v, diag_precdt z = s / AmC0.  Point-wise.
template <typename T>
__global__ __launch_bounds__(NSX_BLOCK) void k_bicg_b1(Geo g, const T *r, const T *v, T *sv, T *z, Partials P,
                                                        SolverState *st) {
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
  if (!isfinite(al)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->exit_code = EXIT_BREAKDOWN;
      st->done = 1;
    }
    return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) st->b_alpha = al;
  const T alt = (T)al;
  const T *lv = load_levels<T>(g);
  for_each_quad(g, [&](int i, int k0, size_t s0) {
    Rows4<T> R;
    rows4<T>(g, lv, i, k0, R);
    T rr[4], vv[4], ss[4], zz[4];
    ld4(r + s0, rr);
    ld4(v + s0, vv);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const T sn = rr[u] - alt * vv[u];
      ss[u] = R.in[u] ? sn : T(0);
      zz[u] = R.in[u] ? sn / R.A0[u] : T(0);
    }
    st4(sv + s0, ss);
    st4(z + s0, zz);
  });
}

// ------------------------------------------------------------ phase B2 ----
// get_Apk t = A z and the partial sums of <t,s>, <t,t> (:535-536)
template <typename T>
__global__ __launch_bounds__(NSX_BLOCK) void k_bicg_b2(Geo g, const T *z, const T *sv, T *tv, Partials P,
                                                        const SolverState *st) {
  if (st->done) return;
  const T *lv = load_levels<T>(g);
  double acc[2] = {0.0, 0.0};
  for_each_quad(g, [&](int i, int k0, size_t s0) {
    Rows4<T> R;
    rows4<T>(g, lv, i, k0, R);
    T ts[4], ss[4];
    spmv4<T>(g, R, i, k0, s0, z, ts);
    ld4(sv + s0, ss);
    st4(tv + s0, ts);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double td = (double)ts[u];
      acc[0] += td * (double)ss[u];
      acc[1] += td * td;
    }
  });
  write_partials<2>(acc, P.b);
}

// ------------------------------------------------------------- phase C ----
// omega = <t,s>/<t,t>, update_xkrk (:1047-1111) x += alpha y + omega z,
// r = s - omega t, partial |r|^2, |r|_1, <rhat, r>, sum x Vol
template <typename T>
__global__ __launch_bounds__(NSX_BLOCK) void k_bicg_c(Geo g, T *x, T *r, const T *rhat, const T *y, const T *z,
                                                       const T *sv, const T *tv, Partials P, SolverState *st) {
  if (st->done) return;
  double t[2];
  reduce_partials<2>(P.b, gridDim.x, t);
  const double al = st->b_alpha, om = t[0] / t[1];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->c_omega = om;
    st->c_iter = st->a_iter + 1;
  }
  const T alt = (T)al, omt = (T)om;
  const T *lv = load_levels<T>(g);
  const T *area = Prec<T>::area(g);
  // after a reliable update A1 has folded the correction into the FP64
  // solution: the FP32 correction starts again from zero
  const bool fresh = st->upd != 0;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for_each_quad(g, [&](int i, int k0, size_t s0) {
    T xx[4], rr[4], rh[4], yy[4], zz[4], ss[4], tt[4];
    if (fresh) {
      xx[0] = xx[1] = xx[2] = xx[3] = T(0);
    } else {
      ld4(x + s0, xx);
    }
    ld4(rhat + s0, rh);
    ld4(y + s0, yy);
    ld4(z + s0, zz);
    ld4(sv + s0, ss);
    ld4(tv + s0, tt);
    const T ar = area[i];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int k = k0 + u;
      const bool in = k >= 1 && k <= g.NZ - 2;
      const T xn = xx[u] + alt * yy[u] + omt * zz[u];
      const T rn = ss[u] - omt * tt[u];
      xx[u] = in ? xn : xx[u];
      rr[u] = in ? rn : T(0);
      if (in) {
        const double rd = (double)rn;
        acc[0] += rd * rd;
        acc[1] += fabs(rd);
        acc[2] += (double)rh[u] * rd;
        acc[3] += (double)xn * (double)(ar * lv[k]);
      }
    }
    st4(x + s0, xx);
    st4(r + s0, rr);
  });
  write_partials<4>(acc, P.c);
}

// ---------------------------------------------------------- refinement ----
// After a refinement residual: the outer norms (true FP64 residual of the
// current x against the mean-removed rhs, the reference's stopping rule
// min(|r|_1/|b|_1, |r|_2/|b|_2) < eps), the outer stop, and the reset of
// the inner FP32 solve.  One block.
__global__ __launch_bounds__(NSX_BLOCK) void k_mp_check(const double *outer, uint32_t G, SolverState *st, double eps,
                                                         double eps_inner, int max_iters, int max_passes) {
  double t[5];
  reduce_partials<5>(outer, G, t);
  if (threadIdx.x != 0) return;
  const int pass = st->mp_pass;
  double b2 = sqrt(t[0]), b1 = t[2];
  if (pass == 0) {
    if (b2 == 0) b2 = 1.0;
    if (b1 == 0) b1 = 1.0;
    st->ob1 = b1;
    st->ob2 = b2;
  } else {
    b1 = st->ob1;
    b2 = st->ob2;
  }
  const double e2 = sqrt(t[1]) / b2, e1 = t[3] / b1;
  st->oe1 = e1;
  st->oe2 = e2;
  st->xvol = t[4];
  const bool conv = e1 < eps || e2 < eps;
  reset_state(st, eps_inner, max_iters);
  if (conv || pass >= max_passes) {
    st->mp_done = 1;
    st->done = 1;  // the queued inner kernels return at once
    st->exit_code = conv ? (pass == 0 ? EXIT_NO_SOLVE : EXIT_CONVERGED) : EXIT_MAX;
  }
  st->mp_pass = pass + 1;
}

// x += e over the interior (the FP32 correction of one refinement pass)
__global__ __launch_bounds__(NSX_BLOCK) void k_mp_update(Geo g, double *x, const float *e) {
  FOR_EACH_CELL(g, c) {
    int i, k;
    cell_of(g, c, i, k);
    const size_t s = (size_t)i * g.NS + k;
    x[s] = x[s] + (double)e[s];
  }
}

// ---------------------------------------------------------- Glob_Norm -----
// bicgstab.F90:598-666: partial sums of f (the rhs, tags = 1)
__global__ __launch_bounds__(NSX_BLOCK) void k_gnorm_sum(Geo g, const double *f, double *P) {
  double acc[1] = {0.0};
  FOR_EACH_CELL(g, c) {
    int i, k;
    cell_of(g, c, i, k);
    acc[0] += f[(size_t)i * g.NS + k];
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
    const size_t s = (size_t)i * g.NS + k;
    const double v = x[s] - avec;
    x[s] = v;
    if (k == 1 && g.bcbot == 1) x[s - 1] = v;
    if (k == g.NZ - 2 && g.bctop == 1) x[s + 1] = v;
  }
}

// ------------------------------------------------------------ launchers ---
static size_t level_lds(const Geo &g, size_t elem) { return 2 * (size_t)g.NS * elem; }

int launch_bicg_init(const Geo &g, hipStream_t s, const double *rhs, const double *phi_in, double *x, double *r,
                     double *rhat, Partials part, SolverState *st, double eps, int max_iters) {
  const dim3 grid(g.grid), blk(NSX_BLOCK);
  hipLaunchKernelGGL(k_bicg_init_state, dim3(1), blk, 0, s, st, eps, max_iters, 0.0);
  hipLaunchKernelGGL(k_resid<double>, grid, blk, 0, s, g, rhs, phi_in, x, r, rhat, (double *)nullptr, part,
                     part.init, (double *)nullptr, 1, st);
  return (int)hipGetLastError();
}

template <typename T>
static int iter_impl(const Geo &g, hipStream_t s, T *x, T *r, const T *rhat, T *p, T *v, T *y, T *z, T *sv, T *tv,
                     Partials part, SolverState *st, hipEvent_t *ev) {
  const dim3 grid(g.grid), blk(NSX_BLOCK);
  const size_t lds = level_lds(g, sizeof(T));
  LAUNCH_TIMED(ev, 0, k_bicg_a1<T>, grid, blk, lds, s, g, (const T *)r, p, (const T *)v, y, part, st,
               (double *)nullptr, (const T *)nullptr);
  LAUNCH_TIMED(ev, 1, k_bicg_a2<T>, grid, blk, lds, s, g, (const T *)y, rhat, v, part, (const SolverState *)st);
  LAUNCH_TIMED(ev, 2, k_bicg_b1<T>, grid, blk, lds, s, g, (const T *)r, (const T *)v, sv, z, part, st);
  LAUNCH_TIMED(ev, 3, k_bicg_b2<T>, grid, blk, lds, s, g, (const T *)z, (const T *)sv, tv, part,
               (const SolverState *)st);
  LAUNCH_TIMED(ev, 4, k_bicg_c<T>, grid, blk, lds, s, g, x, r, rhat, (const T *)y, (const T *)z, (const T *)sv,
               (const T *)tv, part, st);
  return (int)hipGetLastError();
}

int launch_bicg_iter(const Geo &g, hipStream_t s, double *x, double *r, const double *rhat, double *p, double *v,
                     double *y, double *z, double *sv, double *tv, Partials part, SolverState *st, hipEvent_t *ev) {
  return iter_impl<double>(g, s, x, r, rhat, p, v, y, z, sv, tv, part, st, ev);
}

int launch_bicg_iter32(const Geo &g, hipStream_t s, float *x, float *r, const float *rhat, float *p, float *v,
                       float *y, float *z, float *sv, float *tv, Partials part, SolverState *st, hipEvent_t *ev) {
  return iter_impl<float>(g, s, x, r, rhat, p, v, y, z, sv, tv, part, st, ev);
}

int launch_mp_begin(const Geo &g, hipStream_t s, SolverState *st) {
  hipLaunchKernelGGL(k_bicg_init_state, dim3(1), dim3(NSX_BLOCK), 0, s, st, 0.0, 0, 0.0);
  return (int)hipGetLastError();
}

int launch_rel_init(const Geo &g, hipStream_t s, const double *rhs, const double *phi_in, double *x, float *r32,
                    float *rhat32, float *e32, Partials part, SolverState *st, double eps, double delta,
                    int max_iters) {
  const dim3 grid(g.grid), blk(NSX_BLOCK);
  hipLaunchKernelGGL(k_bicg_init_state, dim3(1), blk, 0, s, st, eps, max_iters, delta);
  // the FP64 norms of r0 = b - A x0 seed the first A1 (P.init); r32 = rhat32
  hipLaunchKernelGGL(k_resid<float>, grid, blk, 0, s, g, rhs, phi_in, x, r32, rhat32, e32, part, part.init,
                     (double *)nullptr, phi_in != x ? 1 : 0, st);
  return (int)hipGetLastError();
}

int launch_rel_iter(const Geo &g, hipStream_t s, const double *rhs, double *x, float *e, float *r, const float *rhat,
                    float *p, float *v, float *y, float *z, float *sv, float *tv, Partials part, SolverState *st,
                    hipEvent_t *ev) {
  const dim3 grid(g.grid), blk(NSX_BLOCK);
  const size_t lds = level_lds(g, sizeof(float));
  LAUNCH_TIMED(ev, 0, k_rel_u, grid, blk, (size_t)0, s, g, rhs, (const double *)x, (const float *)e, r, rhat, part,
               st);
  LAUNCH_TIMED(ev, 1, k_bicg_a1<float>, grid, blk, lds, s, g, (const float *)r, p, (const float *)v, y, part, st, x,
               (const float *)e);
  LAUNCH_TIMED(ev, 2, k_bicg_a2<float>, grid, blk, lds, s, g, (const float *)y, rhat, v, part,
               (const SolverState *)st);
  LAUNCH_TIMED(ev, 3, k_bicg_b1<float>, grid, blk, lds, s, g, (const float *)r, (const float *)v, sv, z, part, st);
  LAUNCH_TIMED(ev, 4, k_bicg_b2<float>, grid, blk, lds, s, g, (const float *)z, (const float *)sv, tv, part,
               (const SolverState *)st);
  LAUNCH_TIMED(ev, 5, k_bicg_c<float>, grid, blk, lds, s, g, e, r, rhat, (const float *)y, (const float *)z,
               (const float *)sv, (const float *)tv, part, st);
  return (int)hipGetLastError();
}

int launch_mp_resid(const Geo &g, hipStream_t s, const double *rhs, const double *phi_in, double *x, float *r32,
                    float *rhat32, float *e32, Partials part, SolverState *st, double eps, double eps_inner,
                    int max_iters, int max_passes, int first_pass) {
  const dim3 grid(g.grid), blk(NSX_BLOCK);
  // the first pass copies phi_in into x and forms the residual of phi_in;
  // later passes form the residual of the corrected x
  hipLaunchKernelGGL(k_resid<float>, grid, blk, 0, s, g, rhs, first_pass ? phi_in : (const double *)x, x, r32,
                     rhat32, e32, part, part.outer, part.init, first_pass && phi_in != x ? 1 : 0, st);
  hipLaunchKernelGGL(k_mp_check, dim3(1), blk, 0, s, part.outer, grid.x, st, eps, eps_inner, max_iters, max_passes);
  return (int)hipGetLastError();
}

int launch_mp_update(const Geo &g, hipStream_t s, double *x, const float *e32) {
  hipLaunchKernelGGL(k_mp_update, dim3(g.grid), dim3(NSX_BLOCK), 0, s, g, x, e32);
  return (int)hipGetLastError();
}

// the regenerated diagonal AmC0 for the set-up check against PDE_center's
__global__ __launch_bounds__(NSX_BLOCK) void k_ppe_diag(Geo g, double *A0) {
  FOR_EACH_CELL(g, c) {
    int i, k;
    cell_of(g, c, i, k);
    A0[(size_t)i * g.NS + k] = ppe_row(g, i, k).A0;
  }
}

int launch_ppe_diag(const Geo &g, hipStream_t s, double *A0) {
  hipLaunchKernelGGL(k_ppe_diag, dim3(nsx_grid_for(g.ncell)), dim3(NSX_BLOCK), 0, s, g, A0);
  return (int)hipGetLastError();
}

int launch_gnorm_sum(const Geo &g, hipStream_t s, const double *x, Partials part) {
  hipLaunchKernelGGL(k_gnorm_sum, dim3(g.grid), dim3(NSX_BLOCK), 0, s, g, x, part.src);
  return (int)hipGetLastError();
}

int launch_gnorm_apply(const Geo &g, hipStream_t s, double *x, const SolverState *st) {
  hipLaunchKernelGGL(k_gnorm_apply, dim3(g.grid), dim3(NSX_BLOCK), 0, s, g, x, st);
  return (int)hipGetLastError();
}

// Resident blocks per CU that every kernel taking part in the solver's
// partial-sum chain reaches (they all run on one grid, g.grid, so that a
// consumer sums exactly its producer's partials); the grid is sized so that
// all of its blocks are resident at once, without a second, partial round.
int solver_blocks_per_cu(const Geo &g) {
  struct K {
    const void *f;
    size_t lds;
  } ks[] = {
      {(const void *)k_bicg_a1<double>, level_lds(g, 8)}, {(const void *)k_bicg_a2<double>, level_lds(g, 8)},
      {(const void *)k_bicg_b1<double>, level_lds(g, 8)}, {(const void *)k_bicg_b2<double>, level_lds(g, 8)},
      {(const void *)k_bicg_c<double>, level_lds(g, 8)},  {(const void *)k_bicg_a1<float>, level_lds(g, 4)},
      {(const void *)k_bicg_a2<float>, level_lds(g, 4)},  {(const void *)k_bicg_b1<float>, level_lds(g, 4)},
      {(const void *)k_bicg_b2<float>, level_lds(g, 4)},  {(const void *)k_bicg_c<float>, level_lds(g, 4)},
      {(const void *)k_resid<double>, 0},                 {(const void *)k_resid<float>, 0},
      {(const void *)k_rel_u, 0},
  };
  int best = 1 << 30;
  for (const K &k : ks) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k.f, NSX_BLOCK, k.lds) != hipSuccess || n < 1) return -1;
    if (n < best) best = n;
  }
  return best;
}
