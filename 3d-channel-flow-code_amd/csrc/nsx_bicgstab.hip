// nsx_bicgstab.hip -- Jacobi-preconditioned BiCGSTAB on the 7-point prism
// stencil (bicgstab.F90:303-591), in FP64 (the reference's arithmetic) and
// as the FP32 inner solver of an FP64 iterative refinement.
//
// One iteration is three kernels, one per global synchronisation point of
// the algorithm (see "iteration" below for the scaled vectors they carry):
//
//   A  stop test of k-1, beta; y = rt + beta q at the seven points;
//      v = A y; vt = v/A0; partial <rhat, v>
//                                     update_pk, diag_precdt, get_Apk  :726-1001
//   B  alpha; z = rt - alpha vt; t = A z; s = A0 z; partial <t,s>, <t,t>
//                                     update_sk, diag_precdt, get_Apk  :1007-1041
//   C  omega; y = rt + beta q again; r = s - omega t; x += alpha y +
//      omega z; rt = r/A0; q = y - omega vt; partial |r|^2, |r|_1,
//      <rhat, r>, sum x Vol
//                                     update_xkrk                      :1047-1111
//
// The matrix is PDE_center's, regenerated per row from the per-edge
// coefficient ppec, the cell area and the sigma spacings (bit-identical to
// AmC0..AmCB, checked at set-up), so no coefficient array is read.
//
// Work decomposition: a thread owns four consecutive layers k0..k0+3 of one
// column (k0 a multiple of 4, the column stride a multiple of 4), so every
// field access is a 16-byte-aligned vector load (2x double2 or 1x float4),
// the column data (area, ppec, neighbours) is fetched once per four cells,
// and the five divisions area/dsig(k0-1..k0+3) give AB and AT of all four
// rows.  The sigma spacings of the block live in LDS.  Slots outside the
// interior layers (ghost layers, padding) are written as zero in the work
// vectors and left unchanged in x.
//
// The scalars of a phase come from the partial sums of the phase before it:
// B sums A's partials and forms alpha itself, C forms omega from B's, and A
// of iteration k+1 forms the norms, the stopping test and rho_{k+1} from C's
// (see reduce_partials).  Block 0 records the scalars in SolverState for the
// host and for later phases.  Each kernel exits at once when st->done is set,
// so the host can queue many iterations ahead and poll rarely.  Inner
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
  if (g.trans) {
    // row i of A^T: the diagonal and the (symmetric) vertical couplings are
    // row i's, the lateral ones are those of the neighbours' rows towards i
    r.a1 = g.ppec_t[i * 3] * sv;
    r.a2 = g.ppec_t[i * 3 + 1] * sv;
    r.a3 = g.ppec_t[i * 3 + 2] * sv;
  }
  return r;
}

// ------------------------------------------------- four-layer helpers -----
template <typename T>
struct Prec;
template <>
struct Prec<double> {
  __device__ static const double *area(const Geo &g) { return g.area; }
  __device__ static const double *ppec(const Geo &g) { return g.ppec; }
  __device__ static const double *ppec_t(const Geo &g) { return g.ppec_t; }
  __device__ static const double *dsig(const Geo &g) { return g.dsig; }
  __device__ static const double *dsigv(const Geo &g) { return g.dsigv; }
};
template <>
struct Prec<float> {
  __device__ static const float *area(const Geo &g) { return g.area_f; }
  __device__ static const float *ppec(const Geo &g) { return g.ppec_f; }
  __device__ static const float *ppec_t(const Geo &g) { return g.ppec_t_f; }
  __device__ static const float *dsig(const Geo &g) { return g.dsig_f; }
  __device__ static const float *dsigv(const Geo &g) { return g.dsigv_f; }
};


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

// Column data of the PPE rows of column i: the area, the edge coefficients
// ppec that enter the diagonal, and the lateral couplings (ppec again, or
// ppec_t for the rows of A^T, see ppe_row).
template <typename T>
struct Col {
  T area, c[3], d[3];
};

template <typename T>
__device__ inline void load_col(const Geo &g, int i, Col<T> &C) {
  C.area = Prec<T>::area(g)[i];
  const T *pc = Prec<T>::ppec(g) + i * 3;
#pragma unroll
  for (int j = 0; j < 3; ++j) C.c[j] = pc[j];
  if (g.trans) {
    const T *pt = Prec<T>::ppec_t(g) + i * 3;
#pragma unroll
    for (int j = 0; j < 3; ++j) C.d[j] = pt[j];
  } else {
#pragma unroll
    for (int j = 0; j < 3; ++j) C.d[j] = C.c[j];
  }
}

// q[u] = area / dsig(k0-1+u), u = 0..4: AB of layer k0+u is q[u] and its AT
// is q[u+1], so five divisions serve the four rows of a quad
template <typename T>
__device__ inline void quot5(const Col<T> &C, const T *lv, int k0, T (&q)[5]) {
#pragma unroll
  for (int u = 0; u < 5; ++u) {
    const int k = k0 - 1 + u;
    q[u] = C.area / lv[k < 0 ? 0 : k];
  }
}

// quot5 of the single-reduction CG: the LDS level table also holds 1/dsig
// at lv[2NS..3NS), and q[u] = area * (1/dsig) is a product, not a division
// (its operator rounds differently from area/dsig, in the last bit; every
// launch of that CG forms its rows this way)
template <typename T>
__device__ inline void quot5_r(const Col<T> &C, const T *lv, int NS, int k0, T (&q)[5]) {
#pragma unroll
  for (int u = 0; u < 5; ++u) {
    const int k = k0 - 1 + u;
    q[u] = C.area * lv[2 * NS + (k < 0 ? 0 : k)];
  }
}

// the row of layer k0+u of the column (ppe_row's arithmetic); `in` is false
// outside the interior layers, where the solver's vectors hold zero
template <typename T>
struct Row1 {
  T A0, a1, a2, a3, AT, AB;
  bool in;
};

template <typename T>
__device__ inline Row1<T> row_of(const Geo &g, const T *lv, const Col<T> &C, const T (&q)[5], int k0, int u) {
  const int k = k0 + u;
  Row1<T> R;
  R.in = k >= 1 && k <= g.NZ - 2;
  const T sv = lv[g.NS + (k >= 1 ? k - 1 : 0)];
  R.AT = (k == g.NZ - 2 && g.bctop == 1) ? T(0) : q[u + 1];
  R.AB = (k == 1 && g.bcbot == 1) ? T(0) : q[u];
  R.a1 = C.c[0] * sv;
  R.a2 = C.c[1] * sv;
  R.a3 = C.c[2] * sv;
  R.A0 = T(-1) * (R.AT + R.AB + R.a1 + R.a2 + R.a3);
  if (g.trans) {
    R.a1 = C.d[0] * sv;
    R.a2 = C.d[1] * sv;
    R.a3 = C.d[2] * sv;
  }
  return R;
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
// into `outer`.  FP64 solve (TO = double): rt = r/A0, rhat = r, x = x0.
// Refinement pass of the mixed solve (TO = float): rt32 = (float)(r/A0),
// rhat32 = (float) r, the inner solution e = 0, and the inner norms |r32|^2,
// |r32|_1 into `inner` laid out as the initial partials of an FP32 solve with
// zero start.
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
    r[s] = (TO)(res / A.A0);  // the solver keeps r / A0
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
    st->m_stop = 0;
    st->m_iters = 0;
    st->delta = delta;
    st->ureq[0] = st->ureq[1] = 0;
  }
}

// ----------------------------------------------------------- iteration ----
// The iteration carries the preconditioned (scaled) vectors: rt = r/A0,
// q = (p - omega v)/A0, vt = v/A0, so that the two vectors the stencils
// need, y = p/A0 = rt + beta q and z = s/A0 = rt - alpha vt, are formed at
// the seven points from stored vectors with one multiply-add, without the
// pointwise passes that would otherwise write them.  The unscaled s = A0 z
// and r = s - omega t are formed where needed and never stored.  In exact
// arithmetic these are the reference's vectors (update_pk, diag_precdt,
// update_sk, update_xkrk, bicgstab.F90:726-1111); the rounding differs at
// the last bit per operation.
//
// C forms y = rt + beta q again from the rt and q it reads anyway (the same
// expression on the same values, so the same bits as A's), instead of
// reading a y that A would have to store.
//
// Per interior prism (b = vector width): A reads rt, q, rhat, writes vt
// (4b); B reads rt, vt, writes t (3b); C reads x, rt, q, vt, t, rhat, writes
// x, rt, q (9b): 16b, against 23b for five passes that store y, z, s.

// Closes iteration k-1 (norms and the stopping test of bicgstab.F90:544-575,
// the breakdown guards of :479-490; at k = 0 the norms of :453-469) from C's
// partial sums (P.init at k = 0).  Returns
// false when the solve stops; sets beta (0 at k = 0).
__device__ inline bool begin_iteration(const Geo &g, Partials P, SolverState *st, double &beta_out,
                                       bool &first_out) {
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
  first_out = first;
  if (stop) return false;
  beta_out = first ? 0.0 : (rho / st->a_rho[(k_it - 1) & 1]) * (st->b_alpha / st->c_omega);
  return true;
}

// a + f b over the four layers at `off` (b not read when first)
template <typename T>
__device__ inline void axpy4(const T *a, const T *b, T f, bool first, size_t off, T (&o)[4]) {
  ld4(a + off, o);
  if (!first) {
    T bb[4];
    ld4(b + off, bb);
#pragma unroll
    for (int u = 0; u < 4; ++u) o[u] = o[u] + f * bb[u];
  }
}

template <typename T>
__device__ inline T axpy1(const T *a, const T *b, T f, bool first, size_t off) {
  return first ? a[off] : a[off] + f * b[off];
}

// (A f) at layer k0+u of column i from the values of f at the seven points
// (get_Apk, bicgstab.F90:976-998; Neumann vertical ghosts, periodic partners
// through nbr); rows outside the interior give 0
template <typename T>
__device__ inline T apply_row(const Geo &g, const Row1<T> &R, int k0, int u, const T (&fc)[4], T fm, T fp, T f0,
                              T f1, T f2) {
  const int k = k0 + u;
  const T fs = fc[u];
  const T fu = (k + 1 <= g.NZ - 2) ? (u < 3 ? fc[u < 3 ? u + 1 : 3] : fp) : fs;
  const T fd = (k - 1 >= 1) ? (u > 0 ? fc[u > 0 ? u - 1 : 0] : fm) : fs;
  return R.in ? R.A0 * fs + R.a1 * f0 + R.a2 * f1 + R.a3 * f2 + R.AT * fu + R.AB * fd : T(0);
}

// apply_row of the FP32 solvers as a chain of fused multiply-adds (the FP64
// form keeps the reference's separate products and sums)
template <>
__device__ inline float apply_row<float>(const Geo &g, const Row1<float> &R, int k0, int u, const float (&fc)[4],
                                         float fm, float fp, float f0, float f1, float f2) {
  const int k = k0 + u;
  const float fs = fc[u];
  const float fu = (k + 1 <= g.NZ - 2) ? (u < 3 ? fc[u < 3 ? u + 1 : 3] : fp) : fs;
  const float fd = (k - 1 >= 1) ? (u > 0 ? fc[u > 0 ? u - 1 : 0] : fm) : fs;
  float r = R.A0 * fs;
  r = __builtin_fmaf(R.a1, f0, r);
  r = __builtin_fmaf(R.a2, f1, r);
  r = __builtin_fmaf(R.a3, f2, r);
  r = __builtin_fmaf(R.AT, fu, r);
  r = __builtin_fmaf(R.AB, fd, r);
  return R.in ? r : 0.f;
}

// Windows of NSX_BLOCK consecutive quads, swept band by band.  Blocks are
// dealt round-robin to the XCDs, so hardware block b runs on XCD b % 8.  XCD
// x owns the x-th contiguous eighth of the windows, and its G/8 blocks sweep
// that slice together: block j takes windows j, j + G/8, j + 2G/8, ... of it.
// At any time one XCD works on one band of G/8 adjacent windows, so the
// y-neighbour column of a quad (one mesh row, a few hundred columns away)
// usually lies in the band its own L2 is serving.
//
// With rev the blocks sweep their slices from the far end: block j takes the
// slice's windows from its last one down.  The single-reduction CG alternates
// the direction launch by launch, so a launch first reads the windows the
// launch before it wrote last, while they are still in the chip's last-level
// cache.
struct Sweep {
  uint32_t w, end, step, lo, hi;
  bool rev;
  // the window that sweep position wi (w <= wi < end) stands for
  __device__ uint32_t at(uint32_t wi) const { return rev ? lo + hi - 1 - wi : wi; }
};

__device__ inline Sweep window_sweep(uint32_t nwin, bool rev = false) {
  const uint32_t G = gridDim.x, b = blockIdx.x;
  Sweep S;
  S.rev = rev;
  if ((G % NSX_NXCD) == 0) {
    const uint32_t per = (nwin + NSX_NXCD - 1) / NSX_NXCD, x = b % NSX_NXCD;
    const uint32_t s0 = x * per < nwin ? x * per : nwin;
    S.lo = s0;
    S.hi = S.end = s0 + per < nwin ? s0 + per : nwin;
    S.w = s0 + b / NSX_NXCD;
    S.step = G / NSX_NXCD;
  } else {
    S.lo = 0;
    S.w = b;
    S.hi = S.end = nwin;
    S.step = G;
  }
  return S;
}

// for_each_quad in the order of window_sweep: windows of NSX_BLOCK quads,
// band by band per XCD, from the far end with rev
template <class F>
__device__ inline void for_each_quad_swept(const Geo &g, bool rev, F &&f) {
  const uint32_t nwin = (g.nquad + NSX_BLOCK - 1) / NSX_BLOCK;
  const Sweep S = window_sweep(nwin, rev);
  for (uint32_t wi = S.w; wi < S.end; wi += S.step) {
    const uint32_t t = S.at(wi) * NSX_BLOCK + threadIdx.x;
    if (t < g.nquad) {
      const uint32_t i = fdiv(t, g.divNQ);
      const int k0 = 4 * (int)(t - i * (uint32_t)g.NQ);
      f((int)i, k0, (size_t)i * g.NS + k0);
    }
  }
}

// Seven-point values of a vector f formed from stored vectors around every
// quad.  Each thread forms f on its own four layers (own4) and publishes them
// in LDS; the vertical neighbours and the neighbour columns whose quad lies
// in the block's current window of NSX_BLOCK quads (the in-quad partner and
// usually the x-neighbour) are read from there, the rest (the y-neighbour,
// window edges) are formed from global memory (val4 for four layers, val1
// for one) with the same arithmetic, so every value is bit-identical
// whichever way it was obtained.  Every global load of a window (own values,
// neighbour indices, column data, the out-of-window neighbours and what `pre`
// loads) is issued before the window's barrier, so their latencies overlap;
// body(i, k0, s0, C, fc, fm, fp, f0, f1, f2) runs after it.  The windows hold
// W <= NSX_BLOCK quads; post(act) runs on every thread of the block after
// the bodies of a window (it may synchronise the block).
template <typename T, class Own4, class Val4, class Val1, class Pre, class Body, class Post>
__device__ inline void for_each_quad_formed_w(const Geo &g, uint32_t W, Own4 &&own4, Val4 &&val4, Val1 &&val1,
                                              Pre &&pre, Body &&body, Post &&post, bool rev = false) {
  __shared__ __align__(16) T win[NSX_BLOCK][4];
  const uint32_t nwin = (g.nquad + W - 1) / W;
  const Sweep S = window_sweep(nwin, rev);
  const uint32_t w0 = S.w, wend = S.end, step = S.step;
  // the neighbour columns of this thread's quad at sweep position wi, loaded
  // one window ahead: the gathers of a window would otherwise wait on them
  int nc[3] = {0, 0, 0};
  auto nbr_of = [&](uint32_t wi) {
    const uint32_t t = S.at(wi) * W + threadIdx.x;
    if (wi < wend && threadIdx.x < W && t < g.nquad) {
      const int e = 3 * (int)fdiv(t, g.divNQ);
      nc[0] = g.nbr[e];
      nc[1] = g.nbr[e + 1];
      nc[2] = g.nbr[e + 2];
    }
  };
  nbr_of(w0);
  for (uint32_t wi = w0; wi < wend; wi += step) {
    const uint32_t base = S.at(wi) * W;
    const uint32_t end = base + W < g.nquad ? base + W : g.nquad;
    const uint32_t t = base + threadIdx.x;
    const bool act = t < end;
    const uint32_t lim = end - base;
    const bool mwin = threadIdx.x >= 1, pwin = threadIdx.x + 1 < lim;
    int i = 0, k0 = 0;
    size_t s0 = 0;
    T fc[4] = {T(0), T(0), T(0), T(0)}, fn[3][4], fm = T(0), fp = T(0);
    uint32_t d[3] = {0u, 0u, 0u};
    Col<T> C;
    if (act) {
      i = (int)fdiv(t, g.divNQ);
      k0 = 4 * (int)(t - (uint32_t)i * (uint32_t)g.NQ);
      s0 = (size_t)i * g.NS + k0;
      own4(s0, fc);
      const int n[3] = {nc[0], nc[1], nc[2]};
      load_col(g, i, C);
      pre(s0);
      const uint32_t qq = (uint32_t)(k0 >> 2);
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        d[j] = (uint32_t)n[j] * (uint32_t)g.NQ + qq - base;
        if (d[j] >= lim) val4((size_t)n[j] * g.NS + k0, fn[j]);
      }
      if (k0 >= 1 && !mwin) fm = val1(s0 - 1);
      if (k0 + 4 <= g.NZ - 2 && !pwin) fp = val1(s0 + 4);
    }
    nbr_of(wi + step);
#pragma unroll
    for (int u = 0; u < 4; ++u) win[threadIdx.x][u] = fc[u];
    __syncthreads();
    if (act) {
      if (k0 >= 1 && mwin) fm = win[threadIdx.x - 1][3];
      if (k0 + 4 <= g.NZ - 2 && pwin) fp = win[threadIdx.x + 1][0];
#pragma unroll
      for (int j = 0; j < 3; ++j)
        if (d[j] < lim) {
#pragma unroll
          for (int u = 0; u < 4; ++u) fn[j][u] = win[d[j]][u];
        }
      body(i, k0, s0, C, fc, fm, fp, fn[0], fn[1], fn[2]);
    }
    post(act);
    __syncthreads();
  }
}

template <typename T, class Own4, class Val4, class Val1, class Pre, class Body>
__device__ inline void for_each_quad_formed(const Geo &g, Own4 &&own4, Val4 &&val4, Val1 &&val1, Pre &&pre,
                                            Body &&body, bool rev = false) {
  for_each_quad_formed_w<T>(g, NSX_BLOCK, own4, val4, val1, pre, body, [](bool) {}, rev);
}

// the seven-point values of f = a + c b (f = a when first)
template <typename T, class Pre, class Body>
__device__ inline void for_each_quad_gathered(const Geo &g, const T *a, const T *b, T c, bool first, Pre &&pre,
                                              Body &&body, bool rev = false) {
  auto v4 = [&](size_t off, T(&o)[4]) { axpy4(a, b, c, first, off, o); };
  auto v1 = [&](size_t off) { return axpy1(a, b, c, first, off); };
  for_each_quad_formed<T>(g, v4, v4, v1, pre, body, rev);
}

// ------------------------------------------------------------ phase A ----
// Stop test of k-1 and beta; y = p/A0 = rt + beta q (y = rt at k = 0),
// v = A y (get_Apk), vt = v/A0, partial sums of <rhat, v> (:509).
template <typename T>
__global__ __launch_bounds__(NSX_BLOCK) void k_bicg_a(Geo g, const T *rt, const T *q, const T *rhat, T *vt, Partials P,
                                                       SolverState *st) {
  if (st->done) return;
  // the three phases alternate their sweep direction launch by launch
  // (window_sweep): A and C of iteration k sweep from the far end when k is
  // odd, B the other way
  const bool rev = (st->c_iter & 1) != 0;
  double beta_d = 0.0;
  bool first = false;
  if (!begin_iteration(g, P, st, beta_d, first)) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) st->a_beta = beta_d;
  const T beta = (T)beta_d;
  const T *lv = load_levels<T>(g);
  double acc[1] = {0.0};
  T rh[4];
  for_each_quad_gathered(
      g, rt, q, beta, first, [&](size_t s0) { ld4(rhat + s0, rh); },
      [&](int i, int k0, size_t s0, const Col<T> &C, const T(&yc)[4], T ym, T yp, const T(&y0)[4],
          const T(&y1)[4], const T(&y2)[4]) {
        T qv[5], vo[4];
        quot5(C, lv, k0, qv);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const Row1<T> R = row_of(g, lv, C, qv, k0, u);
          const T v = apply_row(g, R, k0, u, yc, ym, yp, y0[u], y1[u], y2[u]);
          vo[u] = R.in ? v / R.A0 : T(0);
          acc[0] += (double)rh[u] * (double)v;
        }
        st4(vt + s0, vo);
      },
      rev);
  write_partials<1>(acc, P.a);
}

// ------------------------------------------------------------ phase B ----
// alpha = rho/<rhat, v> (breakdown guard of :512-517); z = s/A0 = rt -
// alpha vt, t = A z (get_Apk), s = A0 z, partial sums of <t,s>, <t,t>.
template <typename T>
__global__ __launch_bounds__(NSX_BLOCK) void k_bicg_b(Geo g, const T *rt, const T *vt, T *tv, Partials P,
                                                       SolverState *st) {
  if (st->done) return;
  double tr[1];
  reduce_partials<1>(P.a, gridDim.x, tr);
  const int k_it = st->a_iter;
  const double al = st->a_rho[k_it & 1] / tr[0];
  if (fabs(tr[0]) < TINY || !isfinite(al)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->exit_code = EXIT_BREAKDOWN;
      st->done = 1;
    }
    return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) st->b_alpha = al;
  const bool rev = (k_it & 1) == 0;
  const T alm = (T)(-al);
  const T *lv = load_levels<T>(g);
  double acc[2] = {0.0, 0.0};
  for_each_quad_gathered(
      g, rt, vt, alm, false, [&](size_t) {},
      [&](int i, int k0, size_t s0, const Col<T> &C, const T(&zc)[4], T zm, T zp, const T(&z0)[4],
          const T(&z1)[4], const T(&z2)[4]) {
        T qv[5], t[4];
        quot5(C, lv, k0, qv);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const Row1<T> R = row_of(g, lv, C, qv, k0, u);
          t[u] = apply_row(g, R, k0, u, zc, zm, zp, z0[u], z1[u], z2[u]);
          const double sd = R.in ? (double)(R.A0 * zc[u]) : 0.0, td = (double)t[u];
          acc[0] += td * sd;
          acc[1] += td * td;
        }
        st4(tv + s0, t);
      },
      rev);
  write_partials<2>(acc, P.b);
}

// ------------------------------------------------------------- phase C ----
// omega = <t,s>/<t,t>; y = rt + beta q (A's y, formed again), z = rt -
// alpha vt, s = A0 z, r = s - omega t, x += alpha y + omega z (update_xkrk
// :1047-1111), rt = r/A0, q = y - omega vt; partial |r|^2, |r|_1, <rhat, r>,
// sum x Vol.
template <typename T>
__global__ __launch_bounds__(NSX_BLOCK) void k_bicg_c(Geo g, T *x, T *rt, T *q, const T *rhat, const T *vt,
                                                       const T *tv, Partials P, SolverState *st) {
  if (st->done) return;
  double tr[2];
  reduce_partials<2>(P.b, gridDim.x, tr);
  const double al = st->b_alpha, om = tr[0] / tr[1];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->c_omega = om;
    st->c_iter = st->a_iter + 1;
  }
  const T alt = (T)al, omt = (T)om, bet = (T)st->a_beta;
  const bool first = st->a_iter == 0;
  const bool rev = (st->a_iter & 1) != 0;
  const T *lv = load_levels<T>(g);
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for_each_quad_swept(g, rev, [&](int i, int k0, size_t s0) {
    T xx[4], rr[4], vv[4], yy[4], tt[4], rh[4], qq[4], qv[5];
    // Scheduling fences keep the two 16-byte halves of each vector's run
    // back to back.  Across a wave one half touches every cache line of the
    // wave's span, so the second half then hits the lines the first one
    // brought in; interleaving the six vectors' halves re-requests each line
    // later and costs 15 % of the kernel in FP64.
    ld4(x + s0, xx);
    __builtin_amdgcn_sched_barrier(0);
    ld4(rt + s0, rr);
    __builtin_amdgcn_sched_barrier(0);
    ld4(vt + s0, vv);
    __builtin_amdgcn_sched_barrier(0);
    ld4(q + s0, yy);
    __builtin_amdgcn_sched_barrier(0);
    ld4(tv + s0, tt);
    __builtin_amdgcn_sched_barrier(0);
    ld4(rhat + s0, rh);
    __builtin_amdgcn_sched_barrier(0);
    Col<T> C;
    load_col(g, i, C);
    quot5(C, lv, k0, qv);
#pragma unroll
    for (int u = 0; u < 4; ++u) yy[u] = first ? rr[u] : rr[u] + bet * yy[u];  // y = rt + beta q, as in A
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int k = k0 + u;
      const Row1<T> R = row_of(g, lv, C, qv, k0, u);
      const bool in = R.in;
      const T z = rr[u] - alt * vv[u];
      const T sn = R.A0 * z;
      const T rn = sn - omt * tt[u];
      const T xn = xx[u] + alt * yy[u] + omt * z;
      xx[u] = in ? xn : xx[u];
      rr[u] = in ? rn / R.A0 : T(0);
      qq[u] = in ? yy[u] - omt * vv[u] : T(0);
      if (in) {
        const double rd = (double)rn;
        acc[0] += rd * rd;
        acc[1] += fabs(rd);
        acc[2] += (double)rh[u] * rd;
        acc[3] += (double)xn * (double)(C.area * lv[k]);
      }
    }
    st4(x + s0, xx);
    st4(rt + s0, rr);
    st4(q + s0, qq);
  });
  write_partials<4>(acc, P.c);
}

// ------------------------------------------------ mixed-precision CG ----
// The inner solver of NSX_REFINE_CG: Jacobi-preconditioned conjugate
// gradients in FP32 on the correction e, with reliable updates (Clark et al.
// 2010) against the FP64 solution x.  PDE_center's operator is symmetric
// (up to the last bit) and negative semi-definite with the constants as its
// null space, and the mean-removed right-hand side lies in its range, so CG
// with M = diag(A) applies (PCG on -A with -M).  Unlike FP32 BiCGSTAB, whose
// residual peaks on this anisotropic operator push the recursive residual
// away from the true one, FP32 CG with reliable updates converges in the
// iteration count of FP64 CG.
//
// Start (once per solve, k_cg_start): the norms of b and of r0 = b - A x0
// from k_resid's partials, the reference's no-solve test, and r0 again with
// rho_0 = <r0, M^-1 r0> into P.u.  One iteration is then two kernels:
//
//   W  after a true residual: x += e and the reference's stopping rule on
//      it.  Otherwise, from X's partials: when |r| < delta * max|r| since
//      the last update or the recursive residual meets the stopping rule,
//      ask this iteration's X for a true residual.  beta = rho_k /
//      rho_{k-1}; p = rt + beta p at the seven points; w = A p; <p, w>
//   X  alpha = rho_k / <p, w>; p formed again (same bits as W's); e += alpha
//      p (from zero after a true residual); then either the recursive
//      residual, r = A0 rt - alpha w, rt = r/A0, partial rho_{k+1} = <r, rt>,
//      |r|^2, |r|_1 into P.c; or, when W asked for it, the true residual
//      r = b - A (x + e) in FP64 with the new e formed at the seven points
//      from the stored vectors (the same expressions as the threads that
//      store it), rt = r/A0, the same partials and sum (x+e) Vol into P.u.
//      The true-residual pass reads e, rt, p at neighbouring prisms, so it
//      writes the new e, rt, p to a second set of buffers (e_b, rt_b, p_b),
//      which W and X of the next iteration read; that X writes the first
//      set again.  W never asks right after a true residual, so two
//      consecutive iterations never both take this pass.
//
// Per interior prism W reads rt, p and writes w (3 x 4 B), X reads e, rt, p,
// w and writes e, rt, p (7 x 4 B): 40 B per iteration, against 128 B of an
// FP64 BiCGSTAB iteration, which does two products with A.

// the true residual of x + e at interior prism (i,k), with e(q) given by a
// function of the slot; `res` and the prism's x + e returned
template <class E>
__device__ inline void true_residual(const Geo &g, const double *rhs, const double *x, double avec, int i, int k,
                                     E &&eq, double &res, double &xs) {
  const size_t s = (size_t)i * g.NS + k;
  const double b = rhs[s] - avec * g.area[i] * g.dsig[k];
  const Row A = ppe_row(g, i, k);
  const int ee = i * 3;
  const size_t n0 = (size_t)g.nbr[ee] * g.NS + k, n1 = (size_t)g.nbr[ee + 1] * g.NS + k,
               n2 = (size_t)g.nbr[ee + 2] * g.NS + k;
  xs = x[s] + (double)eq(s, true);
  const double Ax = A.A0 * xs + A.a1 * (x[n0] + (double)eq(n0, true)) + A.a2 * (x[n1] + (double)eq(n1, true)) +
                    A.a3 * (x[n2] + (double)eq(n2, true)) + A.AT * (x[s + 1] + (double)eq(s + 1, k + 1 <= g.NZ - 2)) +
                    A.AB * (x[s - 1] + (double)eq(s - 1, k - 1 >= 1));
  res = b - Ax;
}

template <typename T>
__global__ __launch_bounds__(NSX_BLOCK) void k_cg_start(Geo g, const double *rhs, const double *x, T *rt,
                                                         Partials P, SolverState *st) {
  // the norms of b and r0 = b - A x0 (k_resid's partials) and the
  // reference's no-solve test (bicgstab.F90:453-469)
  double t[5];
  reduce_partials<5>(P.init, gridDim.x, t);
  double b2 = sqrt(t[0]), b1 = t[2];
  const double r2 = sqrt(t[1]), r1 = t[3];
  if (b2 == 0) b2 = 1.0;
  if (b1 == 0) b1 = 1.0;
  const double e1 = r1 / b1, e2 = r2 / b2;
  const bool stop = e1 < st->eps || e2 < st->eps;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->b1 = b1;
    st->b2 = b2;
    st->e1 = e1;
    st->e2 = e2;
    st->xvol = t[4];
    st->cgs_runs = 0;
    if (stop) {
      st->exit_code = EXIT_NO_SOLVE;
      st->done = 1;
    }
  }
  if (stop) return;
  const double avec = st->avec;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  FOR_EACH_CELL(g, c) {
    int i, k;
    cell_of(g, c, i, k);
    double res, xs;
    true_residual(g, rhs, x, avec, i, k, [](size_t, bool) { return T(0); }, res, xs);
    const Row A = ppe_row(g, i, k);
    const T rtf = (T)(res / A.A0);
    rt[(size_t)i * g.NS + k] = rtf;
    acc[0] += res * (double)rtf;
    acc[1] += res * res;
    acc[2] += fabs(res);
    acc[3] += xs * (g.area[i] * g.dsig[k]);
  }
  write_partials<4>(acc, P.u);
}

// whether iteration k starts from a true residual (the start, or X of k-1)
__device__ inline bool cg_fresh(const SolverState *st, int k) { return k == 0 || st->ureq[(k - 1) & 1]; }

__global__ __launch_bounds__(NSX_BLOCK) void k_cg_w(Geo g, const float *rt, const float *p, const float *rt_b,
                                                     const float *p_b, const float *e_b, float *w, double *x,
                                                     Partials P, SolverState *st) {
  if (st->done) return;
  const int k_it = st->c_iter;
  const bool upd = cg_fresh(st, k_it);
  const bool alt = upd && k_it > 0;  // X of k-1 took the true-residual pass into the second set
  double rho;
  int stop = 0, code = EXIT_CONVERGED, req = 0;
  if (upd) {
    double t[4];
    reduce_partials<4>(P.u, (uint32_t)g.grid, t);  // the start and X run on the solver grid
    rho = t[0];
    const double r2 = sqrt(t[1]), e1 = t[2] / st->b1, e2 = r2 / st->b2;
    // x + e is the solution the true residual belongs to (e = 0 at k = 0);
    // an overflowed correction is not folded, and the FP64 solve finishes
    // from x
    if (alt && isfinite(t[0]) && isfinite(t[1])) {
      FOR_EACH_CELL(g, c) {
        int i, k;
        cell_of(g, c, i, k);
        const size_t s = (size_t)i * g.NS + k;
        x[s] = x[s] + (double)e_b[s];
      }
    }
    if (!isfinite(t[0]) || !isfinite(t[1])) {
      stop = 1;
      code = EXIT_BREAKDOWN;
    } else if (e1 < st->eps || e2 < st->eps) {
      stop = 1;
      code = EXIT_CONVERGED;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->e1 = e1;
      st->e2 = e2;
      st->xvol = t[3];
      st->maxrr[k_it & 1] = r2;
      st->mp_pass += 1;
    }
  } else {
    double t[3];
    reduce_partials<3>(P.c, (uint32_t)g.grid, t);
    rho = t[0];
    const double rN = sqrt(t[1]), e1 = t[2] / st->b1, e2 = rN / st->b2;
    const double mrr = fmax(st->maxrr[(k_it - 1) & 1], rN);
    req = e1 < st->eps || e2 < st->eps || rN < st->delta * mrr;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->e1 = e1;
      st->e2 = e2;
      st->maxrr[k_it & 1] = mrr;
    }
  }
  if (!stop && k_it >= st->max_iters) {
    stop = 1;
    code = EXIT_MAX;
  }
  if (!stop && (rho == 0.0 || !isfinite(rho))) {
    stop = 1;
    code = EXIT_BREAKDOWN;
  }
  const bool first = k_it == 0;
  const double beta_d = first ? 0.0 : rho / st->a_rho[(k_it - 1) & 1];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->a_iter = k_it;
    st->a_rho[k_it & 1] = rho;
    st->a_beta = beta_d;
    st->ureq[k_it & 1] = req;
    if (stop) {
      st->exit_code = code;
      st->done = 1;
    }
  }
  if (stop) return;
  const float *lv = load_levels<float>(g);
  double acc[1] = {0.0};
  for_each_quad_gathered(
      g, alt ? rt_b : rt, alt ? p_b : p, (float)beta_d, first, [&](size_t) {},
      [&](int i, int k0, size_t s0, const Col<float> &C, const float(&pc)[4], float pm, float pp,
          const float(&p0)[4], const float(&p1)[4], const float(&p2)[4]) {
        float qv[5], wo[4];
        quot5(C, lv, k0, qv);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const Row1<float> R = row_of(g, lv, C, qv, k0, u);
          wo[u] = apply_row(g, R, k0, u, pc, pm, pp, p0[u], p1[u], p2[u]);
          acc[0] += (double)pc[u] * (double)wo[u];
        }
        st4(w + s0, wo);
      });
  write_partials<1>(acc, P.a);
}

__global__ __launch_bounds__(NSX_BLOCK) void k_cg_x(Geo g, float *e, float *rt, float *p, float *e_b, float *rt_b,
                                                     float *p_b, const float *w, const double *rhs, const double *x,
                                                     Partials P, SolverState *st) {
  if (st->done) return;
  double tr[1];
  reduce_partials<1>(P.a, (uint32_t)g.grid_cg, tr);  // W runs on the CG grid
  const int k_it = st->a_iter;
  const double al = st->a_rho[k_it & 1] / tr[0];
  if (tr[0] == 0.0 || !isfinite(al)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->exit_code = EXIT_BREAKDOWN;
      st->done = 1;
    }
    return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->b_alpha = al;
    st->c_iter = k_it + 1;
  }
  const float alf = (float)al, bet = (float)st->a_beta;
  const bool first = k_it == 0, fresh = cg_fresh(st, k_it);
  if (st->ureq[k_it & 1]) {
    // the true residual of x + e_{k+1}; e_{k+1} and p_k at any interior slot
    // are formed from the stored vectors exactly as the thread that stores
    // them forms them below (ghost slots keep their stored e)
    auto pn = [&](size_t q) { return first ? rt[q] : rt[q] + bet * p[q]; };
    auto en = [&](size_t q, bool in) { return in ? (fresh ? 0.f : e[q]) + alf * pn(q) : e[q]; };
    const double avec = st->avec;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    FOR_EACH_CELL(g, c) {
      int i, k;
      cell_of(g, c, i, k);
      const size_t s = (size_t)i * g.NS + k;
      double res, xs;
      true_residual(g, rhs, x, avec, i, k, en, res, xs);
      const Row A = ppe_row(g, i, k);
      const float rtf = (float)(res / A.A0);
      acc[0] += res * (double)rtf;
      acc[1] += res * res;
      acc[2] += fabs(res);
      acc[3] += xs * (g.area[i] * g.dsig[k]);
      e_b[s] = en(s, true);
      p_b[s] = pn(s);
      rt_b[s] = rtf;
    }
    write_partials<4>(acc, P.u);
    return;
  }
  const float *lv = load_levels<float>(g);
  const float *rt_in = fresh && !first ? rt_b : rt, *p_in = fresh && !first ? p_b : p;
  double acc[3] = {0.0, 0.0, 0.0};
  for_each_quad(g, [&](int i, int k0, size_t s0) {
    float ee[4], rr[4], pp[4], ww[4], qv[5];
    if (fresh) {
      ee[0] = ee[1] = ee[2] = ee[3] = 0.f;
    } else {
      ld4(e + s0, ee);
    }
    ld4(rt_in + s0, rr);
    ld4(p_in + s0, pp);
    ld4(w + s0, ww);
    Col<float> C;
    load_col(g, i, C);
    quot5(C, lv, k0, qv);
#pragma unroll
    for (int u = 0; u < 4; ++u) pp[u] = first ? rr[u] : rr[u] + bet * pp[u];  // p as W formed it
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const Row1<float> R = row_of(g, lv, C, qv, k0, u);
      const bool in = R.in;
      const float en = ee[u] + alf * pp[u];
      const float rn = R.A0 * rr[u] - alf * ww[u];
      const float rtn = rn / R.A0;
      ee[u] = in ? en : ee[u];
      pp[u] = in ? pp[u] : 0.f;
      rr[u] = in ? rtn : 0.f;
      if (in) {
        const double rd = (double)rn;
        acc[0] += rd * (double)rtn;
        acc[1] += rd * rd;
        acc[2] += fabs(rd);
      }
    }
    st4(e + s0, ee);
    st4(rt + s0, rr);
    st4(p + s0, pp);
  });
  write_partials<3>(acc, P.c);
}

// --------------------------------------- single-reduction FP32 CG ----
// NSX_REFINE_CGSR: the CG of NSX_REFINE_CG in the form of Chronopoulos and
// Gear (J. Comput. Appl. Math. 25, 1989), which needs one global reduction
// per iteration.  With u = M^-1 r and w = A u:
//   gamma_i = <r_i, u_i>, delta_i = <w_i, u_i>, beta_i = gamma_i/gamma_{i-1},
//   alpha_i = gamma_i / (delta_i - beta_i gamma_i / alpha_{i-1}),
//   p_i = u_i + beta_i p_{i-1}, s_i = w_i + beta_i s_{i-1},
//   e += alpha_i p_i, r_{i+1} = r_i - alpha_i s_i,
// (beta_0 = 0, alpha_0 = gamma_0/delta_0), so the next product w_{i+1} =
// A u_{i+1} is formed in the same pass as the updates: u_{i+1} = rt -
// alpha_i (om + beta_i sg) at the seven points, from the stored scaled
// vectors rt = u = r/A0, om = w/A0 and sg = s/A0.  One launch per iteration
// (mode N) reads rt, om, sg and writes rt, om, sg into the other of two
// buffer sets (it reads them at neighbouring prisms).  p and e are needed
// only at the launch's own prisms, and only every other iteration touches
// them: it forms the p the iteration before left pending from that
// iteration's u (rt of the other set) and the stored p, then its own p,
// and adds both alpha p terms to e (p and e in place): 24 or 44 B per prism,
// 34 B on average, against 40 B in the two kernels of NSX_REFINE_CG, in one
// pass with one reduction.  In exact arithmetic the iterates are CG's.
//
// Reliable updates as in NSX_REFINE_CG: when an iteration's residual |r|
// has fallen below delta x its largest value since the last true residual,
// or meets the stopping rule, the next launch (mode T) forms the true FP64
// residual of x + e (with any pending alpha p, p formed as above) into rt;
// the launch after it (mode M) folds e (and the pending alpha p) into x,
// applies the stopping rule to the true residual, forms om = A rt / A0 with
// gamma and delta, and sg = A p / A0 afresh (the recurrence for s = A p
// drifts in FP32, and with it the recursive residual; without the reset the
// solve stalls once updates come less often than every decade).  beta,
// alpha carry on across the update.  The first
// launch of a solve is an M without the fold, after k_cg_start.
//
// Launch j reads the carried scalars from SolverState::cgs[(j-1)&1] and
// block 0 writes cgs[j&1]; the partial sums alternate between two buffers
// the same way, so no block reads a value that its own launch rewrites.
// The launches queued after the stop write their slot too (the stop passed
// on), or the one after would read the slot of a launch before the stop and
// run again from it.
// Partial rows: M gamma, delta; N gamma, delta, |r|^2, |r|_1; T |r|^2,
// |r|_1, sum (x+e) Vol.

// The 4-layer block-Jacobi preconditioner of the single-reduction CG: M is A
// restricted to the interior layers of one quad (the four layers a thread
// owns), with the diagonal A's rows have at the sigma ends (the Neumann
// reflection of apply_row); its couplings to the layers of other quads are
// dropped.  Rows outside the interior are the identity, so the solves below
// give them zero.  M is symmetric (AB of a layer is AT of the layer below,
// the same quotient) and negative definite like A's diagonal blocks.
template <typename T>
struct Blk4 {
  T d[4], lo[4], up[4];
};

template <typename T>
__device__ inline Blk4<T> blk4_of(const Geo &g, const Row1<T> (&R)[4], int k0) {
  Blk4<T> B;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int k = k0 + u;
    if (!R[u].in) {
      B.d[u] = T(1);
      B.lo[u] = B.up[u] = T(0);
      continue;
    }
    T d = R[u].A0;
    if (k + 1 > g.NZ - 2) d = d + R[u].AT;
    if (k - 1 < 1) d = d + R[u].AB;
    B.d[u] = d;
    B.lo[u] = (u > 0 && R[u - 1].in) ? R[u].AB : T(0);
    B.up[u] = (u < 3 && R[u + 1].in) ? R[u].AT : T(0);
  }
  return B;
}

// y = M^-1 f (Thomas, four rows)
// the reciprocal of the preconditioner solves: in FP32 the hardware's
// (v_rcp_f32, about 1 ulp) rather than a correctly rounded division, in
// FP64 the division.  M^-1 needs no correct rounding, only the same
// arithmetic in every launch.
__device__ inline float prec_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline double prec_rcp(double x) { return 1.0 / x; }

// a b + c in one rounding
__device__ inline float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ inline double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <typename T>
__device__ inline void blk4_solve(const Blk4<T> &B, const T (&f)[4], T (&y)[4]) {
  // one reciprocal per row: it scales both the row's factor and its
  // right-hand side
  T c[4], z[4];
  T im = prec_rcp(B.d[0]);
  c[0] = B.up[0] * im;
  z[0] = f[0] * im;
#pragma unroll
  for (int u = 1; u < 4; ++u) {
    im = prec_rcp(fma_t(-B.lo[u], c[u - 1], B.d[u]));
    c[u] = B.up[u] * im;
    z[u] = fma_t(-B.lo[u], z[u - 1], f[u]) * im;
  }
  y[3] = z[3];
#pragma unroll
  for (int u = 2; u >= 0; --u) y[u] = fma_t(-c[u], y[u + 1], z[u]);
}

// y = M x
template <typename T>
__device__ inline void blk4_mul(const Blk4<T> &B, const T (&x)[4], T (&y)[4]) {
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    T t = B.d[u] * x[u];
    if (u > 0) t = fma_t(B.lo[u], x[u - 1], t);
    if (u < 3) t = fma_t(B.up[u], x[u + 1], t);
    y[u] = t;
  }
}

// e with the pending alpha p.  The true-residual launch forms it at every
// slot and the fold after it forms it again: one fused multiply-add in both,
// so x takes exactly the values whose residual was tested (an unfused sum in
// one of them moves x by a rounding of e, which A can amplify past the
// stopping rule).
template <typename T>
__device__ inline T e_pending(T e, T p, T a) { return fma_t(a, p, e); }

// p_i = u_i + beta_i p_{i-1} of an iteration that stored no p (u_i alone
// when it was the first): the same expression wherever it is formed
template <typename T>
__device__ inline T p_lagged(T u, T p, T b, bool first) { return first ? u : u + b * p; }

// The line preconditioner of the single-reduction CG: M is A restricted to
// the interior layers of one column (the vertical couplings, and the diagonal
// with the Neumann reflection at the sigma ends); the lateral couplings are
// dropped.  Line4 holds a quad's four rows: lo[0] couples layer k0 to k0-1
// (the last row of the quad below), up[3] layer k0+3 to k0+4.  Rows outside
// the interior are the identity.  M is symmetric and negative definite, and
// strictly diagonally dominant (its diagonal carries the lateral terms).
struct Line4 {
  float d[4], lo[4], up[4];
};

__device__ inline Line4 line4_identity() {
  Line4 L;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    L.d[u] = 1.f;
    L.lo[u] = L.up[u] = 0.f;
  }
  return L;
}

__device__ inline Line4 line4_of(const Geo &g, const Row1<float> (&R)[4], int k0) {
  Line4 L = line4_identity();
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int k = k0 + u;
    if (!R[u].in) continue;
    float d = R[u].A0;
    if (k + 1 > g.NZ - 2) d = d + R[u].AT;
    if (k - 1 < 1) d = d + R[u].AB;
    L.d[u] = d;
    L.lo[u] = k - 1 >= 1 ? R[u].AB : 0.f;
    L.up[u] = k + 1 <= g.NZ - 2 ? R[u].AT : 0.f;
  }
  return L;
}

// y = M x on a quad, with x at k0-1 and k0+4 in xm and xp
__device__ inline void line4_mul(const Line4 &L, const float (&x)[4], float xm, float xp, float (&y)[4]) {
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    float t = L.d[u] * x[u];
    t = t + L.lo[u] * (u > 0 ? x[u > 0 ? u - 1 : 0] : xm);
    t = t + L.up[u] * (u < 3 ? x[u < 3 ? u + 1 : 3] : xp);
    y[u] = t;
  }
}

// quads per window of the line preconditioner: whole columns
__device__ inline uint32_t line_window(const Geo &g) { return (uint32_t)(NSX_BLOCK / g.NQ) * (uint32_t)g.NQ; }

// the position of this thread's quad in its column, in windows of W quads
// (whole columns from a column's start); -1 past the window's columns
__device__ inline int line_pos(const Geo &g, uint32_t W) {
  return threadIdx.x < W ? (int)(threadIdx.x % (uint32_t)g.NQ) : -1;
}

// y = M^-1 f over the columns of a window of whole columns.  Every thread of
// the block calls it (it synchronises the block); pos is the position of the
// thread's quad in its column (0 .. NQ-1), -1 for a thread without a quad,
// whose rows are the identity.  The threads of a column are consecutive
// lanes, so the neighbours of the quad at `pos` are lanes -1 and +1.
//
// Each quad's rows 0-2 are solved for in terms of the last unknown of the
// quad below (Y-) and its own last unknown (Y): y_u = z_u + a_u Y- + b_u Y.
// Row 3 with quad+1's y_0 substituted is then a tridiagonal system in the
// column's last unknowns, one per quad, solved by parallel cyclic reduction
// (ceil(log2 NQ) steps); rows 0-2 follow from Y- and Y.
__device__ inline void line_solve(const Line4 &L, const float (&f)[4], float (&y)[4], int pos, int nq) {
  __shared__ float xz[3][NSX_BLOCK], pr[2][4][NSX_BLOCK], yl[NSX_BLOCK];
  const int lane = threadIdx.x;
  const bool on = pos >= 0;
  // rows 0-2: Thomas with three right-hand sides (f, -lo0 e0, -up2 e2)
  const float i0 = prec_rcp(L.d[0]);
  const float c0 = L.up[0] * i0;
  const float i1 = prec_rcp(L.d[1] - L.lo[1] * c0);
  const float c1 = L.up[1] * i1;
  const float i2 = prec_rcp(L.d[2] - L.lo[2] * c1);
  float z0 = f[0] * i0, z1 = (f[1] - L.lo[1] * z0) * i1, z2 = (f[2] - L.lo[2] * z1) * i2;
  float a0 = -L.lo[0] * i0, a1 = -L.lo[1] * a0 * i1, a2 = -L.lo[2] * a1 * i2;
  const float b2 = -L.up[2] * i2;
  z1 = z1 - c1 * z2;
  z0 = z0 - c0 * z1;
  a1 = a1 - c1 * a2;
  a0 = a0 - c0 * a1;
  const float b1 = -c1 * b2, b0 = -c0 * b1;
  xz[0][lane] = z0;
  xz[1][lane] = a0;
  xz[2][lane] = b0;
  __syncthreads();
  float zn = 0.f, an = 0.f, bn = 0.f;
  if (on && pos + 1 < nq) {
    zn = xz[0][lane + 1];
    an = xz[1][lane + 1];
    bn = xz[2][lane + 1];
  }
  // row 3: lo3 y2 + d3 Y + up3 y0(quad+1) = f3
  float ra = 0.f, rb = 1.f, rc = 0.f, rg = 0.f;
  if (on) {
    ra = L.lo[3] * a2;
    rb = L.d[3] + L.lo[3] * b2 + L.up[3] * an;
    rc = L.up[3] * bn;
    rg = f[3] - L.lo[3] * z2 - L.up[3] * zn;
  }
  int buf = 0;
  for (int h = 1; h < nq; h <<= 1, buf ^= 1) {
    pr[buf][0][lane] = ra;
    pr[buf][1][lane] = rb;
    pr[buf][2][lane] = rc;
    pr[buf][3][lane] = rg;
    __syncthreads();
    if (on) {
      float k1 = 0.f, k2 = 0.f, am = 0.f, cm = 0.f, gm = 0.f, ap = 0.f, cp = 0.f, gp = 0.f;
      if (pos >= h) {
        am = pr[buf][0][lane - h];
        k1 = ra * prec_rcp(pr[buf][1][lane - h]);
        cm = pr[buf][2][lane - h];
        gm = pr[buf][3][lane - h];
      }
      if (pos + h < nq) {
        ap = pr[buf][0][lane + h];
        k2 = rc * prec_rcp(pr[buf][1][lane + h]);
        cp = pr[buf][2][lane + h];
        gp = pr[buf][3][lane + h];
      }
      rb = rb - cm * k1 - ap * k2;
      rg = rg - gm * k1 - gp * k2;
      ra = -am * k1;
      rc = -cp * k2;
    }
  }
  const float Y = rg * prec_rcp(rb);
  yl[lane] = Y;
  __syncthreads();
  const float Ym = on && pos >= 1 ? yl[lane - 1] : 0.f;
  y[0] = z0 + a0 * Ym + b0 * Y;
  y[1] = z1 + a1 * Ym + b1 * Y;
  y[2] = z2 + a2 * Ym + b2 * Y;
  y[3] = Y;
}

// out = M^-1 A f, f = a + c b (a when first), with the line preconditioner,
// over windows of whole columns; with acc, the partial sums <M f, f> and
// <A f, f> over the interior
__device__ inline void line_product(const Geo &g, const float *lv, const float *a, const float *b, float cc,
                                    bool first, float *out, double *acc, bool rev) {
  const uint32_t W = line_window(g);
  const int lpos = line_pos(g, W);
  Line4 L = line4_identity();
  float wv[4] = {0.f, 0.f, 0.f, 0.f};
  int k0s = 0;
  size_t s0s = 0;
  auto v4 = [&](size_t off, float(&o)[4]) { axpy4(a, b, cc, first, off, o); };
  auto v1 = [&](size_t off) { return axpy1(a, b, cc, first, off); };
  for_each_quad_formed_w<float>(
      g, W, v4, v4, v1, [&](size_t) {},
      [&](int, int k0, size_t s0, const Col<float> &C, const float(&fc)[4], float fm, float fp,
          const float(&f0)[4], const float(&f1)[4], const float(&f2)[4]) {
        float qv[5], mr[4];
        quot5_r(C, lv, g.NS, k0, qv);
        Row1<float> R[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          R[u] = row_of(g, lv, C, qv, k0, u);
          wv[u] = apply_row(g, R[u], k0, u, fc, fm, fp, f0[u], f1[u], f2[u]);
        }
        L = line4_of(g, R, k0);
        if (acc) {
          line4_mul(L, fc, fm, fp, mr);
#pragma unroll
          for (int u = 0; u < 4; ++u)
            if (R[u].in) {
              acc[0] += (double)mr[u] * (double)fc[u];
              acc[1] += (double)wv[u] * (double)fc[u];
            }
        }
        k0s = k0;
        s0s = s0;
      },
      [&](bool act) {
        float wo[4];
        if (!act) {
          L = line4_identity();
#pragma unroll
          for (int u = 0; u < 4; ++u) wv[u] = 0.f;
        }
        line_solve(L, wv, wo, act ? lpos : -1, g.NQ);
        if (act) {
#pragma unroll
          for (int u = 0; u < 4; ++u)
            if (k0s + u < 1 || k0s + u > g.NZ - 2) wo[u] = 0.f;
          st4(out + s0s, wo);
        }
      },
      rev);
}

// PC: the preconditioner, 0 point Jacobi, 1 the 4-layer blocks, 2 the
// lines (above).  With blocks or lines the solve starts with a true-residual
// launch, which forms u = M^-1 r0; with lines every phase runs on windows of
// whole columns.
template <typename T, int PC>
__global__ __launch_bounds__(NSX_BLOCK) void k_cgs(Geo g, int j, CgsVecsT<T> v, const double *rhs, double *x,
                                                    const double *Pin, uint32_t Gin, double *Pout, SolverState *st) {
  constexpr bool BLK = PC == 1, LINE = PC == 2, PRE = PC != 0;
  CgsCarry c;
  if (j == 0) {
    c.gamma = c.alpha = c.beta = c.maxr = 0.0;
    c.mode = -1;
    c.iter = c.set = c.pend = c.pend_e = c.first = c.pad = 0;
    c.done = st->done;  // k_cg_start's no-solve test (launch 0 itself never stops)
  } else {
    c = st->cgs[(j - 1) & 1];
  }
  if (c.done) {
    // pass the stop on: the launch after this one reads this launch's slot
    if (blockIdx.x == 0 && threadIdx.x == 0) st->cgs[j & 1] = c;
    return;
  }
  const int prev = c.mode;
  int mode = 0, stop = 0, code = EXIT_CONVERGED;
  bool fold = false, fold_x = true;
  double gam = 0.0, del = 0.0;
  if (prev == -1) {
    double t[4];
    reduce_partials<4>(Pin, Gin, t);  // k_cg_start: rho, |r|^2, |r|_1, sum x Vol
    c.maxr = sqrt(t[1]);
    if (PRE) mode = 2;  // u0 = M^-1 r0 from a true residual
  } else if (prev == 2) {
    double t[3];
    reduce_partials<3>(Pin, Gin, t);
    const double rN = sqrt(t[0]), e1 = t[1] / st->b1, e2 = rN / st->b2;
    fold = true;
    if (!isfinite(t[0]) || !isfinite(t[1])) {
      // the correction has overflowed: x keeps its last value (the fold
      // below only clears e) and the FP64 solve finishes from it
      fold_x = false;
      stop = 1;
      code = EXIT_BREAKDOWN;
    } else if (e1 < st->eps || e2 < st->eps) {
      stop = 1;
    }
    c.maxr = rN;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->e1 = e1;
      st->e2 = e2;
      st->xvol = t[2];
      if (c.iter > 0) st->mp_pass += 1;  // not the start of a preconditioned solve
    }
  } else if (prev == 0) {
    double t[2];
    reduce_partials<2>(Pin, Gin, t);
    gam = t[0];
    del = t[1];
    mode = 1;
  } else {
    double t[4];
    reduce_partials<4>(Pin, Gin, t);
    gam = t[0];
    del = t[1];
    const double rN = sqrt(t[2]), e1 = t[3] / st->b1, e2 = rN / st->b2;
    c.maxr = fmax(c.maxr, rN);
    if (!isfinite(t[2])) {
      stop = 1;
      code = EXIT_BREAKDOWN;
    }
    mode = (e1 < st->eps || e2 < st->eps || rN < st->delta * c.maxr) ? 2 : 1;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->e1 = e1;
      st->e2 = e2;
    }
  }
  double al = 0.0, be = 0.0;
  const bool first = c.iter == 0;
  if (mode == 1 && !stop) {
    if (c.iter >= st->max_iters) {
      stop = 1;
      code = EXIT_MAX;
    } else {
      be = first ? 0.0 : gam / c.gamma;
      al = first ? gam / del : gam / (del - be * gam / c.alpha);
      if (gam == 0.0 || !isfinite(gam) || !isfinite(al) || al == 0.0) {
        stop = 1;
        code = EXIT_BREAKDOWN;
      }
    }
  }
  const int in = c.set;
  // e and the pending alpha p as they stand when this launch starts
  // p and e as they stand when this launch starts: when pend_in, the last
  // iteration i stored neither; p_i = u_i + beta_i p_{i-1} from u_i (rt of
  // the other set, which no launch rewrites before the next iteration) and
  // the stored p_{i-1}, and e lacks alpha_i p_i unless a fold took it into x
  const bool pend_in = c.pend != 0, pend_e_in = c.pend_e != 0, pfirst_in = c.first != 0;
  const T ap_in = (T)c.alpha, bp_in = (T)c.beta;
  const T *rt_lag = v.rt[in ^ 1];
  if (mode == 1 && !stop) {
    c.gamma = gam;
    c.alpha = al;
    c.beta = first ? 0.0 : be;
    c.first = first ? 1 : 0;
    c.iter += 1;
    c.set ^= 1;
    // iterations alternate: leave p and e alone, then form both terms
    c.pend = c.pend_e = pend_in ? 0 : 1;
  }
  if (fold) c.pend_e = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    CgsCarry o = c;
    o.mode = mode;
    o.done = stop;
    st->cgs[j & 1] = o;
    st->cgs_runs = j + 1;
    st->a_iter = c.iter;
    if (stop) {
      st->exit_code = code;
      st->done = 1;
    }
  }
  if (fold) {
    // x += e (+ the pending alpha p), e = 0: the correction since the last
    // update joins x (interior layers: e holds zero elsewhere)
    for_each_quad(g, [&](int, int k0, size_t s0) {
      T ee[4];
      ld4(v.e + s0, ee);
      if (pend_e_in) {
        T ua[4], pp[4] = {T(0), T(0), T(0), T(0)};
        ld4(rt_lag + s0, ua);
        if (!pfirst_in) ld4(v.p + s0, pp);
#pragma unroll
        for (int u = 0; u < 4; ++u) ee[u] = e_pending(ee[u], p_lagged(ua[u], pp[u], bp_in, pfirst_in), ap_in);
      }
      if (fold_x)
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (k0 + u >= 1 && k0 + u <= g.NZ - 2) x[s0 + u] = x[s0 + u] + (double)ee[u];
      const T z[4] = {T(0), T(0), T(0), T(0)};
      st4(v.e + s0, z);
    });
  }
  if (stop) return;
  // odd launches sweep the windows from the far end (window_sweep)
  const bool rev = (j & 1) != 0;
  T *lvw = load_levels<T>(g);
  for (int k = threadIdx.x; k < g.NS; k += NSX_BLOCK) lvw[2 * g.NS + k] = T(1) / lvw[k];
  __syncthreads();
  const T *lv = lvw;
  if (mode == 2) {
    // the true residual of x + e into rt (T reads x and e at neighbouring
    // prisms and writes rt only at its own)
    const double avec = st->avec;
    const T *e = v.e, *p = v.p;
    T *rt = v.rt[in];
    double acc[3] = {0.0, 0.0, 0.0};
    // e at a slot, with the alpha p the last iteration left pending (the fold
    // of the next launch forms the same sum)
    auto ecur = [&](size_t s, bool) {
      return pend_e_in ? e_pending(e[s], p_lagged(rt_lag[s], pfirst_in ? T(0) : p[s], bp_in, pfirst_in), ap_in) : e[s];
    };
    if (PRE) {
      // window by window: the residuals cell by cell (consecutive threads on
      // consecutive slots, as the vectors are stored, so the neighbour
      // gathers coalesce) into LDS, then u = M^-1 r over each quad's block or
      // each column's line.  A window of W quads from quad `base` covers the
      // slots 4 base .. 4 (base + W) - 1.
      __shared__ __align__(16) T rs[NSX_BLOCK][4];
      const uint32_t W = LINE ? line_window(g) : NSX_BLOCK, nwin = (g.nquad + W - 1) / W;
      const int lpos = LINE ? line_pos(g, W) : -1;
      const Sweep S = window_sweep(nwin, rev);
      for (uint32_t wi = S.w; wi < S.end; wi += S.step) {
        const uint32_t base = S.at(wi) * W;
        const uint32_t lim = (base + W < g.nquad ? base + W : g.nquad) - base;
        T *flat = &rs[0][0];
        for (uint32_t m = threadIdx.x; m < 4 * W; m += NSX_BLOCK) {
          T f = T(0);
          if (m < 4 * lim) {
            const uint32_t q = base + (m >> 2);
            const int i = (int)fdiv(q, g.divNQ);
            const int k = 4 * (int)(q - (uint32_t)i * (uint32_t)g.NQ) + (int)(m & 3);
            if (k >= 1 && k <= g.NZ - 2) {
              double res, xs;
              true_residual(g, rhs, x, avec, i, k, ecur, res, xs);
              f = (T)res;
              acc[0] += res * res;
              acc[1] += fabs(res);
              acc[2] += xs * (g.area[i] * g.dsig[k]);
            }
          }
          flat[m] = f;
        }
        __syncthreads();
        const bool act = threadIdx.x < lim;
        T f[4] = {T(0), T(0), T(0), T(0)}, u4[4] = {T(0), T(0), T(0), T(0)};
        Line4 L = line4_identity();
        if (act) {
          const uint32_t t = base + threadIdx.x;
          const int i = (int)fdiv(t, g.divNQ);
          const int k0 = 4 * (int)(t - (uint32_t)i * (uint32_t)g.NQ);
#pragma unroll
          for (int u = 0; u < 4; ++u) f[u] = rs[threadIdx.x][u];
          Col<T> C;
          load_col(g, i, C);
          T qv[5];
          quot5_r(C, lv, g.NS, k0, qv);
          Row1<T> R[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) R[u] = row_of(g, lv, C, qv, k0, u);
          if constexpr (LINE)
            L = line4_of(g, R, k0);
          else
            blk4_solve(blk4_of(g, R, k0), f, u4);
        }
        if constexpr (LINE) line_solve(L, f, u4, act ? lpos : -1, g.NQ);
        if (act) st4(rt + (size_t)4 * (base + threadIdx.x), u4);
        __syncthreads();  // rs is rewritten by the next window
      }
      write_partials<3>(acc, Pout);
      return;
    }
    FOR_EACH_CELL(g, q) {
      int i, k;
      cell_of(g, q, i, k);
      double res, xs;
      true_residual(g, rhs, x, avec, i, k, ecur, res, xs);
      const Row A = ppe_row(g, i, k);
      rt[(size_t)i * g.NS + k] = (T)(res / A.A0);
      acc[0] += res * res;
      acc[1] += fabs(res);
      acc[2] += xs * (g.area[i] * g.dsig[k]);
    }
    write_partials<3>(acc, Pout);
    return;
  }
  if (mode == 0) {
    // om = A rt / A0, gamma = <A0 rt, rt>, delta = <A rt, rt>
    const T *rt = v.rt[in];
    T *om = v.om[in];
    double acc[2] = {0.0, 0.0};
    if constexpr (LINE) {
      line_product(g, lv, rt, rt, T(0), true, om, acc, rev);
      if (fold && c.iter > 0)
        line_product(g, lv, pend_in ? rt_lag : (const T *)v.p, (const T *)v.p, pend_in ? bp_in : T(0),
                     pend_in ? pfirst_in : true, v.sg[in], (double *)nullptr, rev);
    } else {
      for_each_quad_gathered(
          g, rt, rt, T(0), true, [&](size_t) {},
          [&](int, int k0, size_t s0, const Col<T> &C, const T(&fc)[4], T fm, T fp,
              const T(&f0)[4], const T(&f1)[4], const T(&f2)[4]) {
            T qv[5], wo[4];
            quot5_r(C, lv, g.NS, k0, qv);
            if constexpr (BLK) {
              Row1<T> R[4];
              T wv[4], mr[4];
  #pragma unroll
              for (int u = 0; u < 4; ++u) {
                R[u] = row_of(g, lv, C, qv, k0, u);
                wv[u] = apply_row(g, R[u], k0, u, fc, fm, fp, f0[u], f1[u], f2[u]);
              }
              const auto B = blk4_of(g, R, k0);
              blk4_solve(B, wv, wo);
              blk4_mul(B, fc, mr);
  #pragma unroll
              for (int u = 0; u < 4; ++u)
                if (R[u].in) {
                  acc[0] += (double)mr[u] * (double)fc[u];
                  acc[1] += (double)wv[u] * (double)fc[u];
                }
            } else {
  #pragma unroll
              for (int u = 0; u < 4; ++u) {
                const Row1<T> R = row_of(g, lv, C, qv, k0, u);
                const T wv = apply_row(g, R, k0, u, fc, fm, fp, f0[u], f1[u], f2[u]);
                wo[u] = R.in ? wv * prec_rcp(R.A0) : T(0);
                if (R.in) {
                  acc[0] += (double)R.A0 * (double)fc[u] * (double)fc[u];
                  acc[1] += (double)wv * (double)fc[u];
                }
              }
            }
            st4(om + s0, wo);
          },
          rev);
      if (fold && c.iter > 0) {
        // after a true residual, s = A p afresh: its recurrence drifts in FP32
        // from the product it stands for, and with it the recursive residual
        // (p formed from u and the stored p when it lags)
        T *sg = v.sg[in];
        for_each_quad_gathered(
            g, pend_in ? rt_lag : (const T *)v.p, (const T *)v.p, pend_in ? bp_in : T(0),
            pend_in ? pfirst_in : true, [&](size_t) {},
            [&](int, int k0, size_t s0, const Col<T> &C, const T(&fc)[4], T fm, T fp,
                const T(&f0)[4], const T(&f1)[4], const T(&f2)[4]) {
              T qv[5], so[4];
              quot5_r(C, lv, g.NS, k0, qv);
              if (BLK) {
                Row1<T> R[4];
                T sv[4];
  #pragma unroll
                for (int u = 0; u < 4; ++u) {
                  R[u] = row_of(g, lv, C, qv, k0, u);
                  sv[u] = apply_row(g, R[u], k0, u, fc, fm, fp, f0[u], f1[u], f2[u]);
                }
                blk4_solve(blk4_of(g, R, k0), sv, so);
              } else {
  #pragma unroll
                for (int u = 0; u < 4; ++u) {
                  const Row1<T> R = row_of(g, lv, C, qv, k0, u);
                  const T sv = apply_row(g, R, k0, u, fc, fm, fp, f0[u], f1[u], f2[u]);
                  so[u] = R.in ? sv * prec_rcp(R.A0) : T(0);
                }
              }
              st4(sg + s0, so);
            },
            rev);
      }
    }
    write_partials<2>(acc, Pout);
    return;
  }
  // mode N: iteration i with alpha_i, beta_i
  const T af = (T)al, bf = (T)be;
  const T *rt_i = v.rt[in], *om_i = v.om[in], *sg_i = v.sg[in];
  T *rt_o = v.rt[in ^ 1], *om_o = v.om[in ^ 1], *sg_o = v.sg[in ^ 1];
  // u_{i+1} at a slot, from rt, om, sg there: the same expressions wherever
  // it is formed
  auto form4 = [&](size_t off, T(&r)[4], T(&sg)[4], T(&o)[4]) {
    T w[4], q[4] = {T(0), T(0), T(0), T(0)};
    ld4(rt_i + off, r);
    ld4(om_i + off, w);
    if (!first) ld4(sg_i + off, q);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      sg[u] = first ? w[u] : fma_t(bf, q[u], w[u]);
      o[u] = fma_t(-af, sg[u], r[u]);
    }
  };
  auto val4 = [&](size_t off, T(&o)[4]) {
    T r[4], sg[4];
    form4(off, r, sg, o);
  };
  auto val1 = [&](size_t q) {
    const T sg = first ? om_i[q] : fma_t(bf, sg_i[q], om_i[q]);
    return fma_t(-af, sg, rt_i[q]);
  };
  // p and e are read and written every other iteration: an iteration that
  // finds them current leaves both alone (p_i = u_i + beta_i p_{i-1} and
  // alpha_i p_i become pending); the next forms p_i from u_i, still in rt of
  // its output set, and the stored p_{i-1}, then p_{i+1}, stores p_{i+1} and
  // adds alpha_i p_i + alpha_{i+1} p_{i+1} to e (only the second term after
  // a fold took the first into x).  The first iteration never finds them
  // lagging.
  const bool upd_p = pend_in;
  T ri[4], si[4], pi[4] = {T(0), T(0), T(0), T(0)}, ei[4], ui[4];
  auto own4 = [&](size_t off, T(&o)[4]) { form4(off, ri, si, o); };
  auto pre = [&](size_t off) {
    if (upd_p) {
      ld4(rt_lag + off, ui);  // own slots of rt_o, read before this thread stores them
      if (!pfirst_in) ld4(v.p + off, pi);
      ld4(v.e + off, ei);
    }
  };
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  // with lines, om = M^-1 w waits for the window's columns (post)
  Line4 Ln = line4_identity();
  T wvn[4] = {T(0), T(0), T(0), T(0)};
  int k0n = 0;
  size_t s0n = 0;
  const uint32_t W = LINE ? line_window(g) : NSX_BLOCK;
  const int lpos = LINE ? line_pos(g, W) : -1;
  for_each_quad_formed_w<T>(
      g, W, own4, val4, val1, pre,
      [&](int, int k0, size_t s0, const Col<T> &C, const T(&fc)[4], T fm, T fp,
          const T(&f0)[4], const T(&f1)[4], const T(&f2)[4]) {
        T qv[5], ro[4], wo[4], so[4], po[4], eo[4];
        quot5_r(C, lv, g.NS, k0, qv);
        Row1<T> R[4];
        T wv[4], mr[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          R[u] = row_of(g, lv, C, qv, k0, u);
          wv[u] = apply_row(g, R[u], k0, u, fc, fm, fp, f0[u], f1[u], f2[u]);
        }
        if constexpr (LINE) {
          // r = M u for gamma and the norms
          Ln = line4_of(g, R, k0);
          line4_mul(Ln, fc, fm, fp, mr);
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            wvn[u] = wv[u];
            wo[u] = T(0);
          }
          k0n = k0;
          s0n = s0;
        } else if constexpr (BLK) {
          // om = M^-1 w, and r = M u for gamma and the norms
          const auto B = blk4_of(g, R, k0);
          blk4_solve(B, wv, wo);
          blk4_mul(B, fc, mr);
        } else {
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            wo[u] = R[u].in ? wv[u] * prec_rcp(R[u].A0) : T(0);
            mr[u] = R[u].A0 * fc[u];
          }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const bool in = R[u].in;
          if (upd_p) {
            const T pl = p_lagged(ui[u], pi[u], bp_in, pfirst_in);
            const T pn = fma_t(bf, pl, ri[u]);
            const T eb = pend_e_in ? fma_t(ap_in, pl, ei[u]) : ei[u];
            eo[u] = in ? fma_t(af, pn, eb) : ei[u];
            po[u] = in ? pn : T(0);
          }
          ro[u] = in ? fc[u] : T(0);
          so[u] = in ? si[u] : T(0);
          wo[u] = in ? wo[u] : T(0);
          if (in) {
            const double rd = PRE ? (double)mr[u] : (double)R[u].A0 * (double)fc[u];
            acc[0] += rd * (double)fc[u];
            acc[1] += (double)wv[u] * (double)fc[u];
            acc[2] += rd * rd;
            acc[3] += fabs(rd);
          }
        }
        if (upd_p) {
          st4(v.e + s0, eo);
          st4(v.p + s0, po);
        }
        st4(rt_o + s0, ro);
        if (!LINE) st4(om_o + s0, wo);
        st4(sg_o + s0, so);
      },
      [&](bool act) {
        if constexpr (LINE) {
          T wo[4];
          if (!act) {
            Ln = line4_identity();
#pragma unroll
            for (int u = 0; u < 4; ++u) wvn[u] = T(0);
          }
          line_solve(Ln, wvn, wo, act ? lpos : -1, g.NQ);
          if (act) {
#pragma unroll
            for (int u = 0; u < 4; ++u)
              if (k0n + u < 1 || k0n + u > g.NZ - 2) wo[u] = T(0);
            st4(om_o + s0n, wo);
          }
        }
      },
      rev);
  write_partials<4>(acc, Pout);
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
// k_cgs: dsig, dsigv and 1/dsig in the vectors' precision
static size_t cgs_lds(const Geo &g, size_t elem = sizeof(float)) { return 3 * (size_t)g.NS * elem; }

int launch_bicg_init(const Geo &g, hipStream_t s, const double *rhs, const double *phi_in, double *x, double *r,
                     double *rhat, Partials part, SolverState *st, double eps, int max_iters) {
  const dim3 grid(g.grid), blk(NSX_BLOCK);
  hipLaunchKernelGGL(k_bicg_init_state, dim3(1), blk, 0, s, st, eps, max_iters, 0.0);
  hipLaunchKernelGGL(k_resid<double>, grid, blk, 0, s, g, rhs, phi_in, x, r, rhat, (double *)nullptr, part,
                     part.init, (double *)nullptr, 1, st);
  return (int)hipGetLastError();
}

template <typename T>
static int iter_impl(const Geo &g, hipStream_t s, T *x, T *rt, const T *rhat, T *q, T *vt, T *tv, Partials part,
                     SolverState *st, hipEvent_t *ev) {
  const dim3 grid(g.grid), blk(NSX_BLOCK);
  const size_t lds = level_lds(g, sizeof(T));
  LAUNCH_TIMED(ev, 0, k_bicg_a<T>, grid, blk, lds, s, g, (const T *)rt, (const T *)q, rhat, vt, part, st);
  LAUNCH_TIMED(ev, 1, k_bicg_b<T>, grid, blk, lds, s, g, (const T *)rt, (const T *)vt, tv, part, st);
  LAUNCH_TIMED(ev, 2, k_bicg_c<T>, grid, blk, lds, s, g, x, rt, q, rhat, (const T *)vt, (const T *)tv, part, st);
  return (int)hipGetLastError();
}

int launch_bicg_iter(const Geo &g, hipStream_t s, double *x, double *rt, const double *rhat, double *q, double *vt,
                     double *tv, Partials part, SolverState *st, hipEvent_t *ev) {
  return iter_impl<double>(g, s, x, rt, rhat, q, vt, tv, part, st, ev);
}

int launch_cg_init(const Geo &g, hipStream_t s, const double *rhs, const double *phi_in, double *x, float *rt,
                   float *scratch, float *e, Partials part, SolverState *st, double eps, double delta, int max_iters) {
  const dim3 grid(g.grid), blk(NSX_BLOCK);
  hipLaunchKernelGGL(k_bicg_init_state, dim3(1), blk, 0, s, st, eps, max_iters, delta);
  // r0 = b - A x0 and its norms (P.init), e = 0, x = x0; rt, rho_0 into P.u
  hipLaunchKernelGGL(k_resid<float>, grid, blk, 0, s, g, rhs, phi_in, x, rt, scratch, e, part, part.init,
                     (double *)nullptr, phi_in != x ? 1 : 0, st);
  hipLaunchKernelGGL(k_cg_start<float>, grid, blk, 0, s, g, rhs, (const double *)x, rt, part, st);
  return (int)hipGetLastError();
}

int launch_cg_init64(const Geo &g, hipStream_t s, const double *rhs, const double *phi_in, double *x, double *rt,
                     double *scratch, double *e, Partials part, SolverState *st, double eps, double delta,
                     int max_iters) {
  const dim3 grid(g.grid), blk(NSX_BLOCK);
  hipLaunchKernelGGL(k_bicg_init_state, dim3(1), blk, 0, s, st, eps, max_iters, delta);
  hipLaunchKernelGGL(k_resid<double>, grid, blk, 0, s, g, rhs, phi_in, x, rt, scratch, e, part, part.init,
                     (double *)nullptr, phi_in != x ? 1 : 0, st);
  hipLaunchKernelGGL(k_cg_start<double>, grid, blk, 0, s, g, rhs, (const double *)x, rt, part, st);
  return (int)hipGetLastError();
}

int launch_cg_iter(const Geo &g, hipStream_t s, const double *rhs, double *x, float *const v[7], Partials part,
                   SolverState *st, hipEvent_t *ev) {
  // v: e, rt, p, w, e_b, rt_b, p_b.  The stencil kernel W is latency-bound and
  // runs with every block its occupancy allows; the streaming kernel X runs
  // on the solver grid (more concurrent streams slow it down)
  float *e = v[0], *rt = v[1], *p = v[2], *w = v[3], *e_b = v[4], *rt_b = v[5], *p_b = v[6];
  const dim3 grid(g.grid), grid_w(g.grid_cg), blk(NSX_BLOCK);
  const size_t lds = level_lds(g, sizeof(float));
  LAUNCH_TIMED(ev, 0, k_cg_w, grid_w, blk, lds, s, g, (const float *)rt, (const float *)p, (const float *)rt_b,
               (const float *)p_b, (const float *)e_b, w, x, part, st);
  LAUNCH_TIMED(ev, 1, k_cg_x, grid, blk, lds, s, g, e, rt, p, e_b, rt_b, p_b, (const float *)w, rhs,
               (const double *)x, part, st);
  return (int)hipGetLastError();
}

int launch_cgs(const Geo &g, hipStream_t s, int j, const double *rhs, double *x, const CgsVecs &v, Partials part,
               SolverState *st, hipEvent_t *ev, int pc) {
  // launch 0 reads k_cg_start's partials (part.u, solver grid); then the
  // launches alternate between part.c (even) and part.u (odd) on their grid
  double *Pout = (j & 1) ? part.u : part.c;
  const double *Pin = j == 0 ? part.u : ((j & 1) ? part.c : part.u);
  const uint32_t Gin = (uint32_t)(j == 0 ? g.grid : g.grid_cgs);
  const size_t lds = cgs_lds(g);
  if (pc == 2)
    LAUNCH_TIMED(ev, 0, (k_cgs<float, 2>), dim3(g.grid_cgs), dim3(NSX_BLOCK), lds, s, g, j, v, rhs, x, Pin, Gin,
                 Pout, st);
  else if (pc == 1)
    LAUNCH_TIMED(ev, 0, (k_cgs<float, 1>), dim3(g.grid_cgs), dim3(NSX_BLOCK), lds, s, g, j, v, rhs, x, Pin, Gin,
                 Pout, st);
  else
    LAUNCH_TIMED(ev, 0, (k_cgs<float, 0>), dim3(g.grid_cgs), dim3(NSX_BLOCK), lds, s, g, j, v, rhs, x, Pin, Gin,
                 Pout, st);
  return (int)hipGetLastError();
}

// the FP64 instance (4-layer blocks): the same schedule on its own grid
int launch_cgs64(const Geo &g, hipStream_t s, int j, const double *rhs, double *x, const CgsVecsT<double> &v,
                 Partials part, SolverState *st, hipEvent_t *ev) {
  double *Pout = (j & 1) ? part.u : part.c;
  const double *Pin = j == 0 ? part.u : ((j & 1) ? part.c : part.u);
  const uint32_t Gin = (uint32_t)(j == 0 ? g.grid : g.grid_cgs64);
  const size_t lds = cgs_lds(g, sizeof(double));
  LAUNCH_TIMED(ev, 0, (k_cgs<double, 1>), dim3(g.grid_cgs64), dim3(NSX_BLOCK), lds, s, g, j, v, rhs, x, Pin, Gin,
               Pout, st);
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
// resident blocks per CU of the FP32 CG stencil kernel
int cg_blocks_per_cu(const Geo &g) {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, (const void *)k_cg_w, NSX_BLOCK, level_lds(g, 4)) !=
          hipSuccess ||
      n < 1)
    return -1;
  return n;
}

// resident blocks per CU of the single-reduction FP32 CG kernel
int cgs_blocks_per_cu(const Geo &g) {
  const void *ks[3] = {(const void *)k_cgs<float, 0>, (const void *)k_cgs<float, 1>, (const void *)k_cgs<float, 2>};
  int m = 1 << 30;
  for (const void *k : ks) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, NSX_BLOCK, cgs_lds(g)) != hipSuccess || n < 1)
      return -1;
    m = n < m ? n : m;
  }
  return m;
}

// resident blocks per CU of the FP64 single-reduction CG kernel
int cgs64_blocks_per_cu(const Geo &g) {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, (const void *)k_cgs<double, 1>, NSX_BLOCK,
                                                   cgs_lds(g, sizeof(double))) != hipSuccess ||
      n < 1)
    return -1;
  return n;
}

int solver_blocks_per_cu(const Geo &g) {
  struct K {
    const void *f;
    size_t lds;
  } ks[] = {
      {(const void *)k_bicg_a<double>, level_lds(g, 8)}, {(const void *)k_bicg_b<double>, level_lds(g, 8)},
      {(const void *)k_bicg_c<double>, level_lds(g, 8)}, {(const void *)k_resid<double>, 0},
      {(const void *)k_resid<float>, 0},                  {(const void *)k_cg_start<float>, 0},
      {(const void *)k_cg_start<double>, 0},
      {(const void *)k_cg_x, level_lds(g, 4)},
  };
  int best = 1 << 30;
  for (const K &k : ks) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k.f, NSX_BLOCK, k.lds) != hipSuccess || n < 1) return -1;
    if (n < best) best = n;
  }
  return best;
}

// ------------------------------------- CG with the vertical-mode preconditioner
// (NSX_REFINE_MODAL; the tables and their meaning: nsx_modal.hip).  A
// residual r (FP32, column-contiguous) goes to modal coefficients column by
// column, rh_m = sum_k Phi[k][m] r(k+1); the lowest m0 go through one
// V-cycle of the 2D aggregation multigrid, the others are divided by their
// diagonal d_i + mu_m area_i; u = -Phi x (the operator is -(L (x) S + D (x)
// T)).  Both transforms are GEMMs of the N0 columns against the mode table:
// on the matrix cores (v_mfma_f32_16x16x4_f32, FP32 in and out) when the
// columns have at most 128 slots, else on the vector units.

// --- matrix-core transforms.  A wave forms 16 columns x all modes (or
// interior layers): its lane (j, g) = (lane & 15, lane >> 4) holds float4s of
// column c0 + j at interior layers (or modes) 16 t + 4 g .. +3, which are the
// A operand of four MFMAs (k index 4 g + q of step (t, q)); the table's float4
// for (t, tile, lane) holds the matching B operands, laid out by the host so
// that one ds_read_b128 per lane fetches them.  The result D[i][j] of a 16 x
// 16 tile sits in lane j + 16 (i / 4), register i % 4.  Layers and modes are
// both padded to mp = 16 MT (zero rows and columns of the table).
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline f32x4 mfma4(const float4 &a, const float4 &b, f32x4 c) {
  c = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, b.x, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, b.y, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x4f32(a.z, b.z, c, 0, 0, 0);
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a.w, b.w, c, 0, 0, 0);
}

// the transform kernels' blocks: 4 waves share one operand table in LDS
constexpr int MODAL_TB = 256;

// the operand table (64 MT^2 float4s) into LDS, every load issued before the
// first store
template <int MT>
__device__ inline void lds_table(float4 *dst, const float4 *src) {
  constexpr int n = 64 * MT * MT, U = (n + MODAL_TB - 1) / MODAL_TB;
  float4 v[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int q = threadIdx.x + u * MODAL_TB;
    if (q < n) v[u] = src[q];
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int q = threadIdx.x + u * MODAL_TB;
    if (q < n) dst[q] = v[u];
  }
  __syncthreads();
}

// LDS written by some lanes of a wave, then read by others: ordered at the
// wavefront's scope
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// the A operands of the forward transform: interior layers 16 t + 4 g .. +3 of
// column c (slots one above; four-byte aligned, so one unaligned dwordx4
// where all four are interior, single loads at the top, zero above nzi)
template <int MT>
__device__ inline void load_cols(const float *r, int NS, int N0, int nzi, int c, int gq, float4 (&a)[MT]) {
#pragma unroll
  for (int t = 0; t < MT; ++t) {
    const int k = 16 * t + 4 * gq;
    a[t] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (c >= N0) continue;
    const float *src = r + (size_t)c * NS + 1 + k;
    if (k + 4 <= nzi) {
      __builtin_memcpy(&a[t], src, sizeof(float4));
    } else {
      if (k < nzi) a[t].x = src[0];
      if (k + 1 < nzi) a[t].y = src[1];
      if (k + 2 < nzi) a[t].z = src[2];
    }
  }
}

template <int MT>
__device__ inline void load_modes(const float *xh, int N0, int c, int gq, float4 (&a)[MT]) {
#pragma unroll
  for (int t = 0; t < MT; ++t) {
    a[t] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (c < N0) a[t] = *(const float4 *)(xh + (size_t)c * (16 * MT) + 16 * t + 4 * gq);
  }
}

// the same with the 16 multigrid modes formed here: the V-cycle's last step,
// the level-0 post-smoothing x = y + omega (b - A y) / d with y = omega z +
// x_1(agg) at the column and its neighbours (the arithmetic of k_mg_up<1>),
// for modes 4 g .. +3; the other modes from xh
__device__ inline float4 ld4f(const float *p) { return *(const float4 *)p; }

template <int MT>
__device__ inline void load_modes_v0(const ModalDev &M, int N0, int c, int gq, float4 (&a)[MT]) {
  load_modes<MT>(M.xh, N0, c, gq, a);
  if (c >= N0) return;
  const ModalLevelDev &F = M.lev[0];
  const float om = NSX_MODAL_OMEGA;
  const float *z = M.z[0], *b = M.b[0], *x1 = M.x[1];
  const int m = 4 * gq;
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  for (int e = F.ptr[c]; e < F.ptr[c + 1]; ++e) {
    const float4 zj = ld4f(z + (size_t)F.col[e] * 16 + m), xj = ld4f(x1 + (size_t)F.eagg[e] * 16 + m);
    const float w = F.w[e];
    acc[0] = fmaf(w, fmaf(om, zj.x, xj.x), acc[0]);
    acc[1] = fmaf(w, fmaf(om, zj.y, xj.y), acc[1]);
    acc[2] = fmaf(w, fmaf(om, zj.z, xj.z), acc[2]);
    acc[3] = fmaf(w, fmaf(om, zj.w, xj.w), acc[3]);
  }
  const float4 zc = ld4f(z + (size_t)c * 16 + m), bc = ld4f(b + (size_t)c * 16 + m),
               xc = ld4f(x1 + (size_t)F.agg[c] * 16 + m);
  const float zv[4] = {zc.x, zc.y, zc.z, zc.w}, bv[4] = {bc.x, bc.y, bc.z, bc.w}, xv[4] = {xc.x, xc.y, xc.z, xc.w};
  const float dgc = F.diag[c], arc = F.area[c];
  float o[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float di = dgc + M.mu[m + k] * arc;
    const float yi = fmaf(om, zv[k], xv[k]);
    const float ay = fmaf(di, yi, -acc[k]);
    o[k] = yi + om * (bv[k] - ay) / di;
  }
  a[0] = make_float4(o[0], o[1], o[2], o[3]);
}

// MT = mp / 16.  Two blocks of 4 waves a CU; a wave with more than one tile
// loads its next tile's operands before it forms the current one.  STAGE:
// the wave's 16 rows of xh (one contiguous stretch) are assembled in LDS and
// leave as aligned float4s; the low modes' entries written there are
// overwritten by the V-cycle's last prolongation.  The diagonal scaling
// multiplies by v_rcp_f32 of d (a fixed diagonal: the preconditioner stays
// symmetric).
template <int MT, bool STAGE>
__global__ __launch_bounds__(MODAL_TB, 1) void k_modal_fwd_mfma(Geo g, ModalDev M, const float *r,
                                                                 const SolverState *st) {
  if (st->m_stop) return;
  extern __shared__ __align__(16) unsigned char nsx_dyn_lds[];
  float4 *tab = reinterpret_cast<float4 *>(nsx_dyn_lds);
  const int m0 = M.m0, nzi = M.nzi;
  constexpr int MP = 16 * MT;
  lds_table<MT>(tab, M.tfwd);
  const int lane = threadIdx.x & 63, j = lane & 15, gq = lane >> 4;
  float *stg = reinterpret_cast<float *>(tab + 64 * MT * MT) + (size_t)(threadIdx.x >> 6) * 16 * MP;
  const int ntile = (g.N0 + 15) >> 4, stride = gridDim.x * (MODAL_TB / 64);
  int tile = blockIdx.x * (MODAL_TB / 64) + (threadIdx.x >> 6);
  float4 a[MT], an[MT];
  load_cols<MT>(r, g.NS, g.N0, nzi, tile * 16 + j, gq, a);
  for (; tile < ntile; tile += stride) {
    if (tile + stride < ntile) load_cols<MT>(r, g.NS, g.N0, nzi, (tile + stride) * 16 + j, gq, an);
    const int c0 = tile * 16;
    f32x4 acc[MT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) acc[mt] = f32x4{0.f, 0.f, 0.f, 0.f};
    // one layer tile at a time: the scheduler keeps a step's table operands
    // (MT float4s) in flight, not the whole table
#pragma unroll
    for (int t = 0; t < MT; ++t) {
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) acc[mt] = mfma4(a[t], tab[(t * MT + mt) * 64 + lane], acc[mt]);
      __builtin_amdgcn_sched_barrier(0);
    }
    float dg[4], ar[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = min(c0 + 4 * gq + q, g.N0 - 1);
      dg[q] = M.lev[0].diag[c];
      ar[q] = M.lev[0].area[c];
    }
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      const int m = 16 * mt + j;
      const float mum = M.mu[m];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int c = c0 + 4 * gq + q;
        const float v = acc[mt][q], x = v * __builtin_amdgcn_rcpf(dg[q] + mum * ar[q]);
        if (STAGE) stg[(4 * gq + q) * MP + m] = x;
        if (c >= g.N0) continue;
        if (m < m0) {
          M.b[0][(size_t)c * m0 + m] = v;
          M.z[0][(size_t)c * m0 + m] = x;
        } else if (!STAGE) {
          M.xh[(size_t)c * MP + m] = x;
        }
      }
    }
    if (STAGE) {
      wave_lds_sync();
      const int nf = min(16, g.N0 - c0) * MP;
      float *xt = M.xh + (size_t)c0 * MP;
      for (int o = 4 * lane; o < nf; o += 256) *(float4 *)(xt + o) = *(const float4 *)(stg + o);
      wave_lds_sync();
    }
#pragma unroll
    for (int t = 0; t < MT; ++t) a[t] = an[t];
  }
}

// u = -Phi x from the modal coefficients xh (N0 x mp) at the interior layers,
// zero at the other slots; partial sums of <r, u>.  STAGE: the wave's 16
// columns (one contiguous stretch of u) are assembled in LDS and leave, with
// r read beside them for <r, u>, as aligned float4s (else each lane stores
// its outputs where they go, in 64-byte runs that straddle two segments).
template <int MT, bool STAGE, bool V0>
__global__ __launch_bounds__(MODAL_TB, 1) void k_modal_bwd_mfma(Geo g, ModalDev M, const float *r, float *u,
                                                                 double *P, const SolverState *st) {
  if (st->m_stop) return;
  extern __shared__ __align__(16) unsigned char nsx_dyn_lds[];
  float4 *tab = reinterpret_cast<float4 *>(nsx_dyn_lds);
  const int nzi = M.nzi, NS = g.NS;
  lds_table<MT>(tab, M.tbwd);
  const int lane = threadIdx.x & 63, j = lane & 15, gq = lane >> 4;
  float *stg = reinterpret_cast<float *>(tab + 64 * MT * MT) + (size_t)(threadIdx.x >> 6) * 16 * NS;
  const int ntile = (g.N0 + 15) >> 4, stride = gridDim.x * (MODAL_TB / 64);
  int tile = blockIdx.x * (MODAL_TB / 64) + (threadIdx.x >> 6);
  double acc2[1] = {0.0};
  float4 a[MT], an[MT];
  if (V0)
    load_modes_v0<MT>(M, g.N0, tile * 16 + j, gq, a);
  else
    load_modes<MT>(M.xh, g.N0, tile * 16 + j, gq, a);
  for (; tile < ntile; tile += stride) {
    if (tile + stride < ntile) {
      if (V0)
        load_modes_v0<MT>(M, g.N0, (tile + stride) * 16 + j, gq, an);
      else
        load_modes<MT>(M.xh, g.N0, (tile + stride) * 16 + j, gq, an);
    }
    const int c0 = tile * 16;
    f32x4 acc[MT];
#pragma unroll
    for (int nt = 0; nt < MT; ++nt) acc[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < MT; ++t) {
#pragma unroll
      for (int nt = 0; nt < MT; ++nt) acc[nt] = mfma4(a[t], tab[(t * MT + nt) * 64 + lane], acc[nt]);
      __builtin_amdgcn_sched_barrier(0);
    }
    const int nout = NS - nzi;  // slot 0 and nzi + 1 .. NS - 1
    if (STAGE) {
      for (int q = lane; q < 16 * nout; q += 64) {
        const int cl = q / nout, o = q - cl * nout;
        stg[cl * NS + (o == 0 ? 0 : nzi + o)] = 0.f;
      }
#pragma unroll
      for (int nt = 0; nt < MT; ++nt) {
        const int k = 16 * nt + j;
        if (k >= nzi) continue;
#pragma unroll
        for (int q = 0; q < 4; ++q) stg[(4 * gq + q) * NS + k + 1] = 0.f - acc[nt][q];
      }
      wave_lds_sync();
      const int ncol = min(16, g.N0 - c0), nf = ncol * NS;
      float *ut = u + (size_t)c0 * NS;
      const float *rt = r + (size_t)c0 * NS;
      for (int o = 4 * lane; o < nf; o += 256) {
        const float4 uv = *(const float4 *)(stg + o);
        const float4 rv = *(const float4 *)(rt + o);
        *(float4 *)(ut + o) = uv;
        const int s0 = o % NS;  // NS is a multiple of 4: the float4 stays in one column
        const float uu[4] = {uv.x, uv.y, uv.z, uv.w}, rr[4] = {rv.x, rv.y, rv.z, rv.w};
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (s0 + q >= 1 && s0 + q <= nzi) acc2[0] += (double)rr[q] * (double)uu[q];
      }
      wave_lds_sync();
    } else {
#pragma unroll
      for (int nt = 0; nt < MT; ++nt) {
        const int k = 16 * nt + j;
        if (k >= nzi) continue;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int c = c0 + 4 * gq + q;
          if (c >= g.N0) continue;
          const size_t sl = (size_t)c * NS + k + 1;
          const float uv = 0.f - acc[nt][q];
          u[sl] = uv;
          acc2[0] += (double)r[sl] * (double)uv;
        }
      }
      for (int q = lane; q < 16 * nout; q += 64) {
        const int c = c0 + q / nout, o = q % nout;
        if (c < g.N0) u[(size_t)c * NS + (o == 0 ? 0 : nzi + o)] = 0.f;
      }
    }
#pragma unroll
    for (int t = 0; t < MT; ++t) a[t] = an[t];
  }
  write_partials<1, MODAL_TB>(acc2, P);
}

// --- vector-unit transforms for columns of more than 128 interior layers
constexpr int MODAL_COLS = 32;  // columns a block transforms together
constexpr int MODAL_CSTR = 36;  // LDS row stride of the column tile (float4-aligned, spreads the banks)

__device__ inline bool modal_lds_fits(const ModalDev &M) {
  return (size_t)M.nzp * M.nzp + (size_t)MODAL_CSTR * M.nzp <= 16384;
}
static size_t modal_lds(const ModalDev &M) {
  const size_t t = (size_t)M.nzp * M.nzp + (size_t)MODAL_CSTR * M.nzp;
  return (t <= 16384 ? t : (size_t)MODAL_CSTR * M.nzp) * sizeof(float);
}

// out[c][j] = sum_k tab[k][j] in[k][c] for the block's MODAL_COLS columns: a
// thread forms a 4 x 4 tile (4 columns, 4 of j) from a float4 of the table
// and a float4 of the column tile per k (tab: nzp x nzp row-major, in LDS or
// global; tile: [k][MODAL_CSTR] in LDS); f(c, j, value) stores
template <class F>
__device__ inline void modal_tiles(const float *tab, const float *tile, int nzp, int nk, F &&f) {
  const int nq = nzp / 4, nt = (MODAL_COLS / 4) * nq;
  for (int q = threadIdx.x; q < nt; q += NSX_BLOCK) {
    const int cq = q / nq, jq = q - cq * nq;
    float a[4][4] = {};
    for (int k = 0; k < nk; ++k) {
      const float4 t = *(const float4 *)(tab + (size_t)k * nzp + 4 * jq);
      const float4 x = *(const float4 *)(tile + (size_t)k * MODAL_CSTR + 4 * cq);
      const float tv[4] = {t.x, t.y, t.z, t.w}, xv[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
      for (int cc = 0; cc < 4; ++cc)
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) a[cc][jj] = fmaf(tv[jj], xv[cc], a[cc][jj]);
    }
#pragma unroll
    for (int cc = 0; cc < 4; ++cc)
#pragma unroll
      for (int jj = 0; jj < 4; ++jj) f(4 * cq + cc, 4 * jq + jj, a[cc][jj]);
  }
}

__global__ __launch_bounds__(NSX_BLOCK) void k_modal_fwd_vec(Geo g, ModalDev M, const float *r,
                                                              const SolverState *st) {
  if (st->m_stop) return;
  extern __shared__ __align__(16) unsigned char nsx_dyn_lds[];
  float *lds = reinterpret_cast<float *>(nsx_dyn_lds);
  const int nzp = M.nzp, nzi = M.nzi, m0 = M.m0, MP = M.mp;
  const bool in_lds = modal_lds_fits(M);
  const float *ph = in_lds ? lds : M.phi;
  float *rc = in_lds ? lds + (size_t)nzp * nzp : lds;
  if (in_lds)
    for (int q = threadIdx.x; q < nzp * nzp; q += NSX_BLOCK) lds[q] = M.phi[q];
  const int ngrp = (g.N0 + MODAL_COLS - 1) / MODAL_COLS;
  for (int grp = blockIdx.x; grp < ngrp; grp += gridDim.x) {
    __syncthreads();
    const int c0 = grp * MODAL_COLS;
    for (int q = threadIdx.x; q < MODAL_COLS * nzp; q += NSX_BLOCK) {
      const int c = q / nzp, k = q - c * nzp, i = c0 + c;
      rc[(size_t)k * MODAL_CSTR + c] = (i < g.N0 && k < nzi) ? r[(size_t)i * g.NS + k + 1] : 0.f;
    }
    __syncthreads();
    modal_tiles(ph, rc, nzp, nzi, [&](int c, int m, float v) {
      const int i = c0 + c;
      if (i >= g.N0 || m >= nzi) return;
      const float d = M.lev[0].diag[i] + M.mu[m] * M.lev[0].area[i];
      if (m < m0) {
        M.b[0][(size_t)i * m0 + m] = v;
        M.z[0][(size_t)i * m0 + m] = v / d;
      } else {
        M.xh[(size_t)i * MP + m] = v / d;
      }
    });
  }
}

__global__ __launch_bounds__(NSX_BLOCK) void k_modal_bwd_vec(Geo g, ModalDev M, const float *r, float *u, double *P,
                                                              const SolverState *st) {
  if (st->m_stop) return;
  extern __shared__ __align__(16) unsigned char nsx_dyn_lds[];
  float *lds = reinterpret_cast<float *>(nsx_dyn_lds);
  const int nzp = M.nzp, nzi = M.nzi, MP = M.mp;
  const bool in_lds = modal_lds_fits(M);
  const float *pt = in_lds ? lds : M.phiT;
  float *xc = in_lds ? lds + (size_t)nzp * nzp : lds;
  if (in_lds)
    for (int q = threadIdx.x; q < nzp * nzp; q += NSX_BLOCK) lds[q] = M.phiT[q];
  double acc[1] = {0.0};
  const int ngrp = (g.N0 + MODAL_COLS - 1) / MODAL_COLS;
  for (int grp = blockIdx.x; grp < ngrp; grp += gridDim.x) {
    __syncthreads();
    const int c0 = grp * MODAL_COLS;
    for (int q = threadIdx.x; q < MODAL_COLS * nzp; q += NSX_BLOCK) {
      const int c = q / nzp, m = q - c * nzp, i = c0 + c;
      xc[(size_t)m * MODAL_CSTR + c] = (i < g.N0 && m < nzi) ? M.xh[(size_t)i * MP + m] : 0.f;
    }
    __syncthreads();
    modal_tiles(pt, xc, nzp, nzi, [&](int c, int k, float v) {
      const int i = c0 + c;
      if (i >= g.N0 || k >= nzi) return;
      const size_t s = (size_t)i * g.NS + k + 1;
      u[s] = 0.f - v;
      acc[0] += (double)r[s] * (double)(0.f - v);
    });
  }
  // the slots outside the interior layers
  const uint32_t nout = (uint32_t)(g.NS - nzi);
  for (uint32_t t = blockIdx.x * NSX_BLOCK + threadIdx.x; t < (uint32_t)g.N0 * nout; t += gridDim.x * NSX_BLOCK) {
    const uint32_t i = t / nout, j = t - i * nout;
    const int k = j == 0 ? 0 : nzi + (int)j;
    u[(size_t)i * g.NS + k] = 0.f;
  }
  write_partials<1>(acc, P);
}

// --- the batched 2D V-cycle on n_l x m0 arrays (node-major, mode fastest):
// the operator of mode m at level l is d_i(m) - W, d_i(m) = diag_i + mu_m
// area_i.  One damped Jacobi sweep from zero before the coarse correction
// (x = omega z, z = b / d, written by the producer of b) and one after it; the
// restriction forms the residual of the first where it needs it, and the
// post-smoothing forms the prolonged y = omega z + P x_c at the node and its
// neighbours: one launch per level each way.  Rows with many entries are
// split over S slices of lanes (lane m + 16 s), summed by shuffles in a fixed
// order.
constexpr float MG_OMEGA = NSX_MODAL_OMEGA;

template <int S>
__device__ inline float slice_sum(float v) {
  if (S >= 2) v += __shfl_xor(v, 16, 64);
  if (S >= 4) v += __shfl_xor(v, 32, 64);
  return v;
}

// b_c[J] = sum over the children i of J of (b - A x)_i with x = omega z:
// (1 - omega) sum_i b_i + omega sum_j RW[J][j] z_j (RW = R W, the weights of
// the children summed per fine neighbour), one CSR over the fine nodes with
// a coefficient on b and one on z; z_c = b_c / d_J.  A wave per J.
__global__ __launch_bounds__(NSX_BLOCK) void k_mg_down(ModalLevelDev C, const float *mu, int m0, const float *bf,
                                                        const float *zf, float *bc, float *zc, const SolverState *st) {
  if (st->m_stop) return;
  const uint32_t t = blockIdx.x * NSX_BLOCK + threadIdx.x;
  const int J = (int)(t >> 6), m = (int)(t & 15), sl = (int)((t >> 4) & 3);
  if (J >= C.n) return;  // whole waves
  float v = 0.f;
  if (m < m0) {
    for (int e = C.dptr[J] + sl; e < C.dptr[J + 1]; e += 4) {
      const size_t q = (size_t)C.dcol[e] * m0 + m;
      v = fmaf(C.db[e], bf[q], fmaf(C.dz[e], zf[q], v));
    }
  }
  v = slice_sum<4>(v);
  if (sl == 0 && m < m0) {
    bc[(size_t)J * m0 + m] = v;
    zc[(size_t)J * m0 + m] = v / (C.diag[J] + mu[m] * C.area[J]);
  }
}

// x = C^-1 b on the coarsest level, C^-1 laid out [J][K][m]: a block per J,
// 16 slices of K
__global__ __launch_bounds__(NSX_BLOCK) void k_mg_coarse(int nc, int m0, const float *cinvT, const float *b, float *x,
                                                          int ostride, const SolverState *st) {
  if (st->m_stop) return;
  __shared__ float sh[NSX_BLOCK / 64][16];
  const int J = blockIdx.x, m = threadIdx.x & 15, sl = threadIdx.x >> 4;
  float v = 0.f;
  if (m < m0) {
    const float *row = cinvT + (size_t)J * nc * m0 + m;
    for (int K = sl; K < nc; K += NSX_BLOCK / 16) v = fmaf(row[(size_t)K * m0], b[(size_t)K * m0 + m], v);
  }
  v = slice_sum<4>(v);
  if ((threadIdx.x & 63) < 16) sh[threadIdx.x >> 6][m] = v;
  __syncthreads();
  if (threadIdx.x < 16 && m < m0) {
    float s = 0.f;
    for (int w = 0; w < NSX_BLOCK / 64; ++w) s += sh[w][m];
    x[(size_t)J * ostride + m] = s;
  }
}

// x = y + omega (b - A y) / d with y = omega z + x_c(agg) at the node and its
// neighbours; S slices per row (S = 4: a wave per row)
template <int S>
__global__ __launch_bounds__(NSX_BLOCK) void k_mg_up(ModalLevelDev F, const float *mu, int m0, const float *xc,
                                                      const float *bf, const float *zf, float *xf, int ostride,
                                                      const SolverState *st) {
  if (st->m_stop) return;
  const uint32_t t = blockIdx.x * NSX_BLOCK + threadIdx.x;
  const int m = (int)(t & 15), sl = (int)((t >> 4) & (S - 1));
  const uint32_t i = t / (16 * S);
  if (S > 1 && i >= (uint32_t)F.n) return;  // whole waves
  const bool act = i < (uint32_t)F.n && m < m0;
  float acc = 0.f;
  if (act)
    for (int e = F.ptr[i] + sl; e < F.ptr[i + 1]; e += S)
      acc = fmaf(F.w[e], fmaf(MG_OMEGA, zf[(size_t)F.col[e] * m0 + m], xc[(size_t)F.eagg[e] * m0 + m]), acc);
  acc = slice_sum<S>(acc);
  if (act && sl == 0) {
    const size_t q = (size_t)i * m0 + m;
    const float di = F.diag[i] + mu[m] * F.area[i];
    const float yi = fmaf(MG_OMEGA, zf[q], xc[(size_t)F.agg[i] * m0 + m]);
    const float ay = fmaf(di, yi, -acc);
    xf[(size_t)i * ostride + m] = yi + MG_OMEGA * (bf[q] - ay) / di;
  }
}

// (ev: null, or a start/stop pair the dispatch records)
// the backward transform forms the V-cycle's level-0 post-smoothing where it
// runs on the matrix cores with 16 multigrid modes and the V-cycle has a
// level below the columns
static bool modal_fused_v0(const ModalDev &M) { return M.mt <= 8 && M.m0 == 16 && M.nlev >= 2; }

// staged kernels past 64 KB of LDS: the kernel's limit raised once (false
// when the device does not allow it)
static bool lds_allow(const void *kern, size_t bytes, int optin) {
  if (bytes <= 65536) return true;
  if (bytes > (size_t)optin) return false;
  return hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) == hipSuccess;
}

template <int MT>
static void launch_transforms_t(const Geo &g, hipStream_t s, const ModalDev &M, const float *r, float *u, double *P,
                                const SolverState *st, bool fwd, hipEvent_t *ev) {
  const size_t lds = (size_t)MT * MT * 64 * sizeof(float4);
  if (fwd) {
    // the staged forward transform where table and staging fit the LDS a
    // block may have
    const size_t fstage = (size_t)(MODAL_TB / 64) * 16 * (16 * MT) * sizeof(float);
    static const bool fok =
        lds_allow(reinterpret_cast<const void *>(&k_modal_fwd_mfma<MT, true>), lds + fstage, M.lds_optin);
    if (fok)
      LAUNCH_TIMED(ev, 0, (k_modal_fwd_mfma<MT, true>), dim3(M.agrid), dim3(MODAL_TB), lds + fstage, s, g, M, r, st);
    else
      LAUNCH_TIMED(ev, 0, (k_modal_fwd_mfma<MT, false>), dim3(M.agrid), dim3(MODAL_TB), lds, s, g, M, r, st);
    return;
  }
  // the staged backward transform where table and staging fit the LDS a
  // block may have; it forms the level-0 post-smoothing itself where the
  // V-cycle left it out (modal_fused_v0)
  const size_t stage = (size_t)(MODAL_TB / 64) * 16 * g.NS * sizeof(float);
  const bool v0 = modal_fused_v0(M);
  static size_t allowed = 65536;  // the staged kernels' LDS limit as set so far
  if (lds + stage > allowed &&
      lds_allow(reinterpret_cast<const void *>(&k_modal_bwd_mfma<MT, true, false>), lds + stage, M.lds_optin) &&
      lds_allow(reinterpret_cast<const void *>(&k_modal_bwd_mfma<MT, true, true>), lds + stage, M.lds_optin))
    allowed = lds + stage;
  if (lds + stage <= allowed) {
    if (v0)
      LAUNCH_TIMED(ev, 0, (k_modal_bwd_mfma<MT, true, true>), dim3(M.agrid), dim3(MODAL_TB), lds + stage, s, g, M, r,
                   u, P, st);
    else
      LAUNCH_TIMED(ev, 0, (k_modal_bwd_mfma<MT, true, false>), dim3(M.agrid), dim3(MODAL_TB), lds + stage, s, g, M, r,
                   u, P, st);
  } else {
    if (v0)
      LAUNCH_TIMED(ev, 0, (k_modal_bwd_mfma<MT, false, true>), dim3(M.agrid), dim3(MODAL_TB), lds, s, g, M, r, u, P,
                   st);
    else
      LAUNCH_TIMED(ev, 0, (k_modal_bwd_mfma<MT, false, false>), dim3(M.agrid), dim3(MODAL_TB), lds, s, g, M, r, u, P,
                   st);
  }
}

// the matrix-core kernels take at most 128 interior layers (mp / 16 <= 8);
// above, the vector units
static void launch_transform(const Geo &g, hipStream_t s, const ModalDev &M, const float *r, float *u, double *P,
                             const SolverState *st, bool fwd, hipEvent_t *ev) {
  switch (M.mt) {
    case 1: return launch_transforms_t<1>(g, s, M, r, u, P, st, fwd, ev);
    case 2: return launch_transforms_t<2>(g, s, M, r, u, P, st, fwd, ev);
    case 3: return launch_transforms_t<3>(g, s, M, r, u, P, st, fwd, ev);
    case 4: return launch_transforms_t<4>(g, s, M, r, u, P, st, fwd, ev);
    case 5: return launch_transforms_t<5>(g, s, M, r, u, P, st, fwd, ev);
    case 6: return launch_transforms_t<6>(g, s, M, r, u, P, st, fwd, ev);
    case 7: return launch_transforms_t<7>(g, s, M, r, u, P, st, fwd, ev);
    case 8: return launch_transforms_t<8>(g, s, M, r, u, P, st, fwd, ev);
    default: break;
  }
  const size_t lds = modal_lds(M);
  if (fwd)
    LAUNCH_TIMED(ev, 0, k_modal_fwd_vec, dim3(M.agrid), dim3(NSX_BLOCK), lds, s, g, M, r, st);
  else
    LAUNCH_TIMED(ev, 0, k_modal_bwd_vec, dim3(M.agrid), dim3(NSX_BLOCK), lds, s, g, M, r, u, P, st);
}

int launch_modal_apply(const Geo &g, hipStream_t s, const ModalDev &M, const float *r, float *u, Partials part,
                       const SolverState *st, hipEvent_t *ev) {
  const dim3 blk(NSX_BLOCK);
  launch_transform(g, s, M, r, u, part.a, st, true, ev);
  if (ev) (void)hipEventRecord(ev[2], s);
  const int m0 = M.m0, L = M.nlev;
  auto grid = [](size_t threads) { return dim3((unsigned)((threads + NSX_BLOCK - 1) / NSX_BLOCK)); };
  for (int l = 0; l + 1 < L; ++l)
    hipLaunchKernelGGL(k_mg_down, grid((size_t)M.lev[l + 1].n * 64), blk, 0, s, M.lev[l + 1], M.mu, m0, M.b[l],
                       M.z[l], M.b[l + 1], M.z[l + 1], st);
  // the coarsest solve and the last prolongation write the low modes into xh
  hipLaunchKernelGGL(k_mg_coarse, dim3(M.nc), blk, 0, s, M.nc, m0, M.cinv, M.b[L - 1], L == 1 ? M.xh : M.x[L - 1],
                     L == 1 ? M.mp : m0, st);
  for (int l = L - 2; l >= (modal_fused_v0(M) ? 1 : 0); --l) {
    float *out = l == 0 ? M.xh : M.x[l];
    const int os = l == 0 ? M.mp : m0;
    if (l == 0)
      hipLaunchKernelGGL(k_mg_up<1>, grid((size_t)M.lev[l].n * 16), blk, 0, s, M.lev[l], M.mu, m0, M.x[l + 1],
                         M.b[l], M.z[l], out, os, st);
    else
      hipLaunchKernelGGL(k_mg_up<4>, grid((size_t)M.lev[l].n * 64), blk, 0, s, M.lev[l], M.mu, m0, M.x[l + 1],
                         M.b[l], M.z[l], out, os, st);
  }
  if (ev) (void)hipEventRecord(ev[3], s);
  launch_transform(g, s, M, r, u, part.a, st, false, ev ? ev + 4 : nullptr);
  return (int)hipGetLastError();
}

// p_out = u + beta p_in at the seven points, w = A p_out, <p, w>
__global__ __launch_bounds__(NSX_BLOCK) void k_modal_spmv(Geo g, const float *u, const float *pin, float *pout,
                                                           float *w, Partials P, SolverState *st, int first,
                                                           uint32_t na) {
  if (st->done || st->m_stop) return;
  double t[1];
  reduce_partials<1>(P.a, na, t);
  const double rho = t[0];
  // every block forms the same rho, so all return together
  if (!first && fabs(rho) < NSX_MODAL_DROP * fabs(st->m_rho0)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) st->m_stop = 1;
    return;
  }
  const double beta = first ? 0.0 : rho / st->m_rho_prev;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->m_rho = rho;
    if (first) st->m_rho0 = rho;
  }
  float *lvw = load_levels<float>(g);
  for (int k = threadIdx.x; k < g.NS; k += NSX_BLOCK) lvw[2 * g.NS + k] = 1.f / lvw[k];
  __syncthreads();
  const float *lv = lvw;
  double acc[1] = {0.0};
  for_each_quad_gathered(
      g, u, pin, (float)beta, first != 0, [&](size_t) {},
      [&](int, int k0, size_t s0, const Col<float> &C, const float(&fc)[4], float fm, float fp,
          const float(&f0)[4], const float(&f1)[4], const float(&f2)[4]) {
        float qv[5], wo[4], po[4];
        quot5_r(C, lv, g.NS, k0, qv);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const Row1<float> R = row_of(g, lv, C, qv, k0, v);
          const float wv = apply_row(g, R, k0, v, fc, fm, fp, f0[v], f1[v], f2[v]);
          wo[v] = R.in ? wv : 0.f;
          po[v] = R.in ? fc[v] : 0.f;
          if (R.in) acc[0] += (double)fc[v] * (double)wv;
        }
        st4(pout + s0, po);
        st4(w + s0, wo);
      });
  write_partials<1>(acc, P.b);
}

__global__ __launch_bounds__(NSX_BLOCK) void k_modal_update(Geo g, const float *p, const float *w, float *e, float *r,
                                                             Partials P, SolverState *st) {
  if (st->done || st->m_stop) return;
  double t[1];
  reduce_partials<1>(P.b, gridDim.x, t);
  const float alpha = (float)(st->m_rho / t[0]);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->m_rho_prev = st->m_rho;
    st->m_iters += 1;
  }
  for_each_quad(g, [&](int, int, size_t s0) {
    float pp[4], ww[4], ee[4], rr[4];
    ld4(p + s0, pp);
    ld4(w + s0, ww);
    ld4(e + s0, ee);
    ld4(r + s0, rr);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      ee[v] = fmaf(alpha, pp[v], ee[v]);
      rr[v] = fmaf(-alpha, ww[v], rr[v]);
    }
    st4(e + s0, ee);
    st4(r + s0, rr);
  });
}

__global__ __launch_bounds__(NSX_BLOCK) void k_modal_true(Geo g, const double *rhs, const double *x, const float *e,
                                                           float *r, Partials P, SolverState *st) {
  if (st->done) return;
  const double avec = st->avec;
  double acc[3] = {0.0, 0.0, 0.0};
  FOR_EACH_CELL(g, c) {
    int i, k;
    cell_of(g, c, i, k);
    double res, xs;
    true_residual(g, rhs, x, avec, i, k, [&](size_t q, bool) { return e[q]; }, res, xs);
    r[(size_t)i * g.NS + k] = (float)res;
    acc[0] += res * res;
    acc[1] += fabs(res);
    acc[2] += xs * (g.area[i] * g.dsig[k]);
  }
  write_partials<3>(acc, P.c);
}

__global__ __launch_bounds__(NSX_BLOCK) void k_modal_fold(Geo g, double *x, float *e, Partials P, SolverState *st,
                                                           int cap) {
  if (st->done) return;
  const int iters = st->m_iters;
  double t[3];
  reduce_partials<3>(P.c, gridDim.x, t);
  const double rN = sqrt(t[0]), e1 = t[1] / st->b1, e2 = rN / st->b2;
  const bool finite = isfinite(t[0]) && isfinite(t[1]);
  int stop = 0, code = EXIT_CONVERGED;
  if (!finite) {
    stop = 1;
    code = EXIT_BREAKDOWN;
  } else if (e1 < st->eps || e2 < st->eps) {
    stop = 1;
  } else if (iters >= cap) {
    stop = 1;
    code = EXIT_MAX;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->e1 = e1;
    st->e2 = e2;
    st->xvol = t[2];
    st->a_iter = iters;
    st->mp_pass += 1;
    st->m_stop = 0;
    if (stop) {
      st->done = 1;
      st->exit_code = code;
    }
  }
  // x += e, e = 0 (x keeps its value when the residual was not finite)
  for_each_quad(g, [&](int, int k0, size_t s0) {
    float ee[4];
    ld4(e + s0, ee);
    if (finite)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        if (k0 + v >= 1 && k0 + v <= g.NZ - 2) x[s0 + v] = x[s0 + v] + (double)ee[v];
    const float z[4] = {0.f, 0.f, 0.f, 0.f};
    st4(e + s0, z);
  });
}

int launch_modal_spmv(const Geo &g, hipStream_t s, const ModalDev &M, const float *u, const float *pin, float *pout,
                      float *w, Partials part, SolverState *st, int first, hipEvent_t *ev) {
  LAUNCH_TIMED(ev, 0, k_modal_spmv, dim3(g.grid), dim3(NSX_BLOCK), cgs_lds(g), s, g, u, pin, pout, w, part, st,
               first, (uint32_t)M.agrid);
  return (int)hipGetLastError();
}

int launch_modal_update(const Geo &g, hipStream_t s, const float *p, const float *w, float *e, float *r,
                        Partials part, SolverState *st, hipEvent_t *ev) {
  LAUNCH_TIMED(ev, 0, k_modal_update, dim3(g.grid), dim3(NSX_BLOCK), 0, s, g, p, w, e, r, part, st);
  return (int)hipGetLastError();
}

int launch_modal_true(const Geo &g, hipStream_t s, const double *rhs, double *x, float *e, float *r, Partials part,
                      SolverState *st, int cap) {
  hipLaunchKernelGGL(k_modal_true, dim3(g.grid), dim3(NSX_BLOCK), 0, s, g, rhs, (const double *)x, (const float *)e,
                     r, part, st);
  hipLaunchKernelGGL(k_modal_fold, dim3(g.grid), dim3(NSX_BLOCK), 0, s, g, x, e, part, st, cap);
  return (int)hipGetLastError();
}
