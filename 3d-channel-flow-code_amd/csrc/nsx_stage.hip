// nsx_stage.hip -- the non-solver kernels of one projection stage (hydro.F90).
//
// Each kernel states the reference routine(s) it replaces.  Arithmetic keeps
// the reference's operation order and is compiled without FMA contraction, so
// every kernel here except the Smagorinsky cube root reproduces the CPU
// results bit for bit; the cube root uses the device pow.
//
// HBM roofline: all of these are streaming stencils at < 1 flop/byte, no
// MFMA; per-kernel algorithmic bytes are listed next to each kernel.
#include "nsx_kernels.h"

int nsx_grid_for(uint32_t n) {
  uint32_t blocks = (n + NSX_BLOCK - 1) / NSX_BLOCK;
  uint32_t g = blocks < 2048 ? blocks : 2048;
  g = (g + NSX_NXCD - 1) / NSX_NXCD * NSX_NXCD;
  return g ? (int)g : NSX_NXCD;
}

#define FOR_EACH_CELL(g, c)                                                         \
  uint32_t beg_, end_;                                                              \
  block_range((g).ncell, beg_, end_);                                               \
  for (uint32_t base_ = beg_; base_ < end_; base_ += NSX_BLOCK)                     \
    for (uint32_t c = base_ + threadIdx.x, once_ = 1; once_ && c < end_; once_ = 0)

// The element-wise stencil kernels sweep each XCD's eighth of the cells band
// by band: the XCD's G/8 blocks take consecutive batches of NSX_BLOCK cells
// in turn, so the columns a batch gathers from (its horizontal neighbours,
// mostly in the same or the next batches) are fetched into that XCD's L2
// while the blocks that own them load them too.  With one contiguous chunk
// per block, an XCD's blocks were spread over its whole slice and a
// neighbour column was often evicted between the two loads.  A kernel that
// accumulates partial sums (k_source) adds them in the order of its sweep,
// which the fixed grid makes the same on every run.
__device__ inline void band_range(uint32_t n, uint32_t &beg, uint32_t &end, uint32_t &step) {
  const uint32_t G = gridDim.x, b = blockIdx.x;
  if ((G % NSX_NXCD) == 0) {
    uint32_t per = (n + NSX_NXCD - 1) / NSX_NXCD;
    per = (per + NSX_BLOCK - 1) / NSX_BLOCK * NSX_BLOCK;
    const uint64_t s0 = (uint64_t)(b % NSX_NXCD) * per, s1 = s0 + per;
    beg = (uint32_t)(s0 + (uint64_t)(b / NSX_NXCD) * NSX_BLOCK < n ? s0 + (uint64_t)(b / NSX_NXCD) * NSX_BLOCK : n);
    end = (uint32_t)(s1 < n ? s1 : n);
    step = G / NSX_NXCD * NSX_BLOCK;
  } else {
    beg = b * NSX_BLOCK < n ? b * NSX_BLOCK : n;
    end = n;
    step = G * NSX_BLOCK;
  }
}

#define FOR_EACH_CELL_BAND(g, c)                                                    \
  uint32_t beg_, end_, step_;                                                       \
  band_range((g).ncell, beg_, end_, step_);                                         \
  for (uint32_t base_ = beg_; base_ < end_; base_ += step_)                         \
    for (uint32_t c = base_ + threadIdx.x, once_ = 1; once_ && c < end_; once_ = 0)

// Runs of layers: the stencil kernels below take the layers k0 .. k0+RN-1
// (k0 a multiple of RN) of one column per thread.  The column data
// (neighbours, vertices, edge lengths) is loaded once for the run, each run
// of a cell or face field is read with 16-byte accesses, and a wave's
// accesses to a field span 64 RN consecutive slots.  Every prism's
// arithmetic is the one-prism form's, in its order.  Slots outside the
// interior layers 1..NZ-2 (the ghost layers and the padding, in a column's
// first and last runs) are computed on whatever they hold and never stored.
// RL is the run the kernels use.  k_les keeps one prism per thread: with
// its FP64 cube root per prism, runs of one or two layers took 205 us at
// 6.3M prisms against 175 us.
#ifndef NSX_RUN
#define NSX_RUN 2
#endif
constexpr int RL = NSX_RUN;
static_assert(RL == 1 || RL == 2 || RL == 4, "a run divides the 4-slot column padding");

template <int RN>
struct Run {
  int i, k0;
  size_t s0;
  bool act[RN], full;
};

// runs in the cell fields: N0 columns of NS / RN runs
template <int RN>
__host__ __device__ inline uint32_t nruns(const Geo &g) {
  return (uint32_t)g.N0 * (uint32_t)(g.NS / RN);
}

template <int RN>
__device__ inline Run<RN> run_at(const Geo &g, uint32_t t) {
  Run<RN> R;
  // t / (NS/RN) = (t / (4/RN)) / NQ, NQ = NS/4
  const uint32_t i = fdiv(t / (uint32_t)(4 / RN), g.divNQ);
  R.i = (int)i;
  R.k0 = RN * (int)(t - i * (uint32_t)(g.NS / RN));
  R.s0 = (size_t)i * g.NS + R.k0;
  R.full = true;
#pragma unroll
  for (int u = 0; u < RN; ++u) {
    const int k = R.k0 + u;
    R.act[u] = k >= 1 && k <= g.NZ - 2;
    R.full = R.full && R.act[u];
  }
  return R;
}

template <int RN>
__device__ inline void ldr(const double *p, double (&o)[RN]) {
  if constexpr (RN == 1) {
    o[0] = *p;
  } else {
#pragma unroll
    for (int h = 0; h < RN / 2; ++h) {
      const double2 a = ((const double2 *)p)[h];
      o[2 * h] = a.x;
      o[2 * h + 1] = a.y;
    }
  }
}

// the run's active slots of o into p: 16-byte stores when every slot is active
template <int RN>
__device__ inline void str(double *p, const double (&o)[RN], const Run<RN> &R) {
  if (RN > 1 && R.full) {
#pragma unroll
    for (int h = 0; h < RN / 2; ++h) ((double2 *)p)[h] = make_double2(o[2 * h], o[2 * h + 1]);
  } else {
#pragma unroll
    for (int u = 0; u < RN; ++u)
      if (R.act[u]) p[u] = o[u];
  }
}

#define FOR_EACH_RUN_BAND(g, t, RN)                                                 \
  uint32_t beg_, end_, step_;                                                       \
  band_range(nruns<RN>(g), beg_, end_, step_);                                      \
  for (uint32_t base_ = beg_; base_ < end_; base_ += step_)                         \
    for (uint32_t t = base_ + threadIdx.x, once_ = 1; once_ && t < end_; once_ = 0)

// A cell field around a run: c[u+1] = f at layer k0+u of the column, c[0]
// and c[RN+1] the layers below and above (loaded only where an active layer
// needs them), n[j][u] = f at layer k0+u of neighbour column j, v[j][u] =
// the vertex field at level k0+u of the triangle's j-th vertex (active
// layers only: the levels run 0..NZ-2; none when fv is null).
template <int RN>
struct FieldRun {
  double c[RN + 2], n[3][RN], v[3][RN];
};

template <int RN>
__device__ inline void field_run(const Geo &g, const Run<RN> &R, const int (&nb)[3], const int (&vx)[3],
                                 const double *f, const double *fv, FieldRun<RN> &F) {
  double m[RN];
  ldr(f + R.s0, m);
#pragma unroll
  for (int u = 0; u < RN; ++u) F.c[u + 1] = m[u];
  F.c[0] = R.act[0] ? f[R.s0 - 1] : 0.0;
  F.c[RN + 1] = R.act[RN - 1] ? f[R.s0 + RN] : 0.0;
#pragma unroll
  for (int j = 0; j < 3; ++j) ldr(f + (size_t)nb[j] * g.NS + R.k0, F.n[j]);
  if (!fv) return;
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int u = 0; u < RN; ++u) F.v[j][u] = R.act[u] ? fv[(size_t)vx[j] * g.NL + R.k0 + u] : 0.0;
}

// the column data of a run's edges that the kernels below use
struct EdgeRun {
  int nb[3], vx[3];
};

__device__ inline EdgeRun edges_of(const Geo &g, int i) {
  EdgeRun E;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    E.nb[j] = g.nbr[i * 3 + j];
    E.vx[j] = g.vtx[i * 3 + j];
  }
  return E;
}

// faces_of (below) at layer k0+u from a FieldRun: the same expressions
template <int RN>
__device__ inline void faces_run(const FieldRun<RN> &F, int u, const double (&doe)[3], const double (&lvv)[3],
                                 double (&out)[5]) {
  const double fi = F.c[u + 1];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const double extra = doe[j] * (F.v[(j + 1) % 3][u] - F.v[j][u]) / lvv[j];
    out[j] = F.n[j][u] * 0.5 + fi * 0.5 + extra;
  }
  out[3] = 0.5 * (fi + F.c[u + 2]);
  out[4] = 0.5 * (fi + F.c[u]);
}

// ------------------------------------------------------- cell -> vertex ----
// interpolation2D / interpolation3D(P) + BCvelvertex3D / BCpvertex +
// bc_periodic (interpolation2D.F90:8-150, interpolation3D.F90:8-444,
// BC.F90:603-1408).  One thread per (vertex, level); the periodic average of
// a boundary vertex recomputes its partners' values in-thread.
// Bytes: NF fields x (read ~1 cell field + write 0.5 vertex field).
struct VtxArgs {
  const double *f[3];
  double *fv[3];
  const signed char *act[3];
  int zero_bot[3], zero_top[3];
  int lid_top[3];  // the lid: 1 on the top level of the interior vertices (BC.F90:771-777)
  double inflow[3];  // the value of an inflow vertex (VA_INFLOW, BC.F90:724-731)
  int tagdir;
};

// The inverse-distance sums of vertex v at the layers l, l + 1 and l + 2
// (s[0..NL_-1], NL_ of them), each added in the stencil's order.  The stencil
// entries are taken four at a time so that their loads are in flight
// together; a cell's layers l, l + 1 (16-byte aligned: l is even) are one
// load.  Layers past `nl` are not read (their sums stay 0).
template <int NL_>
__device__ inline void idw_sums(const Geo &g, int v, int l, int nl, const double *f, double (&s)[NL_]) {
  const int b = g.vptr[v], e = g.vptr[v + 1];
#pragma unroll
  for (int q = 0; q < NL_; ++q) s[q] = 0.0;
  auto add = [&](int j0, int n) {
    double a[4][NL_], w[4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (q < n) {
        const double *c = f + (size_t)g.vcell[j0 + q] * g.NS + l;
        w[q] = g.vw[j0 + q];
        const double2 d = *(const double2 *)c;
        a[q][0] = d.x;
        if (NL_ > 1) a[q][NL_ > 1 ? 1 : 0] = nl > 1 ? d.y : 0.0;
        if (NL_ > 2) a[q][NL_ > 2 ? 2 : 0] = nl > 2 ? c[2] : 0.0;
      }
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (q < n)
#pragma unroll
        for (int m = 0; m < NL_; ++m) s[m] = s[m] + w[q] * a[q][m];
  };
  int j = b;
  for (; j + 4 <= e; j += 4) add(j, 4);
  if (j < e) add(j, e - j);
}

// Threads take runs of two vertex levels l0 = 2m, l0 + 1 of one vertex (the
// second only when l0 + 1 < NL): the stencil is read once for both, and
// each cell's layers come as one 16-byte load.  Blocks are dealt round-robin
// to the 8 XCDs; the remap gives each XCD one contiguous range of vertices,
// so the columns that neighbouring vertices share are fetched into one L2
// (grid a multiple of 8).
template <int NF>
__global__ __launch_bounds__(NSX_BLOCK) void k_vertex(Geo g, VtxArgs a) {
  const uint32_t G = gridDim.x, b = blockIdx.x;
  const uint32_t lb = (G % NSX_NXCD) == 0 ? (b % NSX_NXCD) * (G / NSX_NXCD) + b / NSX_NXCD : b;
  const uint32_t t = lb * NSX_BLOCK + threadIdx.x;
  const uint32_t nrv = (uint32_t)(g.NL + 1) / 2;
  if (t >= (uint32_t)g.NV * nrv) return;
  const int v = (int)(t / nrv);
  const int l0 = 2 * (int)(t - (uint32_t)v * nrv);
  const int nl = l0 + 1 < g.NL ? 2 : 1;  // levels in this run
  double dzT[2] = {0.0, 0.0}, dzB[2] = {0.0, 0.0};
  if (a.tagdir == 3)
#pragma unroll
    for (int m = 0; m < 2; ++m)
      if (m < nl) {
        dzT[m] = fabs(g.sigv[l0 + m] - g.sig[l0 + m + 1]);
        dzB[m] = fabs(g.sigv[l0 + m] - g.sig[l0 + m]);
      }
#pragma unroll
  for (int q = 0; q < NF; ++q) {
    const double *f = a.f[q];
    const signed char *act = a.act[q];
    bool zero_wall[2];
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      const int l = l0 + m;
      zero_wall[m] = (a.tagdir == 3) && ((l == 0 && a.zero_bot[q]) || (l == g.NL - 1 && a.zero_top[q]));
    }
    // the values of vertex vv at the run's levels (interpolation3D: the two
    // layers around the level weighted by their sigma distances)
    auto val = [&](int vv, double (&r)[2]) {
      if (a.tagdir == 3) {
        double sm[3];
        idw_sums<3>(g, vv, l0, nl + 1, f, sm);
        const double ds = g.dlVsum[vv];
#pragma unroll
        for (int m = 0; m < 2; ++m) r[m] = (dzB[m] * (sm[m + 1] / ds) + dzT[m] * (sm[m] / ds)) / (dzT[m] + dzB[m]);
      } else {
        double sm[2];
        idw_sums<2>(g, vv, l0, nl, f, sm);
#pragma unroll
        for (int m = 0; m < 2; ++m) r[m] = sm[m] / g.dlVsum[vv];
      }
#pragma unroll
      for (int m = 0; m < 2; ++m) {
        if (act[vv] == VA_ZERO) r[m] = 0.;
        if (act[vv] == VA_INFLOW) r[m] = a.inflow[q];
        if (zero_wall[m]) r[m] = 0.;
      }
    };
    double mine[2];
    val(v, mine);
    if (act[v] == VA_PERIODIC) {
      double o[2] = {0.0, 0.0}, r[2];
      const signed char kind = g.vkind[v];
      if (kind == VK_CORNER) {
        double sd = 0.0;
        for (int q4 = 0; q4 < g.ncorner; ++q4) {
          const int c = g.corner[q4];
          val(c, r);
#pragma unroll
          for (int m = 0; m < 2; ++m) o[m] = o[m] + r[m] * g.dlVsum[c];
          sd = sd + g.dlVsum[c];
        }
#pragma unroll
        for (int m = 0; m < 2; ++m) o[m] = o[m] / sd;
      } else if (kind == VK_PAIR) {
        const int p = g.vpair[v];
        const double sd = g.dlVsum[v] + g.dlVsum[p];
        val(p, r);
#pragma unroll
        for (int m = 0; m < 2; ++m) o[m] = (g.dlVsum[v] * mine[m] + g.dlVsum[p] * r[m]) / sd;
      }
#pragma unroll
      for (int m = 0; m < 2; ++m) mine[m] = o[m];
    }
#pragma unroll
    for (int m = 0; m < 2; ++m)
      if (m < nl) {
        const int l = l0 + m;
        double out = mine[m];
        if (a.tagdir == 3 && a.lid_top[q] && l == g.NL - 1 && g.vkind[v] == VK_INTERIOR) out = 1.0;
        a.fv[q][(size_t)v * g.NL + l] = out;
      }
  }
}

int launch_vertex(const Geo &g, hipStream_t s, int nf, const double *const f[3], double *const fv[3],
                  const signed char *const act[3], const int zero_bot[3], const int zero_top[3], int tagdir,
                  const int *lid_top, const double *inflow) {
  VtxArgs a;
  for (int q = 0; q < 3; ++q) {
    a.f[q] = q < nf ? f[q] : nullptr;
    a.fv[q] = q < nf ? fv[q] : nullptr;
    a.act[q] = q < nf ? act[q] : nullptr;
    a.zero_bot[q] = q < nf ? zero_bot[q] : 0;
    a.zero_top[q] = q < nf ? zero_top[q] : 0;
    a.lid_top[q] = q < nf && lid_top ? lid_top[q] : 0;
    a.inflow[q] = q < nf && inflow ? inflow[q] : 0.0;
  }
  a.tagdir = tagdir;
  const uint32_t n = (uint32_t)g.NV * (uint32_t)((g.NL + 1) / 2);
  const dim3 grid(((n + NSX_BLOCK - 1) / NSX_BLOCK + NSX_NXCD - 1) / NSX_NXCD * NSX_NXCD);
  if (nf == 3) hipLaunchKernelGGL(k_vertex<3>, grid, dim3(NSX_BLOCK), 0, s, g, a);
  else hipLaunchKernelGGL(k_vertex<1>, grid, dim3(NSX_BLOCK), 0, s, g, a);
  return (int)hipGetLastError();
}

// ------------------------------------------------------------ face values --
// calcul_face, Kim-Choi branch (interpolationFace.F90:85-111):
// phi_f = 0.5 phi_nb + 0.5 phi_i + doe (phiv(nv2) - phiv(nv1)) / dlVV
__device__ inline void faces_of(const Geo &g, int i, int k, size_t s, const double *f, const double *fv,
                                double (&out)[5]) {
  const double fi = f[s];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int e = i * 3 + j;
    const int nb = g.nbr[e];
    const int nv1 = g.vtx[e], nv2 = g.vtx[i * 3 + (j + 1) % 3];
    const double extra = g.doe[e] * (fv[(size_t)nv2 * g.NL + k] - fv[(size_t)nv1 * g.NL + k]) / g.dlVV[e];
    out[j] = f[(size_t)nb * g.NS + k] * 0.5 + fi * 0.5 + extra;
  }
  out[3] = 0.5 * (fi + f[s + 1]);
  out[4] = 0.5 * (fi + f[s - 1]);
}

__global__ __launch_bounds__(NSX_BLOCK) void k_calcul_face(Geo g, const double *f, const double *fv, double *o0,
                                                            double *o1, double *o2, double *o3, double *o4) {
  FOR_EACH_CELL_BAND(g, c) {
    int i, k;
    cell_of(g, c, i, k);
    const size_t s = (size_t)i * g.NS + k;
    double fc[5];
    faces_of(g, i, k, s, f, fv, fc);
    o0[s] = fc[0];
    o1[s] = fc[1];
    o2[s] = fc[2];
    o3[s] = fc[3];
    o4[s] = fc[4];
  }
}

int launch_calcul_face(const Geo &g, hipStream_t s, const double *f, const double *fv, double *const out[5]) {
  hipLaunchKernelGGL(k_calcul_face, dim3(nsx_grid_for(g.ncell)), dim3(NSX_BLOCK), 0, s, g, f, fv, out[0], out[1],
                     out[2], out[3], out[4]);
  return (int)hipGetLastError();
}

// ------------------------------------------------------------ Smagorinsky --
// sgles + fvm_gradient + sgdamp (sgm_les.F90:24-502, 1247-1281), channel
// (utau = 1, walld = |sig(k)|), Mason power 2 evaluated as products and a
// square root, as the compiled reference does.  dvdx and dvdy take dlVV of
// edge 3 for edge 2 and of edge 2 for edge 3 (sgm_les.F90:444-446,453-455).
// The length scale, (Cs Delta)^2 with Mason's damping (Delta = (area
// dsig)^(1/3), lwall = karman walld, 1/l^2 = 1/(Cs Delta)^2 + 1/lwall^2, the
// device pow for the cube root), depends on the geometry only: k_les_length
// forms it once at commit with the same expressions, and k_les reads it.
// Bytes/cell of k_les: read u,v,w (+3 vertex fields at 0.5) and the length
// scale = 5.5, write eddy = 1.
__global__ __launch_bounds__(NSX_BLOCK) void k_les_length(Geo g, double *nutg) {
  const double Cs = 0.1, karman = (double)0.415f;
  FOR_EACH_CELL_BAND(g, c) {
    int i, k;
    cell_of(g, c, i, k);
    const double a = g.area[i];
    const double dz = g.dsig[k];
    double DD = Cs * pow(a * dz * 1.0, 1.0 / 3.0);
    const double lwall = karman * g.walld[k];
    const double dn1 = (1.0 / DD) * (1.0 / DD) + (1.0 / lwall) * (1.0 / lwall);
    const double dn2 = 1.0 / dn1;
    const double del = sqrt(dn2);
    DD = del * del;
    nutg[(size_t)i * g.NS + k] = 1.0 * DD;
  }
}

int launch_les_length(const Geo &g, hipStream_t s, double *nutg) {
  hipLaunchKernelGGL(k_les_length, dim3(nsx_grid_for(g.ncell)), dim3(NSX_BLOCK), 0, s, g, nutg);
  return (int)hipGetLastError();
}

__global__ __launch_bounds__(NSX_BLOCK) void k_les(Geo g, const double *u, const double *v, const double *w,
                                                    const double *uv, const double *vv, const double *wv,
                                                    double *eddy) {
  FOR_EACH_CELL_BAND(g, c) {
    int i, k;
    cell_of(g, c, i, k);
    const size_t s = (size_t)i * g.NS + k;
    double U[5], V[5], W[5];
    faces_of(g, i, k, s, u, uv, U);
    faces_of(g, i, k, s, v, vv, V);
    faces_of(g, i, k, s, w, wv, W);
    const int e = i * 3;
    const double x1 = g.xnn[e], x2 = g.xnn[e + 1], x3 = g.xnn[e + 2];
    const double y1 = g.ynn[e], y2 = g.ynn[e + 1], y3 = g.ynn[e + 2];
    const double l1 = g.dlVV[e], l2 = g.dlVV[e + 1], l3 = g.dlVV[e + 2];
    const double a = g.area[i];
    double dudx = U[0] * x1 * l1 + U[1] * x2 * l2 + U[2] * x3 * l3;
    double dvdx = V[0] * x1 * l1 + V[1] * x2 * l3 + V[2] * x3 * l2;
    double dwdx = W[0] * x1 * l1 + W[1] * x2 * l2 + W[2] * x3 * l3;
    double dudy = U[0] * y1 * l1 + U[1] * y2 * l2 + U[2] * y3 * l3;
    double dvdy = V[0] * y1 * l1 + V[1] * y2 * l3 + V[2] * y3 * l2;
    double dwdy = W[0] * y1 * l1 + W[1] * y2 * l2 + W[2] * y3 * l3;
    dudx = dudx / a;
    dvdx = dvdx / a;
    dwdx = dwdx / a;
    dudy = dudy / a;
    dvdy = dvdy / a;
    dwdy = dwdy / a;
    const double dz = g.dsig[k];
    const double dudz = (U[3] - U[4]) / dz, dvdz = (V[3] - V[4]) / dz, dwdz = (W[3] - W[4]) / dz;
    const double nut = g.nutg[s];
    const double sxx = (dudx + dudx) / 2.0, sxy = (dudy + dvdx) / 2.0, sxz = (dudz + dwdx) / 2.0;
    const double syy = (dvdy + dvdy) / 2.0, syz = (dvdz + dwdy) / 2.0, szz = (dwdz + dwdz) / 2.0;
    const double Sij = sqrt(2.0 * (sxx * sxx + syy * syy + szz * szz) + 4.0 * (sxy * sxy + sxz * sxz + syz * syz));
    const double ed = nut * Sij;
    eddy[s] = ed;
    // BCpcenter(eddy): Neumann ghost layers
    if (k == 1 && g.bcbot == 1) eddy[s - 1] = ed;
    if (k == g.NZ - 2 && g.bctop == 1) eddy[s + 1] = ed;
  }
}

int launch_les(const Geo &g, hipStream_t s, const double *u, const double *v, const double *w, const double *uv,
               const double *vv, const double *wv, double *eddy) {
  hipLaunchKernelGGL(k_les, dim3(nsx_grid_for(g.ncell)), dim3(NSX_BLOCK), 0, s, g, u, v, w, uv, vv, wv, eddy);
  return (int)hipGetLastError();
}

// --------------------------------------------- momentum fluxes + predictor --
// AdvDiffNS = diffusion3DNEW (Kim-Choi, diffusion3D.F90:110-191) then
// advection3DNEW (advection3D.F90:112-284), fused with the predictor of
// hydro.F90:322-333 (u* = u + dt(c1 rhs + c2 rhs_old), rhs_old = rhs) and
// with the ghost layers of BCvelcenter3D for u* (hydro.F90:342-345).
// Bytes/cell (fused): read u,v,w, 3 vertex fields (1.5), 5 faces, eddy,
// 3 history = 12.5; write u*,v*,w*, 3 history = 6.
struct MomArgs {
  const double *uvw[3], *vtx[3], *face[5], *eddy, *rhs_in[3];
  const double *gam[3];  // per-component diffusivity fields in place of eddy + nu, or null
  const double *force;  // per-prism streamwise body force in place of bdx, or null
  double *hist[3], *rhs_out[3], *ustar[3];
  double c1, c2;
};

__global__ __launch_bounds__(NSX_BLOCK) void k_momentum(Geo g, MomArgs a) {
  FOR_EACH_RUN_BAND(g, t, RL) {
    const Run<RL> R = run_at<RL>(g, t);
    const int e = R.i * 3;
    int nb[3], vx[3];
    double doe[3], lvv[3], lcc[3], le1[3], le2[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      nb[j] = g.nbr[e + j];
      vx[j] = g.vtx[e + j];
      doe[j] = g.doe[e + j];
      lvv[j] = g.dlVV[e + j];
      lcc[j] = g.dlCC[e + j];
      le1[j] = g.dle1[e + j];
      le2[j] = g.dle2[e + j];
    }
    const double area = g.area[R.i];
    // per layer: dsigv(k-1), dsig(k), dsig(k-1) (1 in the slots never stored)
    double sv[RL], dsk[RL], dsb[RL];
#pragma unroll
    for (int m = 0; m < RL; ++m) {
      const int k = R.k0 + m;
      sv[m] = R.act[m] ? g.dsigv[k - 1] : 1.0;
      dsk[m] = R.act[m] ? g.dsig[k] : 1.0;
      dsb[m] = R.act[m] ? g.dsig[k - 1] : 1.0;
    }
    double ed[RL], fr[5][RL];
    ldr(a.eddy + R.s0, ed);
#pragma unroll
    for (int f = 0; f < 5; ++f) ldr(a.face[f] + R.s0, fr[f]);
    double Vol[RL], Gam[RL], UF[5][RL];
#pragma unroll
    for (int m = 0; m < RL; ++m) {
      Vol[m] = area * sv[m];
      Gam[m] = ed[m] + g.nu;
      UF[0][m] = fr[0][m] * lvv[0] * sv[m];
      UF[1][m] = fr[1][m] * lvv[1] * sv[m];
      UF[2][m] = fr[2][m] * lvv[2] * sv[m];
      UF[3][m] = fr[3][m] * area;
      UF[4][m] = fr[4][m] * area;
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      FieldRun<RL> F;
      field_run(g, R, nb, vx, a.uvw[q], a.vtx[q], F);
      double rin[RL], gm[RL], hs[RL];
      if (a.rhs_in[q]) {
        ldr(a.rhs_in[q] + R.s0, rin);
      } else if (q == 0 && a.force) {
        ldr(a.force + R.s0, rin);
      } else {
#pragma unroll
        for (int m = 0; m < RL; ++m) rin[m] = q == 0 ? g.bdx : (q == 1 ? g.bdy : g.bdz);
      }
      if (a.gam[q]) {
        ldr(a.gam[q] + R.s0, gm);
      } else {
#pragma unroll
        for (int m = 0; m < RL; ++m) gm[m] = Gam[m];
      }
      if (a.ustar[q]) ldr(a.hist[q] + R.s0, hs);
      double rr[RL], us[RL];
#pragma unroll
      for (int m = 0; m < RL; ++m) {
        const double pi = F.c[m + 1];
        double dn[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          const double diff1 = (F.n[j][m] - pi) / lcc[j];
          const double diff2 = (F.v[(j + 1) % 3][m] - F.v[j][m]) / lvv[j];
          dn[j] = diff1 * le1[j] + diff2 * le2[j];
        }
        const double dT = (F.c[m + 2] - pi) / dsk[m];
        const double dB = (F.c[m] - pi) / dsb[m];
        const double src = dn[0] * lvv[0] * sv[m] + dn[1] * lvv[1] * sv[m] + dn[2] * lvv[2] * sv[m] + (dT + dB) * area;
        double r = rin[m];
        r = r + gm[m] * src / Vol[m];
        double fc[5];
        faces_run(F, m, doe, lvv, fc);
        const double flux = fc[0] * UF[0][m] + fc[1] * UF[1][m] + fc[2] * UF[2][m] + fc[3] * UF[3][m] + fc[4] * UF[4][m];
        r = r - flux / Vol[m];
        rr[m] = r;
        if (a.ustar[q]) us[m] = pi + g.dt * (a.c1 * r + a.c2 * hs[m]);
      }
      if (a.rhs_out[q]) str(a.rhs_out[q] + R.s0, rr, R);
      if (a.ustar[q]) {
        str(a.hist[q] + R.s0, rr, R);
        str(a.ustar[q] + R.s0, us, R);
#pragma unroll
        for (int m = 0; m < RL; ++m) {
          const int k = R.k0 + m;
          double gv;
          if (k == 1 && vel_ghost(g.bcbot, g.BotBC, q, us[m], gv)) a.ustar[q][R.s0 + m - 1] = gv;
          if (k == g.NZ - 2 && vel_ghost(g.bctop, g.TopBC, q, us[m], gv)) a.ustar[q][R.s0 + m + 1] = gv;
        }
      }
    }
  }
}

int launch_momentum(const Geo &g, hipStream_t s, const double *const uvw[3], const double *const vtx[3],
                    const double *const face[5], const double *eddy, const double *const rhs_in[3],
                    double *const hist[3], double *const rhs_out[3], double *const ustar[3], double c1, double c2,
                    const double *force, const double *const gam[3]) {
  MomArgs a;
  a.force = force;
  for (int q = 0; q < 3; ++q) {
    a.gam[q] = gam ? gam[q] : nullptr;
    a.uvw[q] = uvw[q];
    a.vtx[q] = vtx[q];
    a.rhs_in[q] = rhs_in ? rhs_in[q] : nullptr;
    a.hist[q] = hist ? hist[q] : nullptr;
    a.rhs_out[q] = rhs_out ? rhs_out[q] : nullptr;
    a.ustar[q] = ustar ? ustar[q] : nullptr;
  }
  for (int q = 0; q < 5; ++q) a.face[q] = face[q];
  a.eddy = eddy;
  a.c1 = c1;
  a.c2 = c2;
  hipLaunchKernelGGL(k_momentum, dim3(nsx_grid_for(nruns<RL>(g))), dim3(NSX_BLOCK), 0, s, g, a);
  return (int)hipGetLastError();
}

// ------------------------------------------ FP32 flux and eddy viscosity --
// The precision form's FP32 arithmetic for the momentum fluxes and the
// Smagorinsky eddy viscosity (NSX_PREC_MIXED_FLUX32): the stencil
// differences are formed in FP64 from the stored FP64 fields and rounded to
// FP32 once (a difference of two nearby values keeps its relative accuracy
// that way, where the FP32 values themselves would lose it to cancellation);
// every product, quotient and sum of the fluxes and of the strain rate is
// FP32; the predictor u* = u + dt r adds the FP32 r to the FP64 u.  In exact
// arithmetic these are k_momentum and k_les: the advective flux is formed as
// sum_j (phi_f - phi_i) UF_j + phi_i sum_j UF_j and the face gradients from
// phi_f - phi_i (plus the constant part that fvm_gradient's swapped edge
// lengths leave in dv/dx, dv/dy).  The
// fields stay FP64 in memory, so the bytes are k_momentum's and k_les's.

// the five face values of f at (i,k) minus f(i,k), in FP32 from FP64 differences
__device__ inline void faces_dev32(const Geo &g, int i, int k, size_t s, const double *f, const double *fv,
                                   float (&out)[5]) {
  const double fi = f[s];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int e = i * 3 + j;
    const int nb = g.nbr[e];
    const int nv1 = g.vtx[e], nv2 = g.vtx[i * 3 + (j + 1) % 3];
    const float dv = (float)(fv[(size_t)nv2 * g.NL + k] - fv[(size_t)nv1 * g.NL + k]);
    const float dc = (float)(f[(size_t)nb * g.NS + k] - fi);
    out[j] = dc * 0.5f + (float)g.doe[e] * dv / (float)g.dlVV[e];
  }
  out[3] = 0.5f * (float)(f[s + 1] - fi);
  out[4] = 0.5f * (float)(f[s - 1] - fi);
}

__global__ __launch_bounds__(NSX_BLOCK) void k_les32(Geo g, const double *u, const double *v, const double *w,
                                                      const double *uv, const double *vv, const double *wv,
                                                      double *eddy) {
  const float Cs = 0.1f, karman = 0.415f;
  FOR_EACH_CELL_BAND(g, c) {
    int i, k;
    cell_of(g, c, i, k);
    const size_t s = (size_t)i * g.NS + k;
    float U[5], V[5], W[5];
    faces_dev32(g, i, k, s, u, uv, U);
    faces_dev32(g, i, k, s, v, vv, V);
    faces_dev32(g, i, k, s, w, wv, W);
    const int e = i * 3;
    const float x1 = (float)g.xnn[e], x2 = (float)g.xnn[e + 1], x3 = (float)g.xnn[e + 2];
    const float y1 = (float)g.ynn[e], y2 = (float)g.ynn[e + 1], y3 = (float)g.ynn[e + 2];
    const float l1 = (float)g.dlVV[e], l2 = (float)g.dlVV[e + 1], l3 = (float)g.dlVV[e + 2];
    const float a = (float)g.area[i];
    // the face values are deviations from the prism's own value, whose
    // contribution sum_j n_j l_j phi_i vanishes on the closed triangle -- but
    // not for dvdx, dvdy, whose edge lengths 2 and 3 are swapped: their
    // constant part v_i sum_j n_j l_j' is kept, as the reference has it
    const float vi = (float)v[s];
    const float dudx = (U[0] * x1 * l1 + U[1] * x2 * l2 + U[2] * x3 * l3) / a;
    const float dvdx = (V[0] * x1 * l1 + V[1] * x2 * l3 + V[2] * x3 * l2) / a + vi * ((x1 * l1 + x2 * l3 + x3 * l2) / a);
    const float dwdx = (W[0] * x1 * l1 + W[1] * x2 * l2 + W[2] * x3 * l3) / a;
    const float dudy = (U[0] * y1 * l1 + U[1] * y2 * l2 + U[2] * y3 * l3) / a;
    const float dvdy = (V[0] * y1 * l1 + V[1] * y2 * l3 + V[2] * y3 * l2) / a + vi * ((y1 * l1 + y2 * l3 + y3 * l2) / a);
    const float dwdy = (W[0] * y1 * l1 + W[1] * y2 * l2 + W[2] * y3 * l3) / a;
    const float dz = (float)g.dsig[k];
    const float dudz = (U[3] - U[4]) / dz, dvdz = (V[3] - V[4]) / dz, dwdz = (W[3] - W[4]) / dz;
    const float DD = Cs * cbrtf(a * dz);
    const float lwall = karman * (float)g.walld[k];
    const float dn1 = (1.0f / DD) * (1.0f / DD) + (1.0f / lwall) * (1.0f / lwall);
    const float nut = 1.0f / dn1;
    const float sxx = dudx, sxy = (dudy + dvdx) * 0.5f, sxz = (dudz + dwdx) * 0.5f;
    const float syy = dvdy, syz = (dvdz + dwdy) * 0.5f, szz = dwdz;
    const float Sij = sqrtf(2.0f * (sxx * sxx + syy * syy + szz * szz) + 4.0f * (sxy * sxy + sxz * sxz + syz * syz));
    const double ed = (double)(nut * Sij);
    eddy[s] = ed;
    if (k == 1 && g.bcbot == 1) eddy[s - 1] = ed;
    if (k == g.NZ - 2 && g.bctop == 1) eddy[s + 1] = ed;
  }
}

int launch_les32(const Geo &g, hipStream_t s, const double *u, const double *v, const double *w, const double *uv,
                 const double *vv, const double *wv, double *eddy) {
  hipLaunchKernelGGL(k_les32, dim3(nsx_grid_for(g.ncell)), dim3(NSX_BLOCK), 0, s, g, u, v, w, uv, vv, wv, eddy);
  return (int)hipGetLastError();
}

__global__ __launch_bounds__(NSX_BLOCK) void k_momentum32(Geo g, MomArgs a) {
  FOR_EACH_CELL_BAND(g, c) {
    int i, k;
    cell_of(g, c, i, k);
    const size_t s = (size_t)i * g.NS + k;
    const int e = i * 3;
    const float area = (float)g.area[i];
    const float sv = (float)g.dsigv[k - 1];
    const float Vol = area * sv;
    const float Gam = (float)(a.eddy[s] + g.nu);
    float UF[5];
#pragma unroll
    for (int j = 0; j < 3; ++j) UF[j] = (float)a.face[j][s] * (float)g.dlVV[e + j] * sv;
    UF[3] = (float)a.face[3][s] * area;
    UF[4] = (float)a.face[4][s] * area;
    const float divf = UF[0] + UF[1] + UF[2] + UF[3] + UF[4];
    const float bd[3] = {(float)g.bdx, (float)g.bdy, (float)g.bdz};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double *phi = a.uvw[q];
      const double *phiv = a.vtx[q];
      const double pi = phi[s];
      float dn[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const int ej = e + j;
        const int nb = g.nbr[ej];
        const int jv1 = g.vtx[ej], jv2 = g.vtx[e + (j + 1) % 3];
        const float diff1 = (float)(phi[(size_t)nb * g.NS + k] - pi) / (float)g.dlCC[ej];
        const float diff2 = (float)(phiv[(size_t)jv2 * g.NL + k] - phiv[(size_t)jv1 * g.NL + k]) / (float)g.dlVV[ej];
        dn[j] = diff1 * (float)g.dle1[ej] + diff2 * (float)g.dle2[ej];
      }
      const float dT = (float)(phi[s + 1] - pi) / (float)g.dsig[k];
      const float dB = (float)(phi[s - 1] - pi) / (float)g.dsig[k - 1];
      const float src = dn[0] * (float)g.dlVV[e] * sv + dn[1] * (float)g.dlVV[e + 1] * sv +
                        dn[2] * (float)g.dlVV[e + 2] * sv + (dT + dB) * area;
      float fc[5];
      faces_dev32(g, i, k, s, phi, phiv, fc);
      const float flux = fc[0] * UF[0] + fc[1] * UF[1] + fc[2] * UF[2] + fc[3] * UF[3] + fc[4] * UF[4] +
                         (float)pi * divf;
      const float r0 = a.rhs_in[q] ? (float)a.rhs_in[q][s] : ((q == 0 && a.force) ? (float)a.force[s] : bd[q]);
      const float r = r0 + Gam * src / Vol - flux / Vol;
      if (a.rhs_out[q]) a.rhs_out[q][s] = (double)r;
      if (a.ustar[q]) {
        const double us = pi + g.dt * (a.c1 * (double)r + a.c2 * a.hist[q][s]);
        a.hist[q][s] = (double)r;
        a.ustar[q][s] = us;
        double gv;
        if (k == 1 && vel_ghost(g.bcbot, g.BotBC, q, us, gv)) a.ustar[q][s - 1] = gv;
        if (k == g.NZ - 2 && vel_ghost(g.bctop, g.TopBC, q, us, gv)) a.ustar[q][s + 1] = gv;
      }
    }
  }
}

int launch_momentum32(const Geo &g, hipStream_t s, const double *const uvw[3], const double *const vtx[3],
                      const double *const face[5], const double *eddy, double *const hist[3],
                      double *const ustar[3], double c1, double c2, const double *force) {
  MomArgs a;
  a.force = force;
  for (int q = 0; q < 3; ++q) {
    a.gam[q] = nullptr;
    a.uvw[q] = uvw[q];
    a.vtx[q] = vtx[q];
    a.rhs_in[q] = nullptr;
    a.hist[q] = hist[q];
    a.rhs_out[q] = nullptr;
    a.ustar[q] = ustar[q];
  }
  for (int q = 0; q < 5; ++q) a.face[q] = face[q];
  a.eddy = eddy;
  a.c1 = c1;
  a.c2 = c2;
  hipLaunchKernelGGL(k_momentum32, dim3(nsx_grid_for(g.ncell)), dim3(NSX_BLOCK), 0, s, g, a);
  return (int)hipGetLastError();
}

// --------------------------------------------------------- PPE right side --
// Source3D / interpolateU (source3D.F90:92-127, interpolationU.F90:122-160):
// face-normal velocities of u* and their divergence; fused (rhsp_out != 0)
// with rhsp/dt (hydro.F90:365-373), update_rhsp's lagged Kim-Choi flux
// (source3D.F90:255-288) and the partial sums of Glob_Norm(rhs,1)
// (bicgstab.F90:635-653).
// Bytes/cell (fused): read u*,v*,w*, p vertex (0.5) = 3.5; write 5 faces, rhsp = 6.
__global__ __launch_bounds__(NSX_BLOCK) void k_source(Geo g, const double *u, const double *v, const double *w,
                                                       double *F1, double *F2, double *F3, double *FT, double *FB,
                                                       double *src_out, double *rhsp_out, const double *pv,
                                                       double *partials) {
  double acc[1] = {0.0};
  FOR_EACH_RUN_BAND(g, t, RL) {
    const Run<RL> R = run_at<RL>(g, t);
    const int e = R.i * 3;
    const EdgeRun E = edges_of(g, R.i);
    double xn[3], yn[3], lvv[3], le2[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      xn[j] = g.xnn[e + j];
      yn[j] = g.ynn[e + j];
      lvv[j] = g.dlVV[e + j];
      le2[j] = g.dle2[e + j];
    }
    const double area = g.area[R.i];
    double ur[RL], vr[RL], wc[RL + 2], un[3][RL], vn[3][RL];
    ldr(u + R.s0, ur);
    ldr(v + R.s0, vr);
    {
      double m[RL];
      ldr(w + R.s0, m);
#pragma unroll
      for (int q = 0; q < RL; ++q) wc[q + 1] = m[q];
      wc[0] = R.act[0] ? w[R.s0 - 1] : 0.0;
      wc[RL + 1] = R.act[RL - 1] ? w[R.s0 + RL] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      ldr(u + (size_t)E.nb[j] * g.NS + R.k0, un[j]);
      ldr(v + (size_t)E.nb[j] * g.NS + R.k0, vn[j]);
    }
    double pvr[3][RL];
    if (rhsp_out)
#pragma unroll
      for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int q = 0; q < RL; ++q) pvr[j][q] = R.act[q] ? pv[(size_t)E.vx[j] * g.NL + R.k0 + q] : 0.0;
    double o1[RL], o2[RL], o3[RL], oT[RL], oB[RL], rs[RL], rp_[RL];
#pragma unroll
    for (int m = 0; m < RL; ++m) {
      const int k = R.k0 + m;
      const double ui = ur[m], vi = vr[m];
      double Uf[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double Ui = 0.5 * (ui + un[j][m]);
        const double Vi = 0.5 * (vi + vn[j][m]);
        Uf[j] = Ui * xn[j] + Vi * yn[j];
      }
      const double UT = 0.5 * (wc[m + 1] + wc[m + 2]);
      const double UB = -1.0 * (0.5 * (wc[m + 1] + wc[m]));
      o1[m] = Uf[0];
      o2[m] = Uf[1];
      o3[m] = Uf[2];
      oT[m] = UT;
      oB[m] = UB;
      const double sv = R.act[m] ? g.dsigv[k - 1] : 1.0;
      double resid = 0.0;
#pragma unroll
      for (int j = 0; j < 3; ++j) resid = resid + Uf[j] * lvv[j] * sv;
      resid = resid + UT * area;
      resid = resid + UB * area;
      rs[m] = resid;
      if (rhsp_out) {
        double rp = resid / g.dt;
        double dn[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          const double diff2 = (pvr[(j + 1) % 3][m] - pvr[j][m]) / lvv[j];
          dn[j] = diff2 * le2[j];
        }
        const double corr = dn[0] * lvv[0] * sv + dn[1] * lvv[1] * sv + dn[2] * lvv[2] * sv;
        rp = rp - corr;
        rp_[m] = rp;
        if (R.act[m]) acc[0] += rp;
      }
    }
    str(F1 + R.s0, o1, R);
    str(F2 + R.s0, o2, R);
    str(F3 + R.s0, o3, R);
    str(FT + R.s0, oT, R);
    str(FB + R.s0, oB, R);
    if (src_out) str(src_out + R.s0, rs, R);
    if (rhsp_out) str(rhsp_out + R.s0, rp_, R);
  }
  if (!rhsp_out) return;
  write_partials<1>(acc, partials);
}

int launch_source(const Geo &g, hipStream_t s, const double *const uvw[3], double *const face[5], double *src_out,
                  double *rhsp_out, const double *pv, Partials part) {
  hipLaunchKernelGGL(k_source, dim3(g.grid), dim3(NSX_BLOCK), 0, s, g, uvw[0], uvw[1], uvw[2],
                     face[0], face[1], face[2], face[3], face[4], src_out, rhsp_out, pv, part.src);
  return (int)hipGetLastError();
}

// update_rhsp alone (source3D.F90:255-288) for single-routine use
__global__ __launch_bounds__(NSX_BLOCK) void k_update_rhsp(Geo g, double *rhs, const double *pv) {
  FOR_EACH_CELL_BAND(g, c) {
    int i, k;
    cell_of(g, c, i, k);
    const size_t s = (size_t)i * g.NS + k;
    const int e = i * 3;
    const double sv = g.dsigv[k - 1];
    double dn[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int jv1 = g.vtx[e + j], jv2 = g.vtx[e + (j + 1) % 3];
      const double diff2 = (pv[(size_t)jv2 * g.NL + k] - pv[(size_t)jv1 * g.NL + k]) / g.dlVV[e + j];
      dn[j] = diff2 * g.dle2[e + j];
    }
    const double corr = dn[0] * g.dlVV[e] * sv + dn[1] * g.dlVV[e + 1] * sv + dn[2] * g.dlVV[e + 2] * sv;
    rhs[s] = rhs[s] - corr;
  }
}

int launch_update_rhsp(const Geo &g, hipStream_t s, double *rhs, const double *pv) {
  hipLaunchKernelGGL(k_update_rhsp, dim3(nsx_grid_for(g.ncell)), dim3(NSX_BLOCK), 0, s, g, rhs, pv);
  return (int)hipGetLastError();
}

// ----------------------------------------------------------- LSM gradient --
// grandientLSM (gradientLSM.F90:153-201): 3-neighbour least squares in the
// horizontal, 3-point least squares in sigma, then BCpcenter on the outputs:
// the column's least-squares sums, and the gradient at layer k0+m of a run
struct LsmCol {
  double dxc[3], dyc[3], sx2, sy2, sxy;
};

__device__ inline LsmCol lsm_col(const Geo &g, int i) {
  LsmCol L;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    L.dxc[j] = g.dxCC[i * 3 + j];
    L.dyc[j] = g.dyCC[i * 3 + j];
  }
  L.sx2 = g.sum_xc2[i];
  L.sy2 = g.sum_yc2[i];
  L.sxy = g.sum_xcyc[i];
  return L;
}

template <int RN>
__device__ inline void lsm_run(const Geo &g, const LsmCol &L, const FieldRun<RN> &F, const Run<RN> &R, int m,
                               double &dx, double &dy, double &ds) {
  const double fi = F.c[m + 1];
  double sumfx = 0.0, sumfy = 0.0;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const double d = F.n[j][m] - fi;
    sumfx = sumfx + L.dxc[j] * d;
    sumfy = sumfy + L.dyc[j] * d;
  }
  const double deter = L.sx2 * L.sy2 - L.sxy * L.sxy;
  dx = (L.sy2 * sumfx - L.sxy * sumfy) / deter;
  dy = (L.sx2 * sumfy - L.sxy * sumfx) / deter;
  const int k = R.k0 + m;
  const bool a = R.act[m];
  const double sk = a ? g.sig[k] : 0.0;
  const double dzp = (a ? g.sig[k + 1] : 1.0) - sk;
  const double dzm = (a ? g.sig[k - 1] : 1.0) - sk;
  const double s2 = dzp * dzp + dzm * dzm;
  ds = ((F.c[m + 2] - fi) * dzp + (F.c[m] - fi) * dzm) / s2;
}

__global__ __launch_bounds__(NSX_BLOCK) void k_grad_lsm(Geo g, const double *f, double *dx, double *dy,
                                                         double *ds) {
  FOR_EACH_RUN_BAND(g, t, RL) {
    const Run<RL> R = run_at<RL>(g, t);
    const EdgeRun E = edges_of(g, R.i);
    const LsmCol L = lsm_col(g, R.i);
    FieldRun<RL> F;
    field_run(g, R, E.nb, E.vx, f, nullptr, F);
    double a[RL], b[RL], d[RL];
#pragma unroll
    for (int m = 0; m < RL; ++m) lsm_run(g, L, F, R, m, a[m], b[m], d[m]);
    str(dx + R.s0, a, R);
    str(dy + R.s0, b, R);
    str(ds + R.s0, d, R);
#pragma unroll
    for (int m = 0; m < RL; ++m) {
      const int k = R.k0 + m;
      const size_t s = R.s0 + m;
      if (k == 1 && g.bcbot == 1) {
        dx[s - 1] = a[m];
        dy[s - 1] = b[m];
        ds[s - 1] = d[m];
      }
      if (k == g.NZ - 2 && g.bctop == 1) {
        dx[s + 1] = a[m];
        dy[s + 1] = b[m];
        ds[s + 1] = d[m];
      }
    }
  }
}

int launch_grad_lsm(const Geo &g, hipStream_t s, const double *f, double *dx, double *dy, double *ds) {
  hipLaunchKernelGGL(k_grad_lsm, dim3(nsx_grid_for(nruns<RL>(g))), dim3(NSX_BLOCK), 0, s, g, f, dx, dy, ds);
  return (int)hipGetLastError();
}

// ------------------------------------------------------------- corrector ---
// gradientP(pfnp) (gradientP.F90:103-105 -> grandientLSM) fused with the
// velocity corrector u = u* - dt grad p (hydro.F90:440-448) and the ghost
// layers of BCvelcenter3D for the new velocity (hydro.F90:452-455).
// Bytes/cell: read p, u*,v*,w* = 4; write u,v,w = 3.
__global__ __launch_bounds__(NSX_BLOCK) void k_corrector(Geo g, const double *p, const double *us,
                                                          const double *vs, const double *ws, double *un,
                                                          double *vn, double *wn) {
  FOR_EACH_RUN_BAND(g, t, RL) {
    const Run<RL> R = run_at<RL>(g, t);
    const EdgeRun E = edges_of(g, R.i);
    const LsmCol L = lsm_col(g, R.i);
    FieldRun<RL> P;
    field_run(g, R, E.nb, E.vx, p, nullptr, P);
    double a[RL], b[RL], c[RL];
    ldr(us + R.s0, a);
    ldr(vs + R.s0, b);
    ldr(ws + R.s0, c);
#pragma unroll
    for (int m = 0; m < RL; ++m) {
      double dpdx, dpdy, dpds;
      lsm_run(g, L, P, R, m, dpdx, dpdy, dpds);
      a[m] = a[m] - g.dt * dpdx;
      b[m] = b[m] - g.dt * dpdy;
      c[m] = c[m] - g.dt * dpds;
    }
    str(un + R.s0, a, R);
    str(vn + R.s0, b, R);
    str(wn + R.s0, c, R);
#pragma unroll
    for (int m = 0; m < RL; ++m) {
      const int k = R.k0 + m;
      const size_t s = R.s0 + m;
      double gv;
      if (k == 1) {
        if (vel_ghost(g.bcbot, g.BotBC, 0, a[m], gv)) un[s - 1] = gv;
        if (vel_ghost(g.bcbot, g.BotBC, 1, b[m], gv)) vn[s - 1] = gv;
        if (vel_ghost(g.bcbot, g.BotBC, 2, c[m], gv)) wn[s - 1] = gv;
      }
      if (k == g.NZ - 2) {
        if (vel_ghost(g.bctop, g.TopBC, 0, a[m], gv)) un[s + 1] = gv;
        if (vel_ghost(g.bctop, g.TopBC, 1, b[m], gv)) vn[s + 1] = gv;
        if (vel_ghost(g.bctop, g.TopBC, 2, c[m], gv)) wn[s + 1] = gv;
      }
    }
  }
}

int launch_corrector(const Geo &g, hipStream_t s, const double *p, const double *const us[3],
                     double *const unp[3]) {
  hipLaunchKernelGGL(k_corrector, dim3(nsx_grid_for(nruns<RL>(g))), dim3(NSX_BLOCK), 0, s, g, p, us[0], us[1], us[2],
                     unp[0], unp[1], unp[2]);
  return (int)hipGetLastError();
}

// ------------------------------------------------------------ face update --
// newu3D Kim-Choi branch (newu3D.F90:121-239): U?FACE -= dt dp/dn, the
// bottom face divided by dsig(k) (newu3D.F90:164).
// Bytes/cell: read p, p vertex (0.5), 5 faces = 6.5; write 5 faces.
__global__ __launch_bounds__(NSX_BLOCK) void k_newu(Geo g, const double *p, const double *pv, double *F1,
                                                     double *F2, double *F3, double *FT, double *FB) {
  FOR_EACH_RUN_BAND(g, t, RL) {
    const Run<RL> R = run_at<RL>(g, t);
    const int e = R.i * 3;
    const EdgeRun E = edges_of(g, R.i);
    double lcc[3], lvv[3], le1[3], le2[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      lcc[j] = g.dlCC[e + j];
      lvv[j] = g.dlVV[e + j];
      le1[j] = g.dle1[e + j];
      le2[j] = g.dle2[e + j];
    }
    FieldRun<RL> P;
    field_run(g, R, E.nb, E.vx, p, pv, P);
    double f1[RL], f2[RL], f3[RL], fT[RL], fB[RL];
    ldr(F1 + R.s0, f1);
    ldr(F2 + R.s0, f2);
    ldr(F3 + R.s0, f3);
    ldr(FT + R.s0, fT);
    ldr(FB + R.s0, fB);
    const double c1 = g.dt;
#pragma unroll
    for (int m = 0; m < RL; ++m) {
      const int k = R.k0 + m;
      const double pi = P.c[m + 1];
      double dn[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double diff1 = (P.n[j][m] - pi) / lcc[j];
        const double diff2 = (P.v[(j + 1) % 3][m] - P.v[j][m]) / lvv[j];
        dn[j] = diff1 * le1[j] + diff2 * le2[j];
      }
      const double dsk = R.act[m] ? g.dsig[k] : 1.0;
      const double dT = (P.c[m + 2] - pi) / dsk;
      const double dB = (P.c[m] - pi) / dsk;
      f1[m] = f1[m] - dn[0] * c1;
      f2[m] = f2[m] - dn[1] * c1;
      f3[m] = f3[m] - dn[2] * c1;
      fT[m] = fT[m] - dT * c1;
      fB[m] = fB[m] - dB * c1;
    }
    str(F1 + R.s0, f1, R);
    str(F2 + R.s0, f2, R);
    str(F3 + R.s0, f3, R);
    str(FT + R.s0, fT, R);
    str(FB + R.s0, fB, R);
  }
}

int launch_newu(const Geo &g, hipStream_t s, const double *p, const double *pv, double *const face[5]) {
  hipLaunchKernelGGL(k_newu, dim3(nsx_grid_for(nruns<RL>(g))), dim3(NSX_BLOCK), 0, s, g, p, pv, face[0], face[1],
                     face[2], face[3], face[4]);
  return (int)hipGetLastError();
}

// ------------------------------------------------------------ RK2 average --
// nsmp3D.F90:531-548: u,v,w(n+1) = 0.5 (u(n) + u(n+1)) over every stored
// entry (ghost layers and wall ghost columns included).  Bytes/cell: read 6, write 3.
__global__ __launch_bounds__(NSX_BLOCK) void k_rk2avg(size_t n, const double *u0, const double *v0,
                                                       const double *w0, double *u1, double *v1, double *w1) {
  for (size_t t = (size_t)blockIdx.x * NSX_BLOCK + threadIdx.x; t < n; t += (size_t)gridDim.x * NSX_BLOCK) {
    u1[t] = 0.5 * (u0[t] + u1[t]);
    v1[t] = 0.5 * (v0[t] + v1[t]);
    w1[t] = 0.5 * (w0[t] + w1[t]);
  }
}

int launch_rk2avg(const Geo &g, hipStream_t s, const double *const un[3], double *const unp[3]) {
  const size_t n = (size_t)(g.N0 + g.NGW) * g.NS;
  hipLaunchKernelGGL(k_rk2avg, dim3(nsx_grid_for((uint32_t)(n < 0xffffffffu ? n : 0xffffffffu))), dim3(NSX_BLOCK),
                     0, s, n, un[0], un[1], un[2], unp[0], unp[1], unp[2]);
  return (int)hipGetLastError();
}

// the initial guess 2 a - b of an extrapolating pressure solve, a the newer
// of two solutions
__global__ __launch_bounds__(NSX_BLOCK) void k_extrapolate(size_t n, const double *a, const double *b, double *o) {
  for (size_t t = (size_t)blockIdx.x * NSX_BLOCK + threadIdx.x; t < n; t += (size_t)gridDim.x * NSX_BLOCK)
    o[t] = a[t] + (a[t] - b[t]);
}

int launch_extrapolate(const Geo &g, hipStream_t s, const double *a, const double *b, double *o) {
  const size_t n = (size_t)(g.N0 + g.NGW) * g.NS;
  hipLaunchKernelGGL(k_extrapolate, dim3(nsx_grid_for((uint32_t)(n < 0xffffffffu ? n : 0xffffffffu))),
                     dim3(NSX_BLOCK), 0, s, n, a, b, o);
  return (int)hipGetLastError();
}

// --------------------------------------------------------- ghost layers ----
// vertical part of BCpcenter (kind 0, BC.F90:919-938) and BCvelcenter3D
// (kinds 1..3 = u,v,w, BC.F90:459-527) on a stored field
__global__ __launch_bounds__(NSX_BLOCK) void k_ghost_layers(Geo g, double *f, int kind) {
  const int i = blockIdx.x * NSX_BLOCK + threadIdx.x;
  if (i >= g.N0) return;
  const size_t b = (size_t)i * g.NS;
  if (kind == 0) {
    if (g.bcbot == 1) f[b] = f[b + 1];
    if (g.bctop == 1) f[b + g.NZ - 1] = f[b + g.NZ - 2];
    return;
  }
  double gv;
  if (vel_ghost(g.bcbot, g.BotBC, kind - 1, f[b + 1], gv)) f[b] = gv;
  if (vel_ghost(g.bctop, g.TopBC, kind - 1, f[b + g.NZ - 2], gv)) f[b + g.NZ - 1] = gv;
}

// ------------------------------------------------------ wall ghost cells ---
// the horizontal part of BCvelcenter3D (kinds 1..3 = u,v,w, BC.F90:540-581)
// and BCpcenter (kind 0, BC.F90:945-964) for wall ghosts: every stored layer
// of the ghost column takes its cell's value, copied (flag 1) or negated
// (flag 0: 0 - phi for the velocities, -phi for p, as the reference writes
// them), or for an inflow side (flag 2) 2 inflow() - phi.  Periodic ghosts need no fill: the neighbour index points at the
// partner.  Run after the sigma ghost layers, whose values it copies too.
__global__ __launch_bounds__(NSX_BLOCK) void k_wall_ghosts(Geo g, double *f, int kind) {
  const uint32_t t = blockIdx.x * NSX_BLOCK + threadIdx.x;
  if (t >= (uint32_t)g.NGW * (uint32_t)g.NZ) return;
  const int w = (int)(t / (uint32_t)g.NZ), k = (int)(t - (uint32_t)w * (uint32_t)g.NZ);
  const double x = f[(size_t)g.gw_owner[w] * g.NS + k];
  double y;
  if ((g.gw_neu[w] >> kind) & 1)
    y = x;
  else if ((g.gw_inf[w] >> kind) & 1)
    y = 2.0 * (kind == 2 ? 1.0 : 0.0) - x;  // inflow() of Functions.F90:109-121: 1 for v, 0 for u, w
  else
    y = kind == 0 ? -x : 0.0 - x;
  f[(size_t)(g.N0 + w) * g.NS + k] = y;
}

int launch_wall_ghosts(const Geo &g, hipStream_t s, double *f, int kind) {
  if (g.NGW == 0) return 0;
  const uint32_t n = (uint32_t)g.NGW * (uint32_t)g.NZ;
  hipLaunchKernelGGL(k_wall_ghosts, dim3((n + NSX_BLOCK - 1) / NSX_BLOCK), dim3(NSX_BLOCK), 0, s, g, f, kind);
  return (int)hipGetLastError();
}

int launch_ghost_layers(const Geo &g, hipStream_t s, double *f, int kind) {
  hipLaunchKernelGGL(k_ghost_layers, dim3((g.N0 + NSX_BLOCK - 1) / NSX_BLOCK), dim3(NSX_BLOCK), 0, s, g, f, kind);
  return (int)hipGetLastError();
}

// dst = src over n doubles: the field copy of nsx_copy_field, also the
// known-size copy that HBM traffic counters are calibrated against (reads and
// writes 8 n bytes)
__global__ __launch_bounds__(NSX_BLOCK) void k_copy(size_t n, const double *src, double *dst) {
  for (size_t q = (size_t)blockIdx.x * NSX_BLOCK + threadIdx.x; q < n; q += (size_t)gridDim.x * NSX_BLOCK)
    dst[q] = src[q];
}

int launch_copy(hipStream_t s, size_t n, const double *src, double *dst) {
  size_t blocks = (n + NSX_BLOCK - 1) / NSX_BLOCK;
  if (blocks > 8192) blocks = 8192;
  if (blocks == 0) return 0;
  hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(NSX_BLOCK), 0, s, n, src, dst);
  return (int)hipGetLastError();
}
