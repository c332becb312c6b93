// nsx_dev.h -- device-side data layout and helpers of the nsx HIP library.
//
// Layout in HBM (gfx950):
//   cell field   double[N0][NS]   column-contiguous: entry (i,k) at i*NS+k, NS =
//                                 NZ rounded up to a multiple of 4 (16-byte
//                                 aligned runs of 2 doubles / 4 floats); the
//                                 ghost layers k=0 and k=NZ-1 are stored; a
//                                 periodic horizontal ghost cell is not: the
//                                 neighbour index nbr[i*3+j] already points at
//                                 its partner column; a wall ghost cell is a
//                                 column of its own after the N0 interior ones
//                                 (N0+w), filled from its cell by the wall rule;
//   vertex field double[NV][NZ-1] entry (v,l) at v*(NZ-1)+l;
//   face field   double[N0][NS]   U1FACE..UBFACE, same indexing as cells.
// A wavefront walks consecutive k of one column (then the next column), so
// every access to the cell itself, to its vertical neighbours and to a
// horizontal neighbour column is a contiguous run.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define NSX_BLOCK 256
#define NSX_NXCD 8

// n / d for 32-bit n and a run-time divisor (Granlund-Montgomery, round-up
// variant with a 33-bit intermediate).
struct FastDiv {
  uint32_t d, m, l;
};

static inline FastDiv make_fastdiv(uint32_t d) {
  FastDiv f;
  f.d = d;
  uint32_t l = 0;
  while ((1ull << l) < d) ++l;
  f.l = l;
  f.m = (uint32_t)((((1ull << l) - d) << 32) / d + 1);
  return f;
}

__host__ __device__ inline uint32_t fdiv(uint32_t n, const FastDiv &f) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t t = __umulhi(f.m, n);
#else
  uint32_t t = (uint32_t)(((uint64_t)f.m * n) >> 32);
#endif
  return (uint32_t)(((uint64_t)t + n) >> f.l);
}

// Vertex boundary action per field: what BCvelvertex3D / BCpvertex and
// bc_periodic do to a vertex value (BC.F90:603-1408).
enum : signed char { VA_NONE = 0, VA_ZERO = 1, VA_INFLOW = 2, VA_PERIODIC = 3 };
// Periodic vertex kind (vertex_edge / vertex_pair of calcul_geometry.F90:920-1025)
enum : signed char { VK_INTERIOR = 0, VK_PAIR = 1, VK_CORNER = 2, VK_OTHER = 3 };

struct Geo {
  int N0, NV, NZ, NZi, NL;  // NZi = NZ-2 interior layers, NL = NZ-1 vertex levels
  int NS;                   // column stride of cell/face fields: NZ rounded up to a multiple of 4
  uint32_t ncell;           // N0*NZi interior prisms
  int NQ;                   // NS/4 four-layer groups per column
  int grid;                 // grid of the kernels that produce or consume solver partial sums
  int grid_cg;              // grid of the FP32 CG stencil kernel (its own occupancy)
  int grid_cgs;             // grid of the single-reduction FP32 CG kernel (its own occupancy)
  int grid_cgs64;           // grid of its FP64 instance
  uint32_t nquad;           // N0*NQ
  FastDiv divNZi, divNL, divNQ;
  // per cell-edge (N0*3, edge-fastest: index i*3+j)
  const int *nbr;           // neighbour column across edge j (periodic partner resolved), 0-based
  const int *vtx;           // No_vp(i,j), 0-based
  const double *dlVV, *dlCC, *dle1, *dle2, *doe, *xnn, *ynn, *dxCC, *dyCC;
  const double *ppec;       // (dle1/dlCC)*dlVV: PPE edge coefficient before the dsigv factor
  const double *ppec_t;     // ppec of the neighbour's edge back towards i: the lateral couplings of A^T
  int trans;                // 1 while the solver applies A^T (the adjoint solve)
  const signed char *eside; // 1 when No_cp(i,j) is not an interior cell (a side face of glob_cfl)
  // wall ghost columns N0 .. N0+NGW-1 (walled sides of the box)
  int NGW;
  const int *gw_owner;        // the interior column whose edge the ghost lies behind
  const unsigned char *gw_neu; // bit q set: flag 1 (copy) for kind q (0 p, 1 u, 2 v, 3 w); clear: flag 0 or 2
  const unsigned char *gw_inf; // bit q set: flag 2 (inflow: 2 inflow() - phi) for kind q
  // per cell (N0)
  const double *area, *sum_xc2, *sum_yc2, *sum_xcyc;
  // per level
  const double *sig, *dsig;  // NZ
  const double *sigv, *dsigv;  // NZ-1
  // FP32 copies for the FP32 inner solver of the mixed-precision form
  const float *area_f, *ppec_f, *dsig_f, *dsigv_f, *ppec_t_f;
  const double *walld;       // NZ
  // per prism (i*NS+k, interior layers): (Cs Delta)^2 of sgles with Mason's
  // wall damping, which depends on the geometry only (set at commit)
  const double *nutg;
  // vertices: inverse-distance stencil in CSR form (interior cells, reference order)
  const int *vptr, *vcell;
  const double *vw, *dlVsum;
  const int *vpair;          // 0-based partner or -1
  const signed char *vkind;  // VK_*
  const signed char *vact_u, *vact_v, *vact_w, *vact_p;  // VA_* per vertex
  int corner[4], ncorner;
  int bctop, bcbot, TopBC, BotBC;
  double TotalVolume, sigIni, sigFin;
  double dt, nu, bdx, bdy, bdz;
};

// the four layers at p (16-byte aligned: a run k0..k0+3 with k0 = 4m) as
// 16-byte accesses
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

// velocity ghost layer rule of BCvelcenter3D (BC.F90:459-527) for ZPB = 0:
// the derivative of the ghost value with respect to the interior value next
// to it, +1 symmetric copy, -1 antisymmetric copy (and the lid's 2 - u), 0
// no fill
__host__ __device__ inline int vel_ghost_sign(int bcflag, int wallbc, int comp /*0 u,1 v,2 w*/) {
  if (bcflag != 1) return 0;
  if (wallbc == 0) return -1;
  if (wallbc == 1) return comp == 2 ? -1 : 1;
  if (wallbc == 2) return -1;
  return 0;
}

// the ghost value itself from the interior value x: no slip -x, free slip x
// (w: -x), the lid (wallbc 2, the top only) 2 - x for u and -x for v, w;
// false when bcflag says there is no wall to fill
__host__ __device__ inline bool vel_ghost(int bcflag, int wallbc, int comp, double x, double &out) {
  if (bcflag != 1) return false;
  if (wallbc == 0) {
    out = -x;
  } else if (wallbc == 1) {
    out = comp == 2 ? -x : x;
  } else if (wallbc == 2) {
    out = comp == 0 ? 2.0 - x : -x;
  } else {
    return false;
  }
  return true;
}

// ------------------------------------------------------------ iteration ---
// Interior prisms are numbered c = col*NZi + (k-1).  A launch uses a fixed
// grid; block b works on one contiguous range of c.  Blocks are dealt
// round-robin to the 8 XCDs, so the logical block index is remapped to give
// each XCD (and its L2) one contiguous slice of the columns.
__device__ inline void block_range(uint32_t n, uint32_t &beg, uint32_t &end) {
  const uint32_t G = gridDim.x, b = blockIdx.x;
  uint32_t lb = b;
  if ((G % NSX_NXCD) == 0) lb = (b % NSX_NXCD) * (G / NSX_NXCD) + b / NSX_NXCD;
  uint64_t chunk = ((uint64_t)n + G - 1) / G;
  chunk = (chunk + NSX_BLOCK - 1) / NSX_BLOCK * NSX_BLOCK;
  uint64_t b0 = (uint64_t)lb * chunk, b1 = b0 + chunk;
  beg = (uint32_t)(b0 < n ? b0 : n);
  end = (uint32_t)(b1 < n ? b1 : n);
}

__device__ inline void cell_of(const Geo &g, uint32_t c, int &col, int &k) {
  uint32_t q = fdiv(c, g.divNZi);
  col = (int)q;
  k = (int)(c - q * (uint32_t)g.NZi) + 1;
}

// ---------------------------------------------------------- reductions ---
// Deterministic grid reductions without device-scope fences.  A kernel that
// produces sums writes one partial per block to P[q*G + block]; the kernel
// that needs the totals runs next on the stream (the kernel boundary makes
// the partials visible) and every one of its blocks sums the G partials in
// the same fixed order, so all blocks hold bit-identical totals and take the
// same decisions.  (A last-block/ticket reduction needs a device-scope
// release fence per block, which on gfx950 writes back the XCD's L2; with
// thousands of blocks that costs more than the stencil itself.)
template <int K>
__device__ inline void wave_sum(double (&v)[K]) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int q = 0; q < K; ++q) v[q] += __shfl_down(v[q], off, 64);
}

// (BLK: the block's threads)
template <int K, int BLK = NSX_BLOCK>
__device__ inline void write_partials(double (&v)[K], double *P) {
  __shared__ double sh[K][BLK / 64];
  wave_sum<K>(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < K; ++q) sh[q][wid] = v[q];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < K; ++q) {
      double s = 0.0;
      for (int w = 0; w < BLK / 64; ++w) s += sh[q][w];
      P[(size_t)q * gridDim.x + blockIdx.x] = s;
    }
}

// totals of the K partial rows written by the previous kernel (grid G); every
// thread of the block receives them
template <int K>
__device__ inline void reduce_partials(const double *P, uint32_t G, double (&tot)[K]) {
  __shared__ double sh[K][NSX_BLOCK / 64];
  double v[K];
#pragma unroll
  for (int q = 0; q < K; ++q) {
    double s = 0.0;
    for (uint32_t b = threadIdx.x; b < G; b += NSX_BLOCK) s += P[(size_t)q * G + b];
    v[q] = s;
  }
  wave_sum<K>(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < K; ++q) sh[q][wid] = v[q];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < K; ++q) {
    double s = 0.0;
    for (int w = 0; w < NSX_BLOCK / 64; ++w) s += sh[q][w];
    tot[q] = s;
  }
  __syncthreads();
}

#define HIPCHECK(x)                                                     \
  do {                                                                  \
    hipError_t e_ = (x);                                                \
    if (e_ != hipSuccess) return nsx_fail_hip(ctx, e_, #x, __LINE__);   \
  } while (0)
