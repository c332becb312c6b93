// nsx_stats.hip -- channel turbulence statistics of the reference
// (src/Stats/stats.F90) on the resident fields: the twelve running sums of
// glob_sample and the plane sums of glob_put_plane_mm (in its summation
// order); nsx_files.hip writes the files.
//
// Layout: the accumulators are cell fields, double[N0][NS] like every other
// cell field of the library, and cover the layers k = 0..NZ-1 (ghost layers
// included, as glob_sample's loop does).
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "nsx_kernels.h"

// glob_sample, stats.F90:142-174: every column i < N_CELL0 and every layer
// (k = 1..NZ of the reference), in its operation order
__global__ __launch_bounds__(NSX_BLOCK) void k_glob_sample(Geo g, const double *u, const double *v,
                                                            const double *w, const double *p, const double *eddy,
                                                            StatsBufs sm) {
  const size_t n = (size_t)g.N0 * g.NS;
  for (size_t t = (size_t)blockIdx.x * NSX_BLOCK + threadIdx.x; t < n; t += (size_t)gridDim.x * NSX_BLOCK) {
    const int k = (int)(t % (size_t)g.NS);
    if (k >= g.NZ) continue;
    const double tu = u[t], tv = v[t], tw = w[t], tp = p[t];
    const double tp2 = tp * tp, tu2 = tu * tu, tv2 = tv * tv, tw2 = tw * tw;
    const double tuv = tu * tv, tvw = tv * tw, twu = tw * tu;
    sm.a[SM_UF][t] = sm.a[SM_UF][t] + tu;
    sm.a[SM_VF][t] = sm.a[SM_VF][t] + tv;
    sm.a[SM_WF][t] = sm.a[SM_WF][t] + tw;
    sm.a[SM_PF][t] = sm.a[SM_PF][t] + tp;
    sm.a[SM_U2F][t] = sm.a[SM_U2F][t] + tu2;
    sm.a[SM_V2F][t] = sm.a[SM_V2F][t] + tv2;
    sm.a[SM_W2F][t] = sm.a[SM_W2F][t] + tw2;
    sm.a[SM_UVF][t] = sm.a[SM_UVF][t] + tuv;
    sm.a[SM_VWF][t] = sm.a[SM_VWF][t] + tvw;
    sm.a[SM_WUF][t] = sm.a[SM_WUF][t] + twu;
    sm.a[SM_P2F][t] = sm.a[SM_P2F][t] + tp2;
    if (eddy) sm.a[SM_EF][t] = sm.a[SM_EF][t] + eddy[t];
  }
}

// glob_put_plane_mm, stats.F90:266-279: S_q(k) = sum over i = 1..N_CELL0, in
// order, of sm_q(i,k)/ns, for the nine velocity moments and the interior
// layers.  One block per layer: the block stages NSX_BLOCK columns of the
// nine accumulators in LDS, then thread q adds its quantity's values in
// column order, so each sum runs in the reference's sequence.
__global__ __launch_bounds__(NSX_BLOCK) void k_plane_sums(Geo g, StatsBufs sm, double ns, double *out) {
  __shared__ double tile[SM_NPLANE][NSX_BLOCK];
  const int k = 1 + (int)blockIdx.x;  // interior layer
  const int q = threadIdx.x;
  double s = 0.0;
  for (int i0 = 0; i0 < g.N0; i0 += NSX_BLOCK) {
    const int i = i0 + threadIdx.x;
    const int m = g.N0 - i0 < NSX_BLOCK ? g.N0 - i0 : NSX_BLOCK;
    if (i < g.N0) {
      const size_t at = (size_t)i * g.NS + k;
#pragma unroll
      for (int r = 0; r < SM_NPLANE; ++r) tile[r][threadIdx.x] = sm.a[r][at];
    }
    __syncthreads();
    if (q < SM_NPLANE)
      for (int j = 0; j < m; ++j) s = s + tile[q][j] / ns;
    __syncthreads();
  }
  if (q < SM_NPLANE) out[(size_t)q * g.NZ + k] = s;
}

int launch_glob_sample(const Geo &g, hipStream_t s, const double *u, const double *v, const double *w,
                       const double *p, const double *eddy, StatsBufs sm) {
  const size_t n = (size_t)g.N0 * g.NS;
  size_t blocks = (n + NSX_BLOCK - 1) / NSX_BLOCK;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_glob_sample, dim3((unsigned)blocks), dim3(NSX_BLOCK), 0, s, g, u, v, w, p, eddy, sm);
  return (int)hipGetLastError();
}

int launch_plane_sums(const Geo &g, hipStream_t s, StatsBufs sm, int ns, double *out) {
  if (g.NZ < 3) return 0;
  hipLaunchKernelGGL(k_plane_sums, dim3((unsigned)(g.NZ - 2)), dim3(NSX_BLOCK), 0, s, g, sm, (double)ns, out);
  return (int)hipGetLastError();
}
