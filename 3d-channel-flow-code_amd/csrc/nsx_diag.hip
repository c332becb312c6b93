// nsx_diag.hip -- glob_cfl on the device (cfl.F90:7-269, sequential branch):
// the largest velocities, the volume means, the CFL number and the mass-flux
// checks of the face velocities, from the resident fields, without a host
// copy of any field.
//
// One pass over the interior prisms writes per-block partials; one block then
// reduces them in a fixed order.  The "largest magnitude" values keep the
// reference's choice among equal magnitudes -- the first in its k-then-i scan
// (strict > in cfl.F90:98-106,175-177) -- by carrying the scan index
// (k-1)*N_CELL0 + i with each candidate.  Sums are in a different order than
// the reference's loops (round-off only).
// Bytes/prism: read u, v, w and 5 face velocities = 8 x 8 B.
#include <mutex>

#include "nsx_kernels.h"

namespace {

constexpr int NSUM = 7;   // sumu, sumv, sumw, flux_l, flux_i, flux_o, flux_w
constexpr int NMAX = 4;   // u, v, w, flux (signed value of largest magnitude)
constexpr int NROWS = NSUM + 1 + 2 * NMAX;  // + cfl max + (value, index) pairs

struct Cand {
  double v;
  double idx;  // scan index; 2^53 for none
};

__device__ inline void take(Cand &a, double v, double idx) {
  const double m = fabs(v), b = fabs(a.v);
  if (m > b || (m == b && m > 0.0 && idx < a.idx)) {
    a.v = v;
    a.idx = idx;
  }
}

__device__ inline void wave_combine(double (&s)[NSUM], double &cfl, Cand (&c)[NMAX]) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int q = 0; q < NSUM; ++q) s[q] += __shfl_down(s[q], off, 64);
    cfl = fmax(cfl, __shfl_down(cfl, off, 64));
#pragma unroll
    for (int q = 0; q < NMAX; ++q) {
      const double v = __shfl_down(c[q].v, off, 64), i = __shfl_down(c[q].idx, off, 64);
      take(c[q], v, i);
    }
  }
}

__global__ __launch_bounds__(NSX_BLOCK) void k_glob_cfl(Geo g, const double *u, const double *v, const double *w,
                                                         const double *U1, const double *U2, const double *U3,
                                                         const double *UT, const double *UB, double *P) {
  double s[NSUM] = {0, 0, 0, 0, 0, 0, 0};
  double cfl = 0.0;
  Cand c[NMAX];
#pragma unroll
  for (int q = 0; q < NMAX; ++q) c[q] = Cand{0.0, 9007199254740992.0};
  uint32_t beg, end;
  block_range(g.ncell, beg, end);
  for (uint32_t base = beg; base < end; base += NSX_BLOCK) {
    const uint32_t cc = base + threadIdx.x;
    if (cc >= end) continue;
    int i, k;
    cell_of(g, cc, i, k);
    const size_t sidx = (size_t)i * g.NS + k;
    const double scan = (double)(k - 1) * (double)g.N0 + (double)i;
    const double area = g.area[i];
    take(c[0], u[sidx], scan);
    take(c[1], v[sidx], scan);
    take(c[2], w[sidx], scan);
    s[0] += u[sidx] * area * g.dsig[k];
    s[1] += v[sidx] * area * g.dsig[k];
    s[2] += w[sidx] * area * g.dsig[k];
    const double uf[3] = {U1[sidx], U2[sidx], U3[sidx]};
    double flux = 0.0, fc = 0.0;
    const int e = i * 3;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double dl = g.dlVV[e + j];
      flux = flux + uf[j] * dl * g.dsig[k - 1];
      fc = fc + fabs(uf[j] * dl * g.dsig[k]);
      if (g.eside[e + j]) s[6] += uf[j] * dl * g.dsig[k];
    }
    const double ut = UT[sidx], ub = UB[sidx];
    flux = flux + ut * area;
    flux = flux + ub * area;
    fc = fc + fabs(ub * area) + fabs(ut * area);
    fc = 0.5 * g.dt * fc / (area * g.dsig[k]);
    cfl = fmax(cfl, fc);
    s[3] += flux;
    take(c[3], flux, scan);
    if (k == 1 && g.bcbot == 1) s[4] += ub * area;
    if (k == g.NZ - 2 && g.bctop == 1) s[5] += ut * area;
  }
  wave_combine(s, cfl, c);
  __shared__ double sh[NROWS][NSX_BLOCK / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < NSUM; ++q) sh[q][wid] = s[q];
    sh[NSUM][wid] = cfl;
#pragma unroll
    for (int q = 0; q < NMAX; ++q) {
      sh[NSUM + 1 + 2 * q][wid] = c[q].v;
      sh[NSUM + 2 + 2 * q][wid] = c[q].idx;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t G = gridDim.x, b = blockIdx.x;
    for (int q = 0; q < NSUM; ++q) {
      double t = 0.0;
      for (int ww = 0; ww < NSX_BLOCK / 64; ++ww) t += sh[q][ww];
      P[(size_t)q * G + b] = t;
    }
    double m = 0.0;
    for (int ww = 0; ww < NSX_BLOCK / 64; ++ww) m = fmax(m, sh[NSUM][ww]);
    P[(size_t)NSUM * G + b] = m;
    for (int q = 0; q < NMAX; ++q) {
      Cand a{0.0, 9007199254740992.0};
      for (int ww = 0; ww < NSX_BLOCK / 64; ++ww) take(a, sh[NSUM + 1 + 2 * q][ww], sh[NSUM + 2 + 2 * q][ww]);
      P[(size_t)(NSUM + 1 + 2 * q) * G + b] = a.v;
      P[(size_t)(NSUM + 2 + 2 * q) * G + b] = a.idx;
    }
  }
}

// one block: the totals in block order, and the 13 outputs of glob_cfl
__global__ __launch_bounds__(NSX_BLOCK) void k_glob_cfl_final(Geo g, const double *P, uint32_t G, double *out) {
  double s[NSUM] = {0, 0, 0, 0, 0, 0, 0};
  double cfl = 0.0;
  Cand c[NMAX];
#pragma unroll
  for (int q = 0; q < NMAX; ++q) c[q] = Cand{0.0, 9007199254740992.0};
  for (uint32_t b = threadIdx.x; b < G; b += NSX_BLOCK) {
#pragma unroll
    for (int q = 0; q < NSUM; ++q) s[q] += P[(size_t)q * G + b];
    cfl = fmax(cfl, P[(size_t)NSUM * G + b]);
#pragma unroll
    for (int q = 0; q < NMAX; ++q) take(c[q], P[(size_t)(NSUM + 1 + 2 * q) * G + b], P[(size_t)(NSUM + 2 + 2 * q) * G + b]);
  }
  wave_combine(s, cfl, c);
  __shared__ double sh[NROWS][NSX_BLOCK / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    for (int q = 0; q < NSUM; ++q) sh[q][wid] = s[q];
    sh[NSUM][wid] = cfl;
    for (int q = 0; q < NMAX; ++q) {
      sh[NSUM + 1 + 2 * q][wid] = c[q].v;
      sh[NSUM + 2 + 2 * q][wid] = c[q].idx;
    }
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double t[NSUM];
  for (int q = 0; q < NSUM; ++q) {
    t[q] = 0.0;
    for (int ww = 0; ww < NSX_BLOCK / 64; ++ww) t[q] += sh[q][ww];
  }
  double m = 0.0;
  for (int ww = 0; ww < NSX_BLOCK / 64; ++ww) m = fmax(m, sh[NSUM][ww]);
  Cand a[NMAX];
  for (int q = 0; q < NMAX; ++q) {
    a[q] = Cand{0.0, 9007199254740992.0};
    for (int ww = 0; ww < NSX_BLOCK / 64; ++ww) take(a[q], sh[NSUM + 1 + 2 * q][ww], sh[NSUM + 2 + 2 * q][ww]);
  }
  const double Vol = g.TotalVolume * (g.sigFin - g.sigIni);
  out[0] = a[0].v;
  out[1] = a[1].v;
  out[2] = a[2].v;
  out[3] = m;
  out[4] = t[0] / Vol;
  out[5] = t[1] / Vol;
  out[6] = t[2] / Vol;
  out[7] = t[3];
  out[8] = a[3].v;
  out[9] = t[4] + t[5] + t[6];
  out[10] = t[4];
  out[11] = t[5];
  out[12] = t[6];
}

// The scratch of nsx_glob_cfl_noalloc: the per-block partial rows of one
// pass (NROWS rows of at most CFL_MAXGRID blocks) and the 13 outputs (16
// slots), 262,272 bytes.  Each GPU the library runs on holds one copy until
// the process ends; every call on a GPU shares it, so calls that overlap on
// one GPU race.  Host code never passes `scratch` itself to a kernel (in
// host code the name is a stand-in, not a device address): it passes the
// device's address that glob_cfl_scratch_init obtained.
constexpr int CFL_MAXGRID = 2048;
constexpr size_t CFL_SCRATCH = (size_t)NROWS * CFL_MAXGRID + 16;
static __device__ __align__(256) double scratch[CFL_SCRATCH];

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
double *scratch_ptr[MAX_DEVICES] = {};

}  // namespace

int glob_cfl_rows() { return NROWS; }

// The device address of `scratch` on device dev, obtained once per device
// with dev current (the library's set-up, nsx_commit, calls this); 0 or a
// hipError_t.
int glob_cfl_scratch_init(int dev) {
  if (dev < 0 || dev >= MAX_DEVICES) return (int)hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (scratch_ptr[dev]) return 0;
  int cur = -1;
  hipError_t e = hipGetDevice(&cur);
  if (e != hipSuccess) return (int)e;
  if (cur != dev && (e = hipSetDevice(dev)) != hipSuccess) return (int)e;
  void *p = nullptr;
  e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (cur != dev) (void)hipSetDevice(cur);
  if (e != hipSuccess) return (int)e;
  scratch_ptr[dev] = (double *)p;
  return 0;
}

// glob_cfl into the scratch of device dev: partial rows, then the outputs
// at out_slot; fails when the grid exceeds the scratch (-1) or the scratch
// of dev was not set up (-2)
int launch_glob_cfl_scratch(const Geo &g, hipStream_t s, int dev, const double *u, const double *v, const double *w,
                            const double *const face[5], double **out_slot) {
  if (g.grid > CFL_MAXGRID) return -1;
  double *base = nullptr;
  {
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (dev >= 0 && dev < MAX_DEVICES) base = scratch_ptr[dev];
  }
  if (!base) return -2;
  double *part = base, *out = base + (size_t)NROWS * CFL_MAXGRID;
  *out_slot = out;
  return launch_glob_cfl(g, s, u, v, w, face, part, out);
}

int launch_glob_cfl(const Geo &g, hipStream_t s, const double *u, const double *v, const double *w,
                    const double *const face[5], double *part, double *out) {
  hipLaunchKernelGGL(k_glob_cfl, dim3(g.grid), dim3(NSX_BLOCK), 0, s, g, u, v, w, face[0], face[1], face[2], face[3],
                     face[4], part);
  hipLaunchKernelGGL(k_glob_cfl_final, dim3(1), dim3(NSX_BLOCK), 0, s, g, (const double *)part, (uint32_t)g.grid,
                     out);
  return (int)hipGetLastError();
}
