// nsx_stats.hip -- channel turbulence statistics of the reference
// (src/Stats/stats.F90) on the resident fields: the twelve running sums of
// glob_sample, the plane sums of glob_put_plane_mm (in its summation order),
// and the three files that routine writes.
//
// Layout: the accumulators are cell fields, double[N0][NS] like every other
// cell field of the library, and cover the layers k = 0..NZ-1 (ghost layers
// included, as glob_sample's loop does).
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

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

// ------------------------------------------------------------- the files ---
// Fortran Ew.d output (F2018 13.7.2.3.3, no scale factor): 0.ddddd, then
// E+xx (E+xxx dropping the E above 99), right-justified in w; the digits are
// the value correctly rounded to d significant digits.
static void fmt_e(char *out, double x, int w, int d) {
  char body[64];
  if (isnan(x)) {
    snprintf(out, 64, "%*s", w, "NaN");
    return;
  }
  if (isinf(x)) {
    snprintf(out, 64, "%*s", w, x < 0 ? "-Inf" : "Inf");
    return;
  }
  if (x == 0.0) {
    char z[32];
    memset(z, '0', (size_t)d);
    z[d] = 0;
    snprintf(body, sizeof body, "%s0.%sE+00", signbit(x) ? "-" : "", z);
  } else {
    char sci[64];
    snprintf(sci, sizeof sci, "%.*e", d - 1, fabs(x));  // d.ddd e+XX, correctly rounded
    char digits[32];
    int nd = 0;
    for (const char *c = sci; *c && *c != 'e'; ++c)
      if (*c >= '0' && *c <= '9') digits[nd++] = *c;
    digits[nd] = 0;
    const int e10 = atoi(strchr(sci, 'e') + 1) + 1;  // exponent of 0.ddd form
    if (e10 >= -99 && e10 <= 99)
      snprintf(body, sizeof body, "%s0.%sE%c%02d", x < 0 ? "-" : "", digits, e10 < 0 ? '-' : '+', abs(e10));
    else
      snprintf(body, sizeof body, "%s0.%s%c%03d", x < 0 ? "-" : "", digits, e10 < 0 ? '-' : '+', abs(e10));
  }
  snprintf(out, 64, "%*s", w, body);
}

// The list-directed REAL(8) output of the compiler the reference is built
// with (amdflang/flang).  Digits: the shortest decimal between the two
// midpoints from x to its neighbours, chosen as that compiler's runtime
// chooses it: the exact expansions of the lower midpoint and of the upper
// one less one unit in its last place are cut back digit by digit until all
// but their last digits agree and the upper one's last digit is not 0; the
// last digit is then the upper one's minus half (rounded down) the
// difference of the two.  (Checked on 1400 values against that compiler.)
// Form: F (no leading zero before the point) when the value rounded to one
// significant digit is 0.d x 10^e with 0 <= e <= 15, E (d.ddE+xx) otherwise.
// The midpoints are exact in the host's 80-bit long double.
static void exact_fixed(long double v, int frac, std::string &out) {
  std::vector<char> buf(frac + 400);
  snprintf(buf.data(), buf.size(), "%.*Lf", frac, v);
  out.clear();
  for (const char *c = buf.data(); *c; ++c)
    if (*c != '.') out.push_back(*c);
}

static void fmt_list_real(char *out, double x) {
  if (x == 0.0) {
    snprintf(out, 64, "%s0.", signbit(x) ? "-" : "");
    return;
  }
  if (!isfinite(x)) {
    snprintf(out, 64, "%s", isnan(x) ? "NaN" : (x < 0 ? "-Inf" : "Inf"));
    return;
  }
  const double ax = fabs(x);
  const long double lo = nextafter(ax, 0.0), hi = nextafter(ax, INFINITY);
  const long double less = ((long double)ax + lo) / 2;
  const long double more = isfinite(hi) ? ((long double)ax + hi) / 2 : (long double)ax + ((long double)ax - lo) / 2;
  const int FRAC = 1130;  // past the last nonzero digit of any normal double's midpoint
  std::string L, M;
  exact_fixed(less, FRAC, L);
  exact_fixed(more, FRAC, M);
  const size_t n = L.size() > M.size() ? L.size() : M.size();
  L.insert(0, n - L.size(), '0');
  M.insert(0, n - M.size(), '0');
  for (size_t j = n; j-- > 0;) {  // the upper bound is open: one unit less
    if (M[j] == '0') {
      M[j] = '9';
    } else {
      M[j] = (char)(M[j] - 1);
      break;
    }
  }
  size_t diff = 0;
  while (diff < n && L[diff] == M[diff]) ++diff;
  size_t len = n;
  while (len > 1 && !(diff >= len - 1 && M[len - 1] != '0')) --len;
  const int r = M[len - 1] - '0', lr = L[len - 1] - '0';
  std::string D = M.substr(0, len - 1);
  D.push_back((char)('0' + (r - ((r - lr) >> 1))));
  // D x 10^(n - len - FRAC): significant digits and the exponent of d.ddd
  size_t first = 0;
  while (first + 1 < D.size() && D[first] == '0') ++first;
  char digits[40];
  int nd = 0;
  for (size_t j = first; j < D.size() && nd < 39; ++j) digits[nd++] = D[j];
  const int e_sci = (int)(D.size() - first) - 1 + (int)(n - len) - FRAC;
  while (nd > 1 && digits[nd - 1] == '0') --nd;
  digits[nd] = 0;
  char one[32];
  snprintf(one, sizeof one, "%.0e", ax);
  const int expo = atoi(strchr(one, 'e') + 1) + 1;  // the 1-digit value is 0.d x 10^expo
  const char *sg = x < 0 ? "-" : "";
  if (expo >= 0 && expo <= 15) {
    // F form: the digits placed around the point, e_sci + 1 of them before it
    const int point = e_sci + 1;
    char b[64];
    int p = 0;
    if (point <= 0) {
      b[p++] = '.';
      for (int z = 0; z < -point; ++z) b[p++] = '0';
      for (int j = 0; j < nd; ++j) b[p++] = digits[j];
    } else {
      for (int j = 0; j < point; ++j) b[p++] = j < nd ? digits[j] : '0';
      b[p++] = '.';
      for (int j = point; j < nd; ++j) b[p++] = digits[j];
    }
    b[p] = 0;
    snprintf(out, 64, "%s%s", sg, b);
  } else {
    const int ae = abs(e_sci);
    snprintf(out, 64, "%s%c.%sE%c%0*d", sg, digits[0], digits + 1, e_sci < 0 ? '-' : '+', ae > 99 ? 3 : 2, ae);
  }
}

extern "C" int nsx_format_real(double x, int kind, int w, int d, char *out, int outlen) {
  char b[64];
  if (kind == 0)
    fmt_e(b, x, w, d);
  else
    fmt_list_real(b, x);
  if (!out || outlen < (int)strlen(b) + 1) return -1;
  strcpy(out, b);
  return 0;
}

// glob_put_plane_mm's output (stats.F90:301-363): for the interior layers
// k = 2..NZ-1 (1-based), format (I3,4(2x,e12.5)), to
//   sm-NNNNN.dat   k, y+, mean u, v, w        and the closing "TKE =" line
//   srms-NNNNN.dat k, y+, rms u, v, w
//   stau-NNNNN.dat k, y+, u'v', v'w', w'u' (mean products minus products of means)
// planes[q*NZ + k-1] holds S_q(k) of the nine moments (uf vf wf u2f v2f
// w2f uvf vwf wuf), n_cell0 the columns the sums ran over.
extern "C" int nsx_write_plane_stats(const double *planes, int NZ, int n_cell0, int ns, double Re, const char *dir) {
  if (!planes || !dir || NZ < 3 || n_cell0 < 1 || ns < 0 || ns > 99999) return -1;
  char fn[3][1024];
  const char *stem[3] = {"sm-", "srms-", "stau-"};
  FILE *f[3];
  for (int q = 0; q < 3; ++q) {
    snprintf(fn[q], sizeof fn[q], "%s/%s%05d.dat", dir, stem[q], ns);
    f[q] = fopen(fn[q], "w");
    if (!f[q]) {
      for (int r = 0; r < q; ++r) fclose(f[r]);
      return -2;
    }
  }
  const double npjx = 1.0 / (double)n_cell0;
  double tkesum = 0.0;
  for (int k = 2; k <= NZ - 1; ++k) {
    const double yplus = Re * (double)(2 * k - 3) / (double)(2 * NZ - 4);
    const double *S = planes + (k - 1);
    const double u1m = S[0 * NZ] * npjx, v1m = S[1 * NZ] * npjx, w1m = S[2 * NZ] * npjx;
    const double u2m = S[3 * NZ] * npjx, v2m = S[4 * NZ] * npjx, w2m = S[5 * NZ] * npjx;
    const double uvm = S[6 * NZ] * npjx, vwm = S[7 * NZ] * npjx, wum = S[8 * NZ] * npjx;
    const double Tuvm = u1m * v1m - uvm, Tvwm = v1m * w1m - vwm, Twum = w1m * u1m - wum;
    const double xrms = u2m - u1m * u1m, yrms = v2m - v1m * v1m, zrms = w2m - w1m * w1m;
    tkesum = tkesum + 0.5 * (xrms + yrms + zrms);
    const double row[3][3] = {{u1m, v1m, w1m}, {sqrt(xrms), sqrt(yrms), sqrt(zrms)}, {Tuvm, Tvwm, Twum}};
    for (int q = 0; q < 3; ++q) {
      char a[64], b[3][64];
      fmt_e(a, yplus, 12, 5);
      for (int j = 0; j < 3; ++j) fmt_e(b[j], row[q][j], 12, 5);
      fprintf(f[q], "%3d  %s  %s  %s  %s\n", k, a, b[0], b[1], b[2]);
    }
  }
  char t[64];
  fmt_list_real(t, tkesum / (double)(NZ - 2));
  fprintf(f[0], " TKE = %s\n", t);
  int rc = 0;
  for (int q = 0; q < 3; ++q)
    if (fclose(f[q]) != 0) rc = -3;
  return rc;
}
