// nsx_files.hip -- the reference's output files, written on the host from
// arrays in the reference layout: glob_put_plane_mm's statistics files
// (stats.F90:301-363) and the Tecplot snapshots of SavetecVertex and
// SavetecCenter (SavetecVertex.F90:185-235, SavetecCenter.F90).  Fortran
// edit descriptors (Aw, Iw, Fw.d, Ew.d) and list-directed REAL(8) output are
// reproduced as the compiler of the reference build prints them, so the
// files are byte-identical to the reference's for the same values.  Host
// code only: no device is needed.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "nsx_kernels.h"

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
  // fmt_e builds the digits in 32-byte buffers and the line in 64 bytes
  if (kind == 0 && (d < 1 || d > 30 || w < 0 || w > 63)) return -1;
  if (kind != 0 && kind != 1) return -1;
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
  for (int q = 0; q < 3; ++q)
    if (snprintf(fn[q], sizeof fn[q], "%s/%s%05d.dat", dir, stem[q], ns) >= (int)sizeof fn[q]) return -1;
  for (int q = 0; q < 3; ++q) {
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
  for (int q = 0; q < 3; ++q) {
    if (ferror(f[q])) rc = -3;
    if (fclose(f[q]) != 0) rc = -3;
  }
  return rc;
}

// ------------------------------------------------------ Fortran editing ---
// Aw output: w - len blanks then the string, or its first w characters
static void put_a(FILE *f, const char *s, int w) {
  const int n = (int)strlen(s);
  if (n >= w)
    fprintf(f, "%.*s", w, s);
  else
    fprintf(f, "%*s%s", w - n, "", s);
}

// Iw output: right-justified, w asterisks when the value does not fit
static void put_i(FILE *f, long long v, int w) {
  char b[32];
  const int n = snprintf(b, sizeof b, "%lld", v);
  if (n > w)
    for (int j = 0; j < w; ++j) fputc('*', f);
  else
    fprintf(f, "%*s", w, b);
}

// Fw.d output: d decimals, leading zero kept, w asterisks when it does not fit
static void put_f(FILE *f, double x, int w, int d) {
  char b[64];
  const int n = snprintf(b, sizeof b, "%.*f", d, x);
  if (n > w)
    for (int j = 0; j < w; ++j) fputc('*', f);
  else
    fprintf(f, "%*s", w, b);
}

// one write statement of format (8(1x,e12.5)) over n values v[j*stride +
// off]: eight values a record, the last record shorter
struct Block {
  const double *v;
  size_t ld;  // leading dimension of the (rows, cols) array
  int r0, r1, c0, c1;  // rows r0..r1-1 (fastest) and columns c0..c1-1
};

static void put_e_block(FILE *f, const Block &b) {
  int on_line = 0;
  char e[64];
  for (int c = b.c0; c < b.c1; ++c)
    for (int r = b.r0; r < b.r1; ++r) {
      fmt_e(e, b.v[(size_t)r + (size_t)c * b.ld], 12, 5);
      fputc(' ', f);
      fputs(e, f);
      if (++on_line == 8) {
        fputc('\n', f);
        on_line = 0;
      }
    }
  if (on_line) fputc('\n', f);
}

// the prism connectivity of both Tecplot writers: for layers k = k1..k2 of
// the reference, vertex offsets N_VERT*(k-k1) below and N_VERT*(k-k1+1) above,
// format (8(1x,i8)), in SavetecVertex's vertex order
static void put_bricks(FILE *f, const int *No_vp, int N_CELL0, int N_VERT, int nlayers) {
  for (int l = 0; l < nlayers; ++l)
    for (int i = 0; i < N_CELL0; ++i) {
      const long long b0 = (long long)N_VERT * l, t0 = (long long)N_VERT * (l + 1);
      const long long v1 = No_vp[i], v2 = No_vp[i + N_CELL0], v3 = No_vp[i + 2 * (size_t)N_CELL0];
      const long long q[8] = {v1 + b0, v1 + b0, v1 + t0, v1 + t0, v2 + b0, v3 + b0, v3 + t0, v2 + t0};
      for (int j = 0; j < 8; ++j) {
        fputc(' ', f);
        put_i(f, q[j], 8);
      }
      fputc('\n', f);
    }
}

// SavetecVertex (SavetecVertex.F90:185-235): <dir>/V-NNNN.tec of the vertex
// coordinates and vertex fields (N_VERT, NZ-1), counter = SaveCounter
extern "C" int nsx_write_tec_vertex(const char *dir, int counter, double time, int N_CELL0, int N_VERT, int NZ,
                                    const int *No_vp, const double *xvt, const double *yvt, const double *zvt,
                                    const double *ufv, const double *vfv, const double *wfv, const double *pfv) {
  if (!dir || !No_vp || !xvt || !yvt || !zvt || !ufv || !vfv || !wfv || !pfv || NZ < 3 || counter < 0 ||
      counter > 9999)
    return -1;
  char fn[1024];
  if (snprintf(fn, sizeof fn, "%s/V-%04d.tec", dir, counter) >= (int)sizeof fn) return -1;  // path too long
  FILE *f = fopen(fn, "w");
  if (!f) return -2;
  const long long totv = (long long)N_VERT * (NZ - 1), tote = (long long)N_CELL0 * (NZ - 2);
  put_a(f, "TITLE     = \"nsmp 3D prism data\"", 75);
  fputc('\n', f);
  put_a(f, "VARIABLES = \"xv\",\"yv\",\"zv\",\"ufv\",\"vfv\",\"wfv\",\"pfv\"", 75);
  fputc('\n', f);
  put_a(f, "ZONE N = ", 11);
  put_i(f, totv, 7);
  put_a(f, ", E = ", 8);
  put_i(f, tote, 7);
  put_a(f, ", DATAPACKING= BLOCK, ZONETYPE=FEBRICK", 40);
  fputc('\n', f);
  put_a(f, "StrandID = ", 11);
  put_i(f, 1, 3);
  put_a(f, ", SolutionTime = ", 17);
  put_f(f, time, 12, 5);
  fputc('\n', f);
  const double *vv[7] = {xvt, yvt, zvt, ufv, vfv, wfv, pfv};
  for (int q = 0; q < 7; ++q) put_e_block(f, Block{vv[q], (size_t)N_VERT, 0, N_VERT, 0, NZ - 1});
  put_bricks(f, No_vp, N_CELL0, N_VERT, NZ - 2);
  const int bad = ferror(f);  // a failed write along the way
  return (fclose(f) == 0 && !bad) ? 0 : -3;
}

// SavetecCenter (SavetecCenter.F90): <dir>/C-NNNNN.tec of the vertex
// coordinates (N_VERT, NZ-1) and the cell fields (N_CELL, NZ), interior cells
extern "C" int nsx_write_tec_center(const char *dir, int counter, double time, int N_CELL0, int N_CELL, int N_VERT,
                                    int NZ, const int *No_vp, const double *xvt, const double *yvt, const double *zvt,
                                    const double *u, const double *v, const double *w, const double *p) {
  if (!dir || !No_vp || !xvt || !yvt || !zvt || !u || !v || !w || !p || NZ < 3 || counter < 0 || counter > 99999)
    return -1;
  char fn[1024];
  if (snprintf(fn, sizeof fn, "%s/C-%05d.tec", dir, counter) >= (int)sizeof fn) return -1;  // path too long
  FILE *f = fopen(fn, "w");
  if (!f) return -2;
  const long long totv = (long long)N_VERT * (NZ - 1), tote = (long long)N_CELL0 * (NZ - 2);
  put_a(f, "TITLE     = \"nsmp 3D prism data\"", 80);
  fputc('\n', f);
  put_a(f, "VARIABLES = \"xc\",\"yc\",\"zc\",\"uf\",\"vf\",\"wf\",\"pf\"", 80);
  fputc('\n', f);
  put_a(f, "ZONE N = ", 11);
  put_i(f, totv, 8);
  put_a(f, ", E = ", 8);
  put_i(f, tote, 8);
  put_a(f, ", DATAPACKING= BLOCK, ZONETYPE=FEBRICK", 40);
  fputc('\n', f);
  put_a(f, "VARLOCATION=([4,5,6,7]=CELLCENTERED)", 40);
  fputc('\n', f);
  put_a(f, "StrandID = ", 11);
  put_i(f, 1, 3);
  put_a(f, ", SolutionTime = ", 17);
  put_f(f, time, 12, 5);
  fputc('\n', f);
  const double *xyz[3] = {xvt, yvt, zvt};
  for (int q = 0; q < 3; ++q) put_e_block(f, Block{xyz[q], (size_t)N_VERT, 0, N_VERT, 0, NZ - 1});
  const double *cc[4] = {u, v, w, p};
  for (int q = 0; q < 4; ++q) put_e_block(f, Block{cc[q], (size_t)N_CELL, 0, N_CELL0, 1, NZ - 1});
  put_bricks(f, No_vp, N_CELL0, N_VERT, NZ - 2);
  const int bad = ferror(f);  // a failed write along the way
  return (fclose(f) == 0 && !bad) ? 0 : -3;
}

// the list-directed records of the VTK writer as the reference build's
// compiler writes them: a leading blank, character items as they are, each
// integer as a blank and its digits, and a blank between an integer and a
// character item after it
static void put_list(FILE *f, const char *a, int nint, const long long *iv, const char *b) {
  fputc(' ', f);
  fputs(a, f);
  for (int j = 0; j < nint; ++j) fprintf(f, " %lld", iv[j]);
  if (b) {
    if (nint > 0) fputc(' ', f);
    fputs(b, f);
  }
  fputc('\n', f);
}

// one write statement of format (n(1x,e12.5)) per vertex of every layer
// k = 0..nl-1, with the values of up to three (N_VERT, nl) arrays (a NULL
// array writes zeros)
static void put_e_rows(FILE *f, int N_VERT, int nl, int n, const double *a, const double *b, const double *c) {
  const double *v[3] = {a, b, c};
  char e[64];
  for (int k = 0; k < nl; ++k)
    for (int i = 0; i < N_VERT; ++i) {
      for (int q = 0; q < n; ++q) {
        fmt_e(e, v[q] ? v[q][(size_t)i + (size_t)k * N_VERT] : 0.0, 12, 5);
        fputc(' ', f);
        fputs(e, f);
      }
      fputc('\n', f);
    }
}

// SaveParaviewVertex (SaveParaviewVertex.F90:185-278): <dir>/V-NNNN.vtk, an
// unstructured grid of the prisms (0-based vertex numbers) with the vertex
// fields.  The reference labels ufv "pf" and usv "ps"; the solid-phase
// velocities usv, vsv, wsv may be NULL (zero, as in the channel)
extern "C" int nsx_write_vtk_vertex(const char *dir, int counter, int N_CELL0, int N_VERT, int NZ, const int *No_vp,
                                    const double *xvt, const double *yvt, const double *zvt, const double *ufv,
                                    const double *vfv, const double *wfv, const double *usv, const double *vsv,
                                    const double *wsv) {
  if (!dir || !No_vp || !xvt || !yvt || !zvt || !ufv || !vfv || !wfv || NZ < 3 || counter < 0 || counter > 9999)
    return -1;
  char fn[1024];
  if (snprintf(fn, sizeof fn, "%s/V-%04d.vtk", dir, counter) >= (int)sizeof fn) return -1;  // path too long
  FILE *f = fopen(fn, "w");
  if (!f) return -2;
  const long long totv = (long long)N_VERT * (NZ - 1), tote = (long long)N_CELL0 * (NZ - 2);
  fputs("# vtk DataFile Version 2.0\nUnstructured Grid Example\nASCII\nDATASET UNSTRUCTURED_GRID\n", f);
  put_list(f, "POINTS ", 1, &totv, " float");
  put_e_rows(f, N_VERT, NZ - 1, 3, xvt, yvt, zvt);
  const long long ce[2] = {tote, 7 * tote};
  put_list(f, "CELLS ", 2, ce, nullptr);
  for (int l = 0; l < NZ - 2; ++l)
    for (int i = 0; i < N_CELL0; ++i) {
      const long long b0 = (long long)N_VERT * l - 1, t0 = (long long)N_VERT * (l + 1) - 1;
      const long long v1 = No_vp[i], v2 = No_vp[i + N_CELL0], v3 = No_vp[i + 2 * (size_t)N_CELL0];
      const long long q[7] = {6, v1 + b0, v2 + b0, v3 + b0, v1 + t0, v2 + t0, v3 + t0};
      for (int j = 0; j < 7; ++j) {
        fputc(' ', f);
        put_i(f, q[j], 8);
      }
      fputc('\n', f);
    }
  put_list(f, "CELL_TYPES ", 1, &tote, nullptr);
  for (long long e = 0; e < tote; ++e) fputs("       13\n", f);
  put_list(f, " ", 0, nullptr, nullptr);
  put_list(f, "POINT_DATA ", 1, &totv, nullptr);
  put_list(f, "SCALARS pf     float", 0, nullptr, nullptr);
  put_list(f, "LOOKUP_TABLE default", 0, nullptr, nullptr);
  put_e_rows(f, N_VERT, NZ - 1, 1, ufv, nullptr, nullptr);
  put_list(f, " ", 0, nullptr, nullptr);
  put_list(f, "SCALARS ps     float", 0, nullptr, nullptr);
  put_list(f, "LOOKUP_TABLE default", 0, nullptr, nullptr);
  put_e_rows(f, N_VERT, NZ - 1, 1, usv, nullptr, nullptr);
  put_list(f, " ", 0, nullptr, nullptr);
  put_list(f, "VECTORS velocity_f float", 0, nullptr, nullptr);
  put_e_rows(f, N_VERT, NZ - 1, 3, ufv, vfv, wfv);
  put_list(f, " ", 0, nullptr, nullptr);
  put_list(f, "VECTORS velocity_s float", 0, nullptr, nullptr);
  put_e_rows(f, N_VERT, NZ - 1, 3, usv, vsv, wsv);
  const int bad = ferror(f);  // a failed write along the way
  return (fclose(f) == 0 && !bad) ? 0 : -3;
}

// SavetecVC (SavetecVertexCenter.F90:212-305): <dir>/VC-NNNNN.tec, the
// vertices (N_VERT, NZ-1) followed by the cell centres of the interior
// layers as nodes, each prism split into five bricks around its centre node
extern "C" int nsx_write_tec_vc(const char *dir, int counter, double time, int N_CELL0, int N_CELL, int N_VERT, int NZ,
                                const int *No_vp, const double *xvt, const double *yvt, const double *zvt,
                                const double *xct, const double *yct, const double *zct, const double *ufv,
                                const double *vfv, const double *wfv, const double *pfv, const double *u,
                                const double *v, const double *w, const double *p) {
  if (!dir || !No_vp || !xvt || !yvt || !zvt || !xct || !yct || !zct || !ufv || !vfv || !wfv || !pfv || !u || !v ||
      !w || !p || NZ < 3 || counter < 0 || counter > 99999)
    return -1;
  char fn[1024];
  if (snprintf(fn, sizeof fn, "%s/VC-%05d.tec", dir, counter) >= (int)sizeof fn) return -1;  // path too long
  FILE *f = fopen(fn, "w");
  if (!f) return -2;
  const long long nodes_v = (long long)N_VERT * (NZ - 1), tote = (long long)N_CELL0 * (NZ - 2);
  put_a(f, "TITLE     = \"nsmp 3D prism data\"", 70);
  fputc('\n', f);
  put_a(f, "VARIABLES = \"x\",\"y\",\"z\"", 70);
  fputc('\n', f);
  put_a(f, "\"uf\",\"vf\",\"wf\",\"pf\"", 70);
  fputc('\n', f);
  put_a(f, "ZONE N = ", 11);
  put_i(f, nodes_v + tote, 7);
  put_a(f, ", E = ", 8);
  put_i(f, 5 * tote, 7);
  put_a(f, ", DATAPACKING= BLOCK, ZONETYPE=FEBRICK", 40);
  fputc('\n', f);
  put_a(f, "StrandID = ", 11);
  put_i(f, 1, 3);
  put_a(f, ", SolutionTime = ", 17);
  put_f(f, time, 12, 5);
  fputc('\n', f);
  const double *vv[7] = {xvt, yvt, zvt, ufv, vfv, wfv, pfv};
  const double *cc[7] = {xct, yct, zct, u, v, w, p};
  for (int q = 0; q < 7; ++q) {
    put_e_block(f, Block{vv[q], (size_t)N_VERT, 0, N_VERT, 0, NZ - 1});
    put_e_block(f, Block{cc[q], (size_t)N_CELL, 0, N_CELL0, 1, NZ - 1});
  }
  for (int l = 0; l < NZ - 2; ++l)
    for (int i = 0; i < N_CELL0; ++i) {
      const long long b0 = (long long)N_VERT * l, t0 = (long long)N_VERT * (l + 1);
      const long long v1 = No_vp[i], v2 = No_vp[i + N_CELL0], v3 = No_vp[i + 2 * (size_t)N_CELL0];
      const long long c = (long long)(i + 1) + (long long)N_CELL0 * l + nodes_v;
      const long long q[5][8] = {{v1 + b0, v1 + b0, c, c, v2 + b0, v3 + b0, c, c},
                                 {c, c, v1 + t0, v1 + t0, c, c, v3 + t0, v2 + t0},
                                 {c, c, c, c, v1 + b0, v2 + b0, v2 + t0, v1 + t0},
                                 {c, c, c, c, v2 + b0, v3 + b0, v3 + t0, v2 + t0},
                                 {c, c, c, c, v3 + b0, v1 + b0, v1 + t0, v3 + t0}};
      for (int r = 0; r < 5; ++r) {
        for (int j = 0; j < 8; ++j) {
          fputc(' ', f);
          put_i(f, q[r][j], 8);
        }
        fputc('\n', f);
      }
    }
  const int bad = ferror(f);  // a failed write along the way
  return (fclose(f) == 0 && !bad) ? 0 : -3;
}
