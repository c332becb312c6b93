/*
 * nsx_case.c -- host-side set-up for the prismatic sigma-grid channel.
 *
 * Builds everything the per-step hot path consumes, in the reference's own
 * array layout (column-major, 1-based connectivity, ghost cells appended after
 * N_CELL0):
 *   - a seeded synthetic mesh: nx*ny quads on [0,Lx]x[0,Ly], each split into
 *     two counter-clockwise triangles along a diagonal drawn per quad from the
 *     seed, optional jitter of interior vertices and an (inert) bed;
 *   - tanh-clustered sigma levels;
 *   - the seeded initial condition (Poiseuille profile + uniform noise);
 *   - the geometry of src/Main/calcul_geometry.F90 (centroids, areas, normals,
 *     periodic ghost cells, IDW weights, Kim-Choi / LSM coefficients);
 *   - the boundary flags of src/BoundaryConditions/update_bc.F90 (adjust_tagv,
 *     adjust_bc) for the doubly periodic box with sigma walls;
 *   - the stored PPE matrix of src/AdvectionDiffusion/PDE.F90 (PDE_center).
 *
 * The geometry routines follow the reference loop by loop and operation by
 * operation (no FMA contraction: build with -ffp-contract=off) so that every
 * array is bit-identical to what the Fortran set-up produces for the same
 * vertices; tests check that against the compiled reference.
 *
 * The box is periodic in x and y (XPB=YPB=1, the reference's channel test)
 * or walled on either pair of sides (XPB or YPB = 0, wall kinds XBC, YBC):
 * a wall ghost keeps the centre mirrored across its edge and gets the flags
 * of update_bc.F90 (wall 0, free slip 1, inflow 2).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "nsx_case.h"

#define IX2(i, j, n) ((size_t)(i) + (size_t)(j) * (size_t)(n)) /* 0-based (i,j) of an (n,*) array */

/* ---------------------------------------------------------------- RNG ---- */
static uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

double nsx_rng_u01(uint64_t seed, uint32_t stream, uint64_t index) {
  uint64_t h = mix64(mix64(seed ^ ((uint64_t)stream << 48)) ^ index);
  return (double)(h >> 11) * 0x1.0p-53;
}

/* ------------------------------------------------------------ helpers ---- */
static void *xcalloc(size_t n, size_t sz) {
  void *p = calloc(n ? n : 1, sz);
  if (!p) {
    fprintf(stderr, "nsx_case: out of memory (%zu x %zu)\n", n, sz);
    abort();
  }
  return p;
}

/* Fortran single-precision literals used in comparisons (1.0E-7 etc.). */
#define F32_1EM7 ((double)1.0e-7f)

typedef struct {
  long key;
  int tri, j;
} edge_rec;

static int edge_cmp(const void *a, const void *b) {
  long ka = ((const edge_rec *)a)->key, kb = ((const edge_rec *)b)->key;
  return (ka > kb) - (ka < kb);
}

/* ------------------------------------------------------------- mesh ------ */
static int build_mesh(nsx_case *c) {
  const int nx = c->nx, ny = c->ny;
  const int N0 = c->N_CELL0, NV = c->N_VERT;
  const double dx = c->Lx / nx, dy = c->Ly / ny;

  for (int jy = 0; jy <= ny; ++jy)
    for (int ix = 0; ix <= nx; ++ix) {
      int v = jy * (nx + 1) + ix;
      double x = (ix == 0) ? 0.0 : (ix == nx ? c->Lx : c->Lx * (double)ix / (double)nx);
      double y = (jy == 0) ? 0.0 : (jy == ny ? c->Ly : c->Ly * (double)jy / (double)ny);
      int bnd = (ix == 0 || ix == nx || jy == 0 || jy == ny);
      if (!bnd && c->jitter > 0.0) {
        x += (2.0 * nsx_rng_u01(c->seed, NSX_STREAM_JITTER_X, (uint64_t)v) - 1.0) * c->jitter * dx;
        y += (2.0 * nsx_rng_u01(c->seed, NSX_STREAM_JITTER_Y, (uint64_t)v) - 1.0) * c->jitter * dy;
      }
      c->xv[v] = x;
      c->yv[v] = y;
      c->nbev[v] = bnd ? 1 : 0;
      c->zbv[v] = c->bed ? 0.05 * c->H * sin(2.0 * M_PI * x / c->Lx) * cos(2.0 * M_PI * y / c->Ly) : 0.0;
    }

  for (int jy = 0; jy < ny; ++jy)
    for (int ix = 0; ix < nx; ++ix) {
      int q = jy * nx + ix;
      int v00 = jy * (nx + 1) + ix, v10 = v00 + 1, v01 = v00 + (nx + 1), v11 = v01 + 1;
      int d = nsx_rng_u01(c->seed, NSX_STREAM_DIAG, (uint64_t)q) < 0.5 ? 0 : 1;
      int t0 = 2 * q, t1 = 2 * q + 1;
      int a[3], b[3];
      if (d == 0) { /* diagonal v00-v11 */
        a[0] = v00; a[1] = v10; a[2] = v11;
        b[0] = v00; b[1] = v11; b[2] = v01;
      } else {      /* diagonal v10-v01 */
        a[0] = v00; a[1] = v10; a[2] = v01;
        b[0] = v10; b[1] = v11; b[2] = v01;
      }
      for (int j = 0; j < 3; ++j) {
        c->No_vp[IX2(t0, j, N0)] = a[j] + 1;
        c->No_vp[IX2(t1, j, N0)] = b[j] + 1;
      }
    }

  /* neighbour across edge j = (No_vp(i,j), No_vp(i,j+1)) */
  edge_rec *e = xcalloc((size_t)3 * N0, sizeof(edge_rec));
  for (int i = 0; i < N0; ++i)
    for (int j = 0; j < 3; ++j) {
      int va = c->No_vp[IX2(i, j, N0)] - 1, vb = c->No_vp[IX2(i, (j + 1) % 3, N0)] - 1;
      int lo = va < vb ? va : vb, hi = va < vb ? vb : va;
      e[3 * i + j].key = (long)lo * NV + hi;
      e[3 * i + j].tri = i;
      e[3 * i + j].j = j;
    }
  qsort(e, (size_t)3 * N0, sizeof(edge_rec), edge_cmp);
  for (int i = 0; i < N0; ++i) c->nbe[i] = 0;
  for (size_t n = 0; n < (size_t)3 * N0;) {
    size_t m = n + 1;
    while (m < (size_t)3 * N0 && e[m].key == e[n].key) ++m;
    if (m - n == 2) {
      c->No_cp[IX2(e[n].tri, e[n].j, c->N_CELL)] = e[n + 1].tri + 1;
      c->No_cp[IX2(e[n + 1].tri, e[n + 1].j, c->N_CELL)] = e[n].tri + 1;
    } else if (m - n == 1) {
      c->No_cp[IX2(e[n].tri, e[n].j, c->N_CELL)] = 0;
      c->nbe[e[n].tri] = 1;
    } else {
      free(e);
      snprintf(c->err, sizeof c->err, "non-manifold edge in generated mesh");
      return -1;
    }
    n = m;
  }
  free(e);
  return 0;
}

/* --------------------------------------------------------- sigma levels -- */
static void build_sigma(nsx_case *c) {
  const int NZ = c->NZ, L = c->NN;
  double *sigv = c->sigv, *dsigv = c->dsigv, *sig = c->sig, *dsig = c->dsig;
  /* vertex (interface) levels z_k = (H/2)(1 + tanh(g(2k/L-1))/tanh g), k=0..L */
  for (int k = 0; k <= L; ++k) {
    double z;
    if (k == 0) z = 0.0;
    else if (k == L) z = c->H;
    else if (c->stretch > 0.0)
      z = 0.5 * c->H * (1.0 + tanh(c->stretch * (2.0 * k / (double)L - 1.0)) / tanh(c->stretch));
    else
      z = c->H * (double)k / (double)L;
    sigv[k] = z; /* sigv(k+1) */
  }
  /* thicknesses and cell centres with the recurrence of calcul_geometry.F90:263-283 */
  for (int k = 0; k < NZ - 2; ++k) dsigv[k] = sigv[k + 1] - sigv[k];
  dsigv[NZ - 2] = dsigv[NZ - 3]; /* top ghost cell mirrors the top cell */
  dsig[0] = dsigv[0];
  sig[0] = sigv[0] - 0.5 * dsigv[0];
  for (int k = 1; k < NZ - 1; ++k) {
    dsig[k] = 0.5 * (dsigv[k] + dsigv[k - 1]);
    sig[k] = sig[k - 1] + dsig[k - 1];
  }
  dsig[NZ - 1] = dsigv[NZ - 2];
  sig[NZ - 1] = sig[NZ - 2] + dsig[NZ - 2];
}

/* ---------------------------------------------------- search indexes --- */
/* The reference finds periodic partners by scanning every vertex or cell for
 * each ghost cell and boundary vertex (calcul_geometry.F90:484-562,
 * 920-1025).  Here the boundary vertices sit in a uniform bucket grid (cell
 * size h >= 1e-6, so every vertex within 3e-7 of a point lies in its 3x3
 * buckets) and the boundary cells in a map from edges to cells.  Candidates
 * are tested with the reference's own conditions in increasing index order,
 * so the matches, their counts and the match kept are those of the scans. */
typedef struct {
  double x0, y0, h;
  long nbx, nby;
  long *key; /* sorted bucket keys, with the vertex of each */
  int *idx;
  int n;
} vgrid;

typedef struct {
  long key;
  int v;
} kv_rec;

static int kv_cmp(const void *a, const void *b) {
  const kv_rec *p = a, *q = b;
  if (p->key != q->key) return (p->key > q->key) - (p->key < q->key);
  return (p->v > q->v) - (p->v < q->v);
}

static int int_cmp(const void *a, const void *b) {
  int p = *(const int *)a, q = *(const int *)b;
  return (p > q) - (p < q);
}

static void vgrid_build(vgrid *g, const double *xv, const double *yv, const int *nbev, int NV) {
  double xmin = 0, xmax = 0, ymin = 0, ymax = 0;
  int nb = 0;
  for (int v = 0; v < NV; ++v) {
    if (nbev[v] == 0) continue;
    if (nb == 0 || xv[v] < xmin) xmin = xv[v];
    if (nb == 0 || xv[v] > xmax) xmax = xv[v];
    if (nb == 0 || yv[v] < ymin) ymin = yv[v];
    if (nb == 0 || yv[v] > ymax) ymax = yv[v];
    ++nb;
  }
  double h = nb > 0 ? 2.0 * ((xmax - xmin) + (ymax - ymin)) / nb : 1.0;
  if (!(h >= 1.0e-6)) h = 1.0e-6;
  g->h = h;
  g->x0 = xmin - 2.0 * h;
  g->y0 = ymin - 2.0 * h;
  g->nbx = (long)((xmax - g->x0) / h) + 3;
  g->nby = (long)((ymax - g->y0) / h) + 3;
  kv_rec *r = xcalloc((size_t)nb, sizeof(kv_rec));
  int m = 0;
  for (int v = 0; v < NV; ++v) {
    if (nbev[v] == 0) continue;
    const long bx = (long)floor((xv[v] - g->x0) / h), by = (long)floor((yv[v] - g->y0) / h);
    r[m].key = bx * g->nby + by;
    r[m].v = v;
    ++m;
  }
  qsort(r, (size_t)nb, sizeof(kv_rec), kv_cmp);
  g->key = xcalloc((size_t)nb, sizeof(long));
  g->idx = xcalloc((size_t)nb, sizeof(int));
  for (int j = 0; j < nb; ++j) {
    g->key[j] = r[j].key;
    g->idx[j] = r[j].v;
  }
  g->n = nb;
  free(r);
}

static void vgrid_free(vgrid *g) {
  free(g->key);
  free(g->idx);
}

/* the boundary vertices in the 3x3 buckets around (x, y), ascending and
 * without repeats, appended to out[*n] (room for cap) */
static void vgrid_near(const vgrid *g, double x, double y, int *out, int *n, int cap) {
  const double fx = floor((x - g->x0) / g->h), fy = floor((y - g->y0) / g->h);
  if (!(fx > -2.0 && fx < (double)g->nbx + 1.0 && fy > -2.0 && fy < (double)g->nby + 1.0)) return;
  const long bx = (long)fx, by = (long)fy;
  for (long dx = -1; dx <= 1; ++dx)
    for (long dy = -1; dy <= 1; ++dy) {
      const long cx = bx + dx, cy = by + dy;
      if (cx < 0 || cx >= g->nbx || cy < 0 || cy >= g->nby) continue;
      const long k = cx * g->nby + cy;
      int lo = 0, hi = g->n;
      while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (g->key[mid] < k) lo = mid + 1;
        else hi = mid;
      }
      for (int j = lo; j < g->n && g->key[j] == k && *n < cap; ++j) out[(*n)++] = g->idx[j];
    }
  qsort(out, (size_t)*n, sizeof(int), int_cmp);
  int m = 0;
  for (int j = 0; j < *n; ++j)
    if (m == 0 || out[j] != out[m - 1]) out[m++] = out[j];
  *n = m;
}

/* ------------------------------------------------------------ geometry --- */
/* calcul_geometry.F90:290-367 (per-cell), 402-578 (ghost cells, sequential
 * periodic branch), 851-1040 (tags, vertex pairs), 1055-1076 (IDW weights),
 * 1343-1517 (Kim-Choi / LSM coefficients). */
static int build_geometry(nsx_case *c) {
  const int N0 = c->N_CELL0, NV = c->N_VERT, NC = c->N_CELL;
  int *No_cp = c->No_cp;
  const int *No_vp = c->No_vp, *nbe = c->nbe, *nbev = c->nbev;
  const double *xv = c->xv, *yv = c->yv;
  double *xc = c->xc, *yc = c->yc, *area = c->areaCell;
  double *dxVV = c->dxVV, *dyVV = c->dyVV, *dlVV = c->dlVV, *xnn = c->xnn, *ynn = c->ynn;
#define VP(i, j) (No_vp[IX2(i, j, N0)] - 1)
#define CP(i, j) No_cp[IX2(i, j, NC)]

  double Volsum = 0.;
  for (int i = 0; i < N0; ++i) {
    int jv1 = VP(i, 0), jv2 = VP(i, 1), jv3 = VP(i, 2);
    xc[i] = (xv[jv1] + xv[jv2] + xv[jv3]) / 3.0;
    yc[i] = (yv[jv1] + yv[jv2] + yv[jv3]) / 3.0;
    area[i] = 0.5 * fabs(xv[jv1] * (yv[jv2] - yv[jv3]) + xv[jv2] * (yv[jv3] - yv[jv1]) +
                         xv[jv3] * (yv[jv1] - yv[jv2]));
    dxVV[IX2(i, 0, N0)] = xv[jv2] - xv[jv1];
    dxVV[IX2(i, 1, N0)] = xv[jv3] - xv[jv2];
    dxVV[IX2(i, 2, N0)] = xv[jv1] - xv[jv3];
    dyVV[IX2(i, 0, N0)] = yv[jv2] - yv[jv1];
    dyVV[IX2(i, 1, N0)] = yv[jv3] - yv[jv2];
    dyVV[IX2(i, 2, N0)] = yv[jv1] - yv[jv3];
    for (int j = 0; j < 3; ++j) {
      double a = dxVV[IX2(i, j, N0)], b = dyVV[IX2(i, j, N0)];
      dlVV[IX2(i, j, N0)] = sqrt(a * a + b * b);
    }
    for (int j = 0; j < 3; ++j) xnn[IX2(i, j, N0)] = dyVV[IX2(i, j, N0)] / dlVV[IX2(i, j, N0)];
    for (int j = 0; j < 3; ++j) ynn[IX2(i, j, N0)] = -dxVV[IX2(i, j, N0)] / dlVV[IX2(i, j, N0)];
    Volsum = Volsum + area[i];
  }
  c->TotalVolume = Volsum;

  /* ghost cells (sequential build, KeyTESTpBC) */
  const double XDIni = c->XDIni, XDFin = c->XDFin, YDIni = c->YDIni, YDFin = c->YDFin;
  vgrid vg;
  vgrid_build(&vg, xv, yv, nbev, NV);
  const int cap = 4 * vg.n + 16;
  int *cand = xcalloc((size_t)cap, sizeof(int));
  /* edges of the boundary cells, (lower, higher vertex) -> cell, ascending */
  int nem = 0;
  kv_rec *em = xcalloc((size_t)3 * N0 + 1, sizeof(kv_rec));
  for (int i0 = 0; i0 < N0; ++i0) {
    if (nbe[i0] == 0) continue;
    for (int j = 0; j < 3; ++j) {
      const long a1 = VP(i0, j), a2 = VP(i0, (j + 1) % 3);
      em[nem].key = (a1 < a2 ? a1 : a2) * (long)NV + (a1 < a2 ? a2 : a1);
      em[nem].v = i0;
      ++nem;
    }
  }
  qsort(em, (size_t)nem, sizeof(kv_rec), kv_cmp);
  int gerr = 0;
  int ii = N0;
  for (int i = 0; i < N0; ++i) {
    if (nbe[i] <= 0) continue;
    for (int j = 0; j < 3; ++j) {
      int nc = CP(i, j);
      if (!(nc < 1 || nc > N0)) continue;
      if (ii >= NC) {
        snprintf(c->err, sizeof c->err, "more ghost cells than N_CELL allows");
        gerr = -1;
        goto done;
      }
      int g = ii; /* 0-based ghost row */
      ii = ii + 1;
      CP(i, j) = ii;
      CP(g, 0) = i + 1;
      CP(g, 1) = j + 1;
      area[g] = area[i];
      int jj = (j + 1) % 3;
      int jv1 = VP(i, j), jv2 = VP(i, jj);
      double xee, yee;
      if (fabs(dyVV[IX2(i, j, N0)]) < F32_1EM7) {
        yee = yv[jv2];
        xee = xc[i];
      } else {
        double dxody = dxVV[IX2(i, j, N0)] / dyVV[IX2(i, j, N0)];
        double dxody2 = dxody * dxody;
        yee = (yc[i] - (xv[jv2] - xc[i]) * dxody + yv[jv2] * dxody2) / (1.0 + dxody2);
        xee = xv[jv2] + (yee - yv[jv2]) * dxody;
      }
      xc[g] = 2.0 * xee - xc[i];
      yc[g] = 2.0 * yee - yc[i];
      CP(g, 2) = -1;

      double dxl = XDFin - XDIni, dyl = YDFin - YDIni;
      if (c->XPB == 1) {
        if (xc[g] > XDFin) {
          dxl = -1.0 * dxl;
          dyl = 0.;
        } else if (xc[g] < XDIni) {
          dyl = 0.;
        }
      }
      if (c->YPB == 1) {
        if (yc[g] > YDFin) {
          dxl = 0.;
          dyl = -1.0 * dyl;
        } else if (yc[g] < YDIni) {
          dxl = 0.;
        }
      }
      /* not periodic: the mirrored centre stays, No_cp(g,3) = -1 */
      if (dxl * dyl > F32_1EM7) continue;
      int jv1p = -1, jv2p = -1, ncount = 0, ncand = 0;
      /* the boundary vertex at each end of the edge shifted across the box */
      xee = xv[jv1] + dxl;
      yee = yv[jv1] + dyl;
      vgrid_near(&vg, xee, yee, cand, &ncand, cap);
      for (int q = 0; q < ncand; ++q) {
        const int nv = cand[q];
        if (fabs(xee - xv[nv]) < F32_1EM7 && fabs(yee - yv[nv]) < F32_1EM7) {
          jv1p = nv;
          ++ncount;
        }
      }
      if (ncount != 1) goto pairfail;
      ncount = 0;
      ncand = 0;
      xee = xv[jv2] + dxl;
      yee = yv[jv2] + dyl;
      vgrid_near(&vg, xee, yee, cand, &ncand, cap);
      for (int q = 0; q < ncand; ++q) {
        const int nv = cand[q];
        if (fabs(xee - xv[nv]) < F32_1EM7 && fabs(yee - yv[nv]) < F32_1EM7) {
          jv2p = nv;
          ++ncount;
        }
      }
      if (ncount != 1) goto pairfail;
      {
        /* partner cell: the boundary triangle holding edge (jv1p, jv2p); the
         * reference tests products of 1-based vertex numbers, which hold
         * only for the cells having both vertices: those of the edge map */
        long p1 = (long)(jv1p + 1) * (jv2p + 1);
        int cc = -1;
        ncount = 0;
        const long ek = (long)(jv1p < jv2p ? jv1p : jv2p) * NV + (jv1p < jv2p ? jv2p : jv1p);
        int lo = 0, hi = nem;
        while (lo < hi) {
          const int mid = (lo + hi) / 2;
          if (em[mid].key < ek) lo = mid + 1;
          else hi = mid;
        }
        for (int q = lo; q < nem && em[q].key == ek; ++q) {
          const int i0 = em[q].v;
          if (q > lo && em[q - 1].v == i0) continue;
          long a10 = VP(i0, 0) + 1, a20 = VP(i0, 1) + 1, a30 = VP(i0, 2) + 1;
          if (a10 * a20 == p1) {
            if (a10 == jv1p + 1 || a10 == jv2p + 1) { cc = (int)a30 - 1; CP(g, 2) = i0 + 1; ++ncount; }
          }
          if (a30 * a20 == p1) {
            if (a30 == jv1p + 1 || a30 == jv2p + 1) { cc = (int)a10 - 1; CP(g, 2) = i0 + 1; ++ncount; }
          }
          if (a10 * a30 == p1) {
            if (a10 == jv1p + 1 || a10 == jv2p + 1) { cc = (int)a20 - 1; CP(g, 2) = i0 + 1; ++ncount; }
          }
        }
        if (ncount != 1) goto pairfail;
        xee = xv[cc] - dxl;
        yee = yv[cc] - dyl;
        xc[g] = (xv[jv1] + xv[jv2] + xee) / 3.0;
        yc[g] = (yv[jv1] + yv[jv2] + yee) / 3.0;
      }
      continue;
    pairfail:
      snprintf(c->err, sizeof c->err, "periodic matching failed for cell %d edge %d", i + 1, j + 1);
      gerr = -1;
      goto done;
    }
  }
done:
  free(em);
  if (gerr) {
    free(cand);
    vgrid_free(&vg);
    return gerr;
  }
  c->N_CELLexact = ii;
  c->N_CELLghost = ii - N0;
  /* every edge of an interior cell now has a cell behind it (a cell that the
   * mesh does not mark as a boundary cell keeps its out-of-range entries) */
  for (int i = 0; i < N0; ++i)
    for (int j = 0; j < 3; ++j)
      if (CP(i, j) < 1 || CP(i, j) > ii) {
        snprintf(c->err, sizeof c->err, "cell %d edge %d has no neighbour (No_cp = %d, nbe = %d)", i + 1, j + 1,
                 CP(i, j), nbe[i]);
        free(cand);
        vgrid_free(&vg);
        return -1;
      }

  /* vertex tags (calcul_geometry.F90:895-918) and periodic pairs (:920-1025) */
  int *tagXV = c->tagXV, *tagYV = c->tagYV;
  for (int nv = 0; nv < NV; ++nv) {
    tagXV[nv] = 0;
    tagYV[nv] = 0;
    if (nbev[nv] != 0) {
      if (fabs(xv[nv] - XDIni) < 1.0e-7) tagXV[nv] = 1;
      if (fabs(xv[nv] - XDFin) < 1.0e-7) tagXV[nv] = 2;
      if (fabs(yv[nv] - YDIni) < 1.0e-7) tagYV[nv] = 1;
      if (fabs(yv[nv] - YDFin) < 1.0e-7) tagYV[nv] = 2;
    }
  }
  for (int nv = 0; nv < NV; ++nv) {
    c->vertex_pair[nv] = -1;
    c->vertex_edge[nv] = -1;
  }
  /* with every boundary vertex inside the box (to 1e-7), a partner across
   * it lies within 2e-7 of the point shifted by the box length, so the
   * buckets around the four shifted points hold every candidate */
  int inbox = 1;
  for (int nv = 0; nv < NV && inbox; ++nv)
    if (nbev[nv] != 0 && !(xv[nv] >= XDIni - 1.0e-7 && xv[nv] <= XDFin + 1.0e-7 && yv[nv] >= YDIni - 1.0e-7 &&
                           yv[nv] <= YDFin + 1.0e-7))
      inbox = 0;
  for (int nv = 0; nv < NV; ++nv) {
    double xee = xv[nv], yee = yv[nv];
    int ncount = (tagXV[nv] * tagYV[nv] > 0) ? 2 : ((tagXV[nv] > 0 || tagYV[nv] > 0) ? 1 : 0);
    if (ncount == 1 || (ncount == 2 && c->XPB * c->YPB != 1)) {
      /* a boundary vertex, or a corner of a box periodic in one direction
       * only (vertex_edge 0, its pair searched the same way) */
      if (ncount == 2) c->vertex_edge[nv] = 0;
      int vcount = 0, ncand = 0;
      if (inbox) {
        const double Lx = XDFin - XDIni, Ly = YDFin - YDIni;
        vgrid_near(&vg, xee, yee + Ly, cand, &ncand, cap);
        vgrid_near(&vg, xee, yee - Ly, cand, &ncand, cap);
        vgrid_near(&vg, xee + Lx, yee, cand, &ncand, cap);
        vgrid_near(&vg, xee - Lx, yee, cand, &ncand, cap);
      }
      const int nscan = inbox ? ncand : NV;
      for (int q = 0; q < nscan; ++q) {
        const int i = inbox ? cand[q] : q;
        if (nbev[i] == 0) continue;
        if (c->YPB == 1 && fabs(xee - xv[i]) < 1.0e-7 && fabs(yee - yv[i]) > (YDFin - YDIni - 1.0e-7)) {
          ++vcount;
          c->vertex_pair[nv] = i + 1;
        }
        if (c->XPB == 1 && fabs(yee - yv[i]) < 1.0e-7 && fabs(xee - xv[i]) > (XDFin - XDIni - 1.0e-7)) {
          ++vcount;
          c->vertex_pair[nv] = i + 1;
        }
      }
      if (vcount == 0) c->vertex_pair[nv] = -1;
      else if (vcount != 1) {
        snprintf(c->err, sizeof c->err, "vertex %d: %d periodic partners", nv + 1, vcount);
        free(cand);
        vgrid_free(&vg);
        return -1;
      }
    }
    else if (ncount == 2) c->vertex_edge[nv] = 1; /* XPB*YPB = 1 */
  }
  free(cand);
  vgrid_free(&vg);

  /* inverse-distance weights, interior cells only (calcul_geometry.F90:1055-1076) */
  for (int nv = 0; nv < NV; ++nv) {
    c->Dimsurrounding[nv] = 0;
    c->dlVsum[nv] = 0.0;
  }
  for (int i = 0; i < N0; ++i)
    for (int j = 0; j < 3; ++j) {
      int nv = VP(i, j);
      int j0 = c->Dimsurrounding[nv];
      if (j0 >= NSX_MAX_SURROUND) {
        snprintf(c->err, sizeof c->err, "vertex %d touches more than %d cells", nv + 1, NSX_MAX_SURROUND);
        return -1;
      }
      double dxCV = xc[i] - xv[nv], dyCV = yc[i] - yv[nv];
      c->surrounding[IX2(nv, j0, NV)] = i + 1;
      c->weight[IX2(nv, j0, NV)] = 1.0 / sqrt(dxCV * dxCV + dyCV * dyCV);
      c->dlVsum[nv] = c->dlVsum[nv] + c->weight[IX2(nv, j0, NV)];
      c->Dimsurrounding[nv] = j0 + 1;
    }

  /* edge points and Kim-Choi offsets, EdgePointOption = 1 (:1346-1387) */
  for (int i = 0; i < N0; ++i)
    for (int j = 0; j < 3; ++j) {
      int jj = (j + 1) % 3;
      int jv1 = VP(i, j), jv2 = VP(i, jj);
      size_t f = IX2(i, j, N0);
      c->xme[f] = 0.5 * (xv[jv2] + xv[jv1]);
      c->yme[f] = 0.5 * (yv[jv2] + yv[jv1]);
      int jc = CP(i, j) - 1;
      double xm1 = xc[i], xm2 = xc[jc], xm3 = xv[jv1], xm4 = xv[jv2];
      double ym1 = yc[i], ym2 = yc[jc], ym3 = yv[jv1], ym4 = yv[jv2];
      double den = ((xm1 - xm2) * (ym3 - ym4) - (ym1 - ym2) * (xm3 - xm4));
      c->xmo[f] = ((xm1 * ym2 - ym1 * xm2) * (xm3 - xm4) - (xm1 - xm2) * (xm3 * ym4 - ym3 * xm4)) / den;
      c->ymo[f] = ((xm1 * ym2 - ym1 * xm2) * (ym3 - ym4) - (ym1 - ym2) * (xm3 * ym4 - ym3 * xm4)) / den;
      xm1 = c->xme[f] - c->xmo[f];
      ym1 = c->yme[f] - c->ymo[f];
      c->doe[f] = sqrt(xm1 * xm1 + ym1 * ym1);
      xm2 = xm1 * dxVV[f] + ym1 * dyVV[f];
      if (xm2 < 0.0) c->doe[f] = -c->doe[f];
    }

  /* LSM sums and Kim-Choi decomposition (:1460-1517, KeyKimChoi) */
  for (int i = 0; i < N0; ++i) {
    double sumX = 0.0, sumY = 0.0, sumXY = 0.0;
    for (int j = 0; j < 3; ++j) {
      size_t f = IX2(i, j, N0);
      int jc = CP(i, j) - 1;
      c->dxCC[f] = xc[jc] - xc[i];
      c->dyCC[f] = yc[jc] - yc[i];
      c->dlCC[f] = sqrt(c->dxCC[f] * c->dxCC[f] + c->dyCC[f] * c->dyCC[f]);
      sumX = sumX + c->dxCC[f] * c->dxCC[f];
      sumY = sumY + c->dyCC[f] * c->dyCC[f];
      sumXY = sumXY + c->dxCC[f] * c->dyCC[f];
      double xee = c->dxCC[f] / c->dlCC[f];
      double yee = c->dyCC[f] / c->dlCC[f];
      double nfn1 = xee * xnn[f] + yee * ynn[f];
      xee = xee / nfn1;
      yee = yee / nfn1;
      c->dle1[f] = 1.0 / nfn1;
      c->dle3[f] = 0.;
      xee = xnn[f] - xee;
      yee = ynn[f] - yee;
      nfn1 = xee * dxVV[f] + yee * dyVV[f];
      if (nfn1 < 0.) c->dle2[f] = -1.0 * sqrt(xee * xee + yee * yee);
      else c->dle2[f] = sqrt(xee * xee + yee * yee);
    }
    c->sum_xc2[i] = sumX;
    c->sum_yc2[i] = sumY;
    c->sum_xcyc[i] = sumXY;
  }
#undef VP
#undef CP
  return 0;
}

/* ------------------------------------------ Tecplot cell-centre elements --- */
/* calcul_geometry.F90:1590-1665: around each vertex, a fan of triangles over
 * the centres of the cells that share it, (initc, neigc, next) with initc
 * fixed, walking from cell to cell across the edges that touch the vertex.
 * The walk starts at the first cell (in cell order) that has the vertex on a
 * ghost edge (a boundary vertex, initc the ghost) or as a corner (an interior
 * vertex) and stops at a ghost or, dropping its last triangle, back at
 * initc.  The reference finds each start by scanning every cell; here the
 * first occurrences are collected in one pass in the same order.  Cell
 * numbers are 1-based as the reference stores them. */
static int build_tecplot(nsx_case *c) {
  const int N0 = c->N_CELL0, NV = c->N_VERT, NC = c->N_CELL;
  const int *No_vp = c->No_vp, *No_cp = c->No_cp;
#define VP(i, j) (No_vp[IX2(i, j, N0)] - 1)
#define CP(i, j) No_cp[IX2(i, j, NC)]
  int *first_b = xcalloc((size_t)NV, sizeof(int)), *first_v = xcalloc((size_t)NV, sizeof(int));
  for (int v = 0; v < NV; ++v) first_b[v] = first_v[v] = -1;
  for (int i = 0; i < N0; ++i)
    for (int j = 0; j < 3; ++j) {
      const int jv = VP(i, j);
      if (first_v[jv] < 0) first_v[jv] = 3 * i + j;
      if (c->nbe[i] != 0 && CP(i, j) > N0) {
        const int jv1 = VP(i, j), jv2 = VP(i, (j + 1) % 3);
        if (first_b[jv1] < 0) first_b[jv1] = 3 * i + j;
        if (first_b[jv2] < 0) first_b[jv2] = 3 * i + j;
      }
    }
  const int cap = 5 * NV;
  int num = 0, initc = 0, lastc = 0, neigc = 0, rc = 0;
  for (int nv = 0; nv < NV && !rc; ++nv) {
    const int f = c->nbev[nv] != 0 ? first_b[nv] : first_v[nv];
    if (f >= 0) {
      const int i = f / 3, j = f % 3;
      if (c->nbev[nv] != 0) {
        initc = CP(i, j);
        lastc = initc;
        neigc = i + 1;
      } else {
        initc = i + 1;
        lastc = i + 1;
        neigc = CP(i, j);
      }
    }
    for (int steps = 0; neigc >= 1 && neigc <= N0; ++steps) {
      if (num >= cap || steps > 64) {
        snprintf(c->err, sizeof c->err, "Tecplot cell-centre walk around vertex %d does not close", nv + 1);
        rc = -1;
        break;
      }
      c->ic1tec[num] = initc;
      c->ic2tec[num] = neigc;
      ++num;
      int next = 0;
      for (int j = 0; j < 3; ++j) {
        const int jc = CP(neigc - 1, j);
        if (jc == lastc) continue;
        const int jv1 = VP(neigc - 1, j), jv2 = VP(neigc - 1, (j + 1) % 3);
        if (jv1 == nv || jv2 == nv) {
          c->ic3tec[num - 1] = jc;
          lastc = neigc;
          neigc = jc;
          next = jc == initc ? 2 : 1;
          break;
        }
      }
      if (next == 2) --num;
      if (next != 1) break;
    }
  }
  c->Numictec[0] = num;
  free(first_b);
  free(first_v);
#undef VP
#undef CP
  return rc;
}

/* -------------------------------------------------------- BC flags ------- */
/* TagXC/TagYC of a cell from its centre (calcul_geometry.F90:857-886): 1 or 2
 * below or above the box in x (y), 0 inside */
static void cell_tags(const nsx_case *c, int n, int *tx, int *ty) {
  *tx = *ty = 0;
  if (c->xc[n] < c->XDIni) *tx = 1;
  if (c->xc[n] > c->XDFin) *tx = 2;
  if (c->yc[n] < c->YDIni) *ty = 1;
  if (c->yc[n] > c->YDFin) *ty = 2;
}

/* the flag of a wall ghost for component q (0 u, 1 v, 2 w, 3 p) from its
 * tags (update_bc.F90:201-516); -9 when no rule applies */
static int wall_flag(const nsx_case *c, int q, int tx, int ty) {
  if (q == 3) return 1; /* all Neumann for p */
  int fB = -9, flag = 0;
  if (tx > 0) {
    ++flag;
    if (c->XBC == 0) fB = 0;
    else if (c->XBC == 1) fB = q == 0 ? 0 : 1;
    else if (c->XBC == 2) fB = tx == 1 ? 2 : 1;
    else fB = -9;
  }
  if (ty > 0) {
    ++flag;
    if (c->YBC == 0) fB = 0;
    else if (c->YBC == 1) fB = q == 0 ? 2 : (q == 1 ? 0 : 1);
    else fB = -9;
  }
  return flag == 1 ? fB : -9;
}

/* the flag of a boundary vertex for component q (update_bc.F90:531-780) */
static int vertex_flag(const nsx_case *c, int q, int nv) {
  int fB = -9, flag = 0;
  const int tx = c->tagXV[nv], ty = c->tagYV[nv];
  if (tx > 0) {
    ++flag;
    if (c->XPB != 0) fB = 3;
    else if (q == 3) fB = 1;
    else if (c->XBC == 0) fB = 0;
    else if (c->XBC == 1) fB = q == 0 ? 0 : 1;
    else if (c->XBC == 2) fB = tx == 1 ? 2 : 1;
  }
  if (ty > 0) {
    ++flag;
    if (c->YPB != 0) fB = 3;
    else if (q == 3) fB = 1;
    else if (c->YBC == 0) fB = 0;
    else if (c->YBC == 1) fB = q == 1 ? 0 : 1;
  }
  return flag == 1 ? fB : -9;
}

/* adjust_tagv / adjust_bc (update_bc.F90:4-848), TopBC/BotBC walls */
static int build_bc(nsx_case *c) {
  const int N0 = c->N_CELL0, NV = c->N_VERT, NC = c->N_CELL;
  for (int i = 0; i < NC; ++i) {
    c->bcuc[i] = c->bcvc[i] = c->bcwc[i] = c->bcpc[i] = -1;
    c->tagsp[i] = 0;
  }
  int *const cf[4] = {c->bcuc, c->bcvc, c->bcwc, c->bcpc};
  for (int i = 0; i < N0; ++i) {
    if (c->nbe[i] == 0) continue;
    for (int j = 0; j < 3; ++j) {
      int nc = c->No_cp[IX2(i, j, NC)];
      if (!(nc < 1 || nc > N0)) continue;
      int ii = c->No_cp[IX2(nc - 1, 2, NC)];
      if (ii == -1) {
        /* a wall: the Kim-Choi split of its face is the plain normal one
         * (update_bc.F90:160-171) */
        const size_t f = IX2(i, j, N0);
        c->tagsp[nc - 1] = 1;
        c->dle1[f] = 1.0;
        c->dle2[f] = 0.;
        c->dle3[f] = 0.0;
      }
      int tx, ty;
      cell_tags(c, nc - 1, &tx, &ty);
      if (tx > 0 && ty > 0) {
        snprintf(c->err, sizeof c->err, "ghost cell %d lies beyond the box in x and y", nc);
        return -1;
      }
      for (int q = 0; q < 4; ++q) {
        int fB;
        if (ii > 0) fB = 3;
        else if (ii == 0) fB = -1;
        else fB = wall_flag(c, q, tx, ty);
        if (fB == -9) {
          snprintf(c->err, sizeof c->err, "no boundary rule for ghost cell %d, component %d", nc, q + 1);
          return -1;
        }
        cf[q][nc - 1] = fB;
      }
    }
  }
  /* adjust_tagv: corners keep the y tag */
  for (int nv = 0; nv < NV; ++nv)
    if (c->nbev[nv] != 0 && c->tagXV[nv] * c->tagYV[nv] != 0) c->tagXV[nv] = 0;
  int *const vf[4] = {c->bcuv, c->bcvv, c->bcwv, c->bcpv};
  for (int nv = 0; nv < NV; ++nv) {
    c->bcuv[nv] = c->bcvv[nv] = c->bcwv[nv] = c->bcpv[nv] = -1;
    if (c->nbev[nv] == 0) continue;
    for (int q = 0; q < 4; ++q) {
      const int fB = vertex_flag(c, q, nv);
      if (fB == -9) {
        snprintf(c->err, sizeof c->err, "no boundary rule for vertex %d, component %d", nv + 1, q + 1);
        return -1;
      }
      vf[q][nv] = fB;
    }
    if (c->vertex_edge[nv] > 0) c->bcpv[nv] = 3;
    if (c->vertex_pair[nv] > 0) c->bcpv[nv] = 3;
  }
  c->bctop = 1;
  c->bcbot = 1;
  return 0;
}

/* ---------------------------------------------------------- PPE matrix --- */
/* PDE_center (PDE.F90:69-121), KeyKimChoi, ZPB = 0 */
static void build_ppe(nsx_case *c) {
  const int N0 = c->N_CELL0, NZ = c->NZ, NC = c->N_CELL;
  for (size_t n = 0; n < (size_t)N0 * NZ; ++n)
    c->AmC0[n] = c->AmC1[n] = c->AmC2[n] = c->AmC3[n] = c->AmCT[n] = c->AmCB[n] = 0.0;
  for (int k = 1; k < NZ - 1; ++k)
    for (int i = 0; i < N0; ++i) {
      double AA[3];
      for (int j = 0; j < 3; ++j) {
        size_t f = IX2(i, j, N0);
        int jc = c->No_cp[IX2(i, j, NC)] - 1;
        double co1 = c->dle1[f] / c->dlCC[f];
        if (c->bcpc[jc] == 1) AA[j] = 0.0;
        else AA[j] = co1 * c->dlVV[f] * c->dsigv[k - 1];
      }
      double AAB = c->areaCell[i] / c->dsig[k - 1];
      double AAT = c->areaCell[i] / c->dsig[k];
      if (k == 1 && c->bcbot == 1) AAB = 0.0;
      if (k == NZ - 2 && c->bctop == 1) AAT = 0.0;
      double AA0 = -1.0 * (AAT + AAB + AA[0] + AA[1] + AA[2]);
      size_t n = IX2(i, k, N0);
      c->AmC0[n] = AA0;
      c->AmC1[n] = AA[0];
      c->AmC2[n] = AA[1];
      c->AmC3[n] = AA[2];
      c->AmCT[n] = AAT;
      c->AmCB[n] = AAB;
    }
}

/* ------------------------------------------------------ initial field ---- */
/* u = 1.5 Ub (1 - (z/h-1)^2) + noise, v = w = noise, noise ~ U(-a,a) Ub, p = 0,
 * at the centroids of interior cells and layers k = 2..NZ-1; ghost entries 0. */
static void build_ic(nsx_case *c) {
  const int NC = c->N_CELL, NZ = c->NZ, N0 = c->N_CELL0;
  const double h = 0.5 * c->H, Ub = c->Ubulk, a = c->noise;
  memset(c->u0, 0, sizeof(double) * NC * NZ);
  memset(c->v0, 0, sizeof(double) * NC * NZ);
  memset(c->w0, 0, sizeof(double) * NC * NZ);
  memset(c->p0, 0, sizeof(double) * NC * NZ);
  for (int k = 1; k < NZ - 1; ++k)
    for (int i = 0; i < N0; ++i) {
      uint64_t idx = (uint64_t)k * (uint64_t)N0 + (uint64_t)i;
      double zeta = c->sig[k] / h - 1.0;
      double up = 1.5 * Ub * (1.0 - zeta * zeta);
      size_t n = IX2(i, k, NC);
      c->u0[n] = up + (2.0 * nsx_rng_u01(c->seed, NSX_STREAM_NOISE_U, idx) - 1.0) * a * Ub;
      c->v0[n] = (2.0 * nsx_rng_u01(c->seed, NSX_STREAM_NOISE_V, idx) - 1.0) * a * Ub;
      c->w0[n] = (2.0 * nsx_rng_u01(c->seed, NSX_STREAM_NOISE_W, idx) - 1.0) * a * Ub;
    }
}

/* -------------------------------------------------------------- API ------ */
void nsx_case_free(nsx_case *c) {
  if (!c) return;
  for (int n = 0; n < c->n_arrays; ++n) free(c->arrays[n].ptr);
  free(c);
}

static void *add_array(nsx_case *c, const char *name, int type, size_t n) {
  size_t sz = type == NSX_T_F64 ? sizeof(double) : sizeof(int);
  void *p = xcalloc(n, sz);
  nsx_case_array *a = &c->arrays[c->n_arrays++];
  snprintf(a->name, sizeof a->name, "%s", name);
  a->type = type;
  a->n = (int64_t)n;
  a->ptr = p;
  return p;
}

static void alloc_case(nsx_case *c) {
  const size_t N0 = c->N_CELL0, NV = c->N_VERT, NC = c->N_CELL, NZ = c->NZ;
  c->No_vp = add_array(c, "No_vp", NSX_T_I32, N0 * 3);
  c->No_cp = add_array(c, "No_cp", NSX_T_I32, NC * 3);
  c->nbe = add_array(c, "nbe", NSX_T_I32, N0);
  c->nbev = add_array(c, "nbev", NSX_T_I32, NV);
  c->xv = add_array(c, "xv", NSX_T_F64, NV);
  c->yv = add_array(c, "yv", NSX_T_F64, NV);
  c->zbv = add_array(c, "zbv", NSX_T_F64, NV);
  c->sig = add_array(c, "sig", NSX_T_F64, NZ);
  c->dsig = add_array(c, "dsig", NSX_T_F64, NZ);
  c->sigv = add_array(c, "sigv", NSX_T_F64, NZ - 1);
  c->dsigv = add_array(c, "dsigv", NSX_T_F64, NZ - 1);
  c->xc = add_array(c, "xc", NSX_T_F64, NC);
  c->yc = add_array(c, "yc", NSX_T_F64, NC);
  c->areaCell = add_array(c, "areaCell", NSX_T_F64, NC);
  c->dxVV = add_array(c, "dxVV", NSX_T_F64, N0 * 3);
  c->dyVV = add_array(c, "dyVV", NSX_T_F64, N0 * 3);
  c->dlVV = add_array(c, "dlVV", NSX_T_F64, N0 * 3);
  c->xnn = add_array(c, "xnn", NSX_T_F64, N0 * 3);
  c->ynn = add_array(c, "ynn", NSX_T_F64, N0 * 3);
  c->dxCC = add_array(c, "dxCC", NSX_T_F64, N0 * 3);
  c->dyCC = add_array(c, "dyCC", NSX_T_F64, N0 * 3);
  c->dlCC = add_array(c, "dlCC", NSX_T_F64, N0 * 3);
  c->xme = add_array(c, "xme", NSX_T_F64, N0 * 3);
  c->yme = add_array(c, "yme", NSX_T_F64, N0 * 3);
  c->xmo = add_array(c, "xmo", NSX_T_F64, N0 * 3);
  c->ymo = add_array(c, "ymo", NSX_T_F64, N0 * 3);
  c->doe = add_array(c, "doe", NSX_T_F64, N0 * 3);
  c->dle1 = add_array(c, "dle1", NSX_T_F64, N0 * 3);
  c->dle2 = add_array(c, "dle2", NSX_T_F64, N0 * 3);
  c->dle3 = add_array(c, "dle3", NSX_T_F64, N0 * 3);
  c->sum_xc2 = add_array(c, "sum_xc2", NSX_T_F64, N0);
  c->sum_yc2 = add_array(c, "sum_yc2", NSX_T_F64, N0);
  c->sum_xcyc = add_array(c, "sum_xcyc", NSX_T_F64, N0);
  c->surrounding = add_array(c, "surrounding", NSX_T_I32, NV * NSX_MAX_SURROUND);
  c->weight = add_array(c, "weight", NSX_T_F64, NV * NSX_MAX_SURROUND);
  c->Dimsurrounding = add_array(c, "Dimsurrounding", NSX_T_I32, NV);
  c->dlVsum = add_array(c, "dlVsum", NSX_T_F64, NV);
  c->vertex_pair = add_array(c, "vertex_pair", NSX_T_I32, NV);
  c->vertex_edge = add_array(c, "vertex_edge", NSX_T_I32, NV);
  c->tagXV = add_array(c, "tagXV", NSX_T_I32, NV);
  c->tagYV = add_array(c, "tagYV", NSX_T_I32, NV);
  c->tagsp = add_array(c, "tagsp", NSX_T_I32, NC);
  c->bcuc = add_array(c, "bcuc", NSX_T_I32, NC);
  c->bcvc = add_array(c, "bcvc", NSX_T_I32, NC);
  c->bcwc = add_array(c, "bcwc", NSX_T_I32, NC);
  c->bcpc = add_array(c, "bcpc", NSX_T_I32, NC);
  c->bcuv = add_array(c, "bcuv", NSX_T_I32, NV);
  c->bcvv = add_array(c, "bcvv", NSX_T_I32, NV);
  c->bcwv = add_array(c, "bcwv", NSX_T_I32, NV);
  c->bcpv = add_array(c, "bcpv", NSX_T_I32, NV);
  c->AmC0 = add_array(c, "AmC0", NSX_T_F64, N0 * NZ);
  c->AmC1 = add_array(c, "AmC1", NSX_T_F64, N0 * NZ);
  c->AmC2 = add_array(c, "AmC2", NSX_T_F64, N0 * NZ);
  c->AmC3 = add_array(c, "AmC3", NSX_T_F64, N0 * NZ);
  c->AmCT = add_array(c, "AmCT", NSX_T_F64, N0 * NZ);
  c->AmCB = add_array(c, "AmCB", NSX_T_F64, N0 * NZ);
  c->walld = add_array(c, "walld", NSX_T_F64, NZ);
  c->u0 = add_array(c, "u0", NSX_T_F64, NC * NZ);
  c->v0 = add_array(c, "v0", NSX_T_F64, NC * NZ);
  c->w0 = add_array(c, "w0", NSX_T_F64, NC * NZ);
  c->p0 = add_array(c, "p0", NSX_T_F64, NC * NZ);
  c->ic1tec = add_array(c, "ic1tec", NSX_T_I32, 5 * NV);
  c->ic2tec = add_array(c, "ic2tec", NSX_T_I32, 5 * NV);
  c->ic3tec = add_array(c, "ic3tec", NSX_T_I32, 5 * NV);
  c->Numictec = add_array(c, "Numictec", NSX_T_I32, 1);
  c->grid = add_array(c, "grid", NSX_T_I32, 3);

}

/* sigma levels, geometry, flags, PPE matrix, wall distance and the initial
 * field of a case whose mesh is set */
static nsx_case *finish_case(nsx_case *c) {
  if ((build_sigma(c), 0) || build_geometry(c) || build_tecplot(c) || build_bc(c)) {
    c->status = -1;
    return c;
  }
  build_ppe(c);
  c->grid[0] = c->nx;
  c->grid[1] = c->ny;
  c->grid[2] = c->NN;
  /* channel wall distance: |sig(k)| (sgm_les.F90:856-862) */
  for (int k = 0; k < c->NZ; ++k) c->walld[k] = fabs(c->sig[k]);
  build_ic(c);
  c->status = 0;
  return c;
}

nsx_case *nsx_case_create(int nx, int ny, int nn, double Lx, double Ly, double H, double stretch,
                          double jitter, int bed, double Ubulk, double noise, uint64_t seed) {
  return nsx_case_create_walls(nx, ny, nn, Lx, Ly, H, stretch, jitter, bed, Ubulk, noise, seed, 1, 1, 0, 1);
}

nsx_case *nsx_case_create_walls(int nx, int ny, int nn, double Lx, double Ly, double H, double stretch,
                                double jitter, int bed, double Ubulk, double noise, uint64_t seed, int XPB, int YPB,
                                int XBC, int YBC) {
  nsx_case *c = xcalloc(1, sizeof(nsx_case));
  c->XPB = XPB;
  c->YPB = YPB;
  c->XBC = XBC;
  c->YBC = YBC;
  c->nx = nx; c->ny = ny; c->NN = nn; c->NZ = nn + 2;
  c->Lx = Lx; c->Ly = Ly; c->H = H; c->stretch = stretch;
  c->jitter = jitter; c->bed = bed; c->Ubulk = Ubulk; c->noise = noise; c->seed = seed;
  c->N_CELL0 = 2 * nx * ny;
  c->N_VERT = (nx + 1) * (ny + 1);
  c->N_CELL = c->N_CELL0 + 100 * nn; /* N_CELLghostMAX = 100*NN (common.mpf:113) */
  c->XDIni = 0.0;
  c->XDFin = Lx;
  c->YDIni = 0.0;
  c->YDFin = Ly;
  if (nx < 2 || ny < 2 || nn < 2 || 2 * (nx + ny) > 100 * nn || (XPB != 0 && XPB != 1) || (YPB != 0 && YPB != 1) ||
      XBC < 0 || XBC > 2 || YBC < 0 || YBC > 1) {
    snprintf(c->err, sizeof c->err, "bad sizes or boundary kinds nx=%d ny=%d nn=%d XPB=%d YPB=%d XBC=%d YBC=%d", nx,
             ny, nn, XPB, YPB, XBC, YBC);
    c->status = -1;
    return c;
  }
  alloc_case(c);
  if (build_mesh(c)) {
    c->status = -1;
    return c;
  }
  return finish_case(c);
}

nsx_case *nsx_case_from_mesh(int N_CELL0, int N_VERT, int nn, const int *No_vp, const double *xv, const double *yv,
                             const int *No_cp, const double *zbv, const int *nbe, const int *nbev, double XDIni,
                             double XDFin, double YDIni, double YDFin, double H, double stretch, double Ubulk,
                             double noise, uint64_t seed) {
  nsx_case *c = xcalloc(1, sizeof(nsx_case));
  c->XPB = c->YPB = 1;
  c->XBC = 0;
  c->YBC = 1;
  c->NN = nn;
  c->NZ = nn + 2;
  c->N_CELL0 = N_CELL0;
  c->N_VERT = N_VERT;
  c->N_CELL = N_CELL0 + 100 * nn; /* N_CELLghostMAX = 100*NN (common.mpf:113) */
  c->Lx = XDFin - XDIni;
  c->Ly = YDFin - YDIni;
  c->XDIni = XDIni;
  c->XDFin = XDFin;
  c->YDIni = YDIni;
  c->YDFin = YDFin;
  c->H = H;
  c->stretch = stretch;
  c->Ubulk = Ubulk;
  c->noise = noise;
  c->seed = seed;
  if (N_CELL0 < 1 || N_VERT < 3 || nn < 2 || !No_vp || !xv || !yv || !No_cp || !nbe || !nbev ||
      !(XDFin > XDIni) || !(YDFin > YDIni)) {
    snprintf(c->err, sizeof c->err, "bad mesh arguments");
    c->status = -1;
    return c;
  }
  alloc_case(c);
  const size_t N0 = (size_t)c->N_CELL0, NC = (size_t)c->N_CELL;
  for (size_t n = 0; n < 3 * N0; ++n)
    if (No_vp[n] < 1 || No_vp[n] > N_VERT) {
      snprintf(c->err, sizeof c->err, "No_vp(%zu,%zu) = %d is not a vertex", n % N0 + 1, n / N0 + 1, No_vp[n]);
      c->status = -1;
      return c;
    }
  memcpy(c->No_vp, No_vp, 3 * N0 * sizeof(int));
  for (int j = 0; j < 3; ++j) memcpy(c->No_cp + (size_t)j * NC, No_cp + (size_t)j * N0, N0 * sizeof(int));
  memcpy(c->nbe, nbe, N0 * sizeof(int));
  memcpy(c->nbev, nbev, (size_t)N_VERT * sizeof(int));
  memcpy(c->xv, xv, (size_t)N_VERT * sizeof(double));
  memcpy(c->yv, yv, (size_t)N_VERT * sizeof(double));
  if (zbv) memcpy(c->zbv, zbv, (size_t)N_VERT * sizeof(double));
  c->bed = 0;
  return finish_case(c);
}

/* ----------------------------------------------------- the mesh file --- */
/* input.F90:147-194.  Each READ starts a record and takes its items field by
 * field from fixed columns, n fields to a record (3I8, 2F15.5, 10F8.3,
 * 20I7), the format starting a new record when it runs out.  A record
 * shorter than its fields is padded with blanks; blanks in a field are
 * ignored and an all-blank field reads as zero. */
typedef struct {
  FILE *f;
  char *buf;
  size_t cap;
  long recno;
} mesh_reader;

static int next_record(mesh_reader *r) {
  size_t n = 0;
  int ch;
  while ((ch = fgetc(r->f)) != EOF && ch != '\n') {
    if (n + 2 > r->cap) {
      r->cap = r->cap ? 2 * r->cap : 256;
      r->buf = realloc(r->buf, r->cap);
      if (!r->buf) return -1;
    }
    r->buf[n++] = (char)ch;
  }
  if (ch == EOF && n == 0) return -1;
  if (!r->buf) {
    r->cap = 256;
    r->buf = malloc(r->cap);
    if (!r->buf) return -1;
  }
  r->buf[n] = 0;
  ++r->recno;
  return 0;
}

/* the non-blank characters of columns [col, col+w) of the record */
static void field_chars(const char *rec, int col, int w, char *out) {
  const int len = (int)strlen(rec);
  int m = 0;
  for (int q = col; q < col + w; ++q)
    if (q < len && rec[q] != ' ') out[m++] = rec[q];
  out[m] = 0;
}

static int parse_iw(const char *t, long *v) {
  if (!*t) {
    *v = 0;
    return 0;
  }
  const char *p = t;
  int neg = 0;
  if (*p == '+' || *p == '-') neg = *p++ == '-';
  if (!*p) return -1;
  long x = 0;
  for (; *p; ++p) {
    if (*p < '0' || *p > '9' || x > 100000000000L) return -1;
    x = 10 * x + (*p - '0');
  }
  *v = neg ? -x : x;
  return 0;
}

/* Fw.d input: [sign] digits [. digits] [exponent], the exponent as E, D, e, d
 * with an optional sign, or a bare sign; without a decimal point the digits
 * are scaled by 10^-d.  The value is the correctly rounded double. */
static int parse_fwd(const char *t, int d, double *v) {
  if (!*t) {
    *v = 0.0;
    return 0;
  }
  char mant[128];
  int m = 0, dot = 0, ndig = 0;
  const char *p = t;
  if (*p == '+' || *p == '-') mant[m++] = *p++;
  for (; *p && m < 100; ++p) {
    if (*p >= '0' && *p <= '9') {
      mant[m++] = *p;
      ++ndig;
    } else if (*p == '.' && !dot) {
      mant[m++] = '.';
      dot = 1;
    } else
      break;
  }
  if (ndig == 0) return -1;
  long ex = 0;
  if (*p) {
    if (*p == 'E' || *p == 'e' || *p == 'D' || *p == 'd') ++p;
    else if (*p != '+' && *p != '-') return -1;
    if (parse_iw(p, &ex) || !*p) return -1;
  }
  if (!dot) ex -= d;
  mant[m] = 0;
  char full[192];
  snprintf(full, sizeof full, "%se%ld", mant, ex);
  char *end = NULL;
  *v = strtod(full, &end);
  return (end && *end == 0) ? 0 : -1;
}

/* n items of format (per_rec(Iw)) or (per_rec(Fw.d)) into iv/dv[0..n) (or,
 * with stride 2, pairs xv(i), yv(i) from 2F15.5 into two arrays) */
static int read_ints(mesh_reader *r, int n, int per_rec, int w, int *iv, int stride_rows, int ncols) {
  /* items in the order of the implied DO: for i, for j (row-major over an
   * (stride_rows, ncols) column-major array when ncols > 1) */
  char fld[64];
  int k = 0;
  for (int item = 0; item < n; ++item) {
    if (k == 0 && next_record(r)) return -1;
    field_chars(r->buf, k * w, w, fld);
    long v;
    if (parse_iw(fld, &v) || v < -2147483647L || v > 2147483647L) return -1;
    const int i = ncols > 1 ? item / ncols : item, j = ncols > 1 ? item % ncols : 0;
    iv[(size_t)i + (size_t)j * stride_rows] = (int)v;
    k = (k + 1) % per_rec;
  }
  return 0;
}

static int read_reals(mesh_reader *r, int n, int per_rec, int w, int d, double *a, double *b) {
  char fld[64];
  int k = 0;
  for (int item = 0; item < n; ++item) {
    if (k == 0 && next_record(r)) return -1;
    field_chars(r->buf, k * w, w, fld);
    double v;
    if (parse_fwd(fld, d, &v)) return -1;
    if (b) {
      if (item % 2 == 0) a[item / 2] = v;
      else b[item / 2] = v;
    } else
      a[item] = v;
    k = (k + 1) % per_rec;
  }
  return 0;
}

int nsx_read_mesh(const char *path, int N_CELL0, int N_VERT, int *No_vp, double *xv, double *yv, int *No_cp,
                  double *zbv, int *nbe, int *nbev) {
  if (!path || N_CELL0 < 1 || N_VERT < 1) return -1;
  mesh_reader r = {fopen(path, "r"), NULL, 0, 0};
  if (!r.f) return -1;
  int rc = 0;
  for (int t = 0; t < 4 && !rc; ++t) rc = next_record(&r); /* three titles and the first section's */
  if (!rc) rc = read_ints(&r, 3 * N_CELL0, 3, 8, No_vp, N_CELL0, 3);
  if (!rc) rc = next_record(&r);
  if (!rc) rc = read_reals(&r, 2 * N_VERT, 2, 15, 5, xv, yv);
  if (!rc) rc = next_record(&r);
  if (!rc) rc = read_ints(&r, 3 * N_CELL0, 3, 8, No_cp, N_CELL0, 3);
  if (!rc) rc = next_record(&r);
  if (!rc) rc = read_reals(&r, N_VERT, 10, 8, 3, zbv, NULL);
  if (!rc) rc = next_record(&r);
  if (!rc) rc = read_ints(&r, N_CELL0, 20, 7, nbe, N_CELL0, 1);
  if (!rc) rc = next_record(&r);
  if (!rc) rc = read_ints(&r, N_VERT, 20, 7, nbev, N_VERT, 1);
  const long at = r.recno;
  fclose(r.f);
  free(r.buf);
  return rc ? (int)-(at + 1) : 0;
}

/* Iw and Fw.d output, asterisks when the value does not fit */
static void out_iw(FILE *f, long v, int w) {
  char b[32];
  const int n = snprintf(b, sizeof b, "%ld", v);
  if (n > w)
    for (int q = 0; q < w; ++q) fputc('*', f);
  else
    fprintf(f, "%*s", w, b);
}

static void out_fwd(FILE *f, double v, int w, int d) {
  char b[64];
  const int n = snprintf(b, sizeof b, "%.*f", d, v);
  if (n > w)
    for (int q = 0; q < w; ++q) fputc('*', f);
  else
    fprintf(f, "%*s", w, b);
}

int nsx_write_mesh(const char *path, int N_CELL0, int N_VERT, const int *No_vp, const double *xv, const double *yv,
                   const int *No_cp, const double *zbv, const int *nbe, const int *nbev) {
  if (!path || N_CELL0 < 1 || N_VERT < 1 || !No_vp || !xv || !yv || !No_cp || !nbe || !nbev) return -1;
  FILE *f = fopen(path, "w");
  if (!f) return -1;
  fprintf(f, " ---------------------------------------------\n");
  fprintf(f, "  prismatic channel mesh: %d cells, %d vertices\n", N_CELL0, N_VERT);
  fprintf(f, " ---------------------------------------------\n");
  fprintf(f, " --------NUMBERING OF CELL VERTICES----\n");
  for (int i = 0; i < N_CELL0; ++i) {
    for (int j = 0; j < 3; ++j) out_iw(f, No_vp[(size_t)i + (size_t)j * N_CELL0], 8);
    fputc('\n', f);
  }
  fprintf(f, " --------COORDINATES OF CELL VERTICES----\n");
  for (int v = 0; v < N_VERT; ++v) {
    out_fwd(f, xv[v], 15, 5);
    out_fwd(f, yv[v], 15, 5);
    fputc('\n', f);
  }
  fprintf(f, " --------NUMBERING OF CELL CENTERS----\n");
  for (int i = 0; i < N_CELL0; ++i) {
    for (int j = 0; j < 3; ++j) out_iw(f, No_cp[(size_t)i + (size_t)j * N_CELL0], 8);
    fputc('\n', f);
  }
  fprintf(f, " --------BOTTOM LEVELS ---------------\n");
  for (int v = 0; v < N_VERT; ++v) {
    out_fwd(f, zbv ? zbv[v] : 0.0, 8, 3);
    if (v % 10 == 9 || v == N_VERT - 1) fputc('\n', f);
  }
  fprintf(f, " --------TYPE OF BOUNDARY CELLS ---------------\n");
  for (int i = 0; i < N_CELL0; ++i) {
    out_iw(f, nbe[i], 7);
    if (i % 20 == 19 || i == N_CELL0 - 1) fputc('\n', f);
  }
  fprintf(f, " --------TYPE OF BOUNDARY VERTEX-----------\n");
  for (int v = 0; v < N_VERT; ++v) {
    out_iw(f, nbev[v], 7);
    if (v % 20 == 19 || v == N_VERT - 1) fputc('\n', f);
  }
  return fclose(f) == 0 ? 0 : -1;
}

/* --------------------------------------------- the cached binary set-up --- */
/* One file holds a finished set-up: the sizes and scalars of the case, then
 * every array by name, type and length, then a 64-bit checksum of all the
 * bytes before it.  Loading it skips the geometry, flag and matrix build. */
static const char CASE_MAGIC[8] = {'N', 'S', 'X', 'C', 'A', 'S', 'E', '1'};

typedef struct {
  FILE *f;
  uint64_t h;
  int err;
} case_io;

/* word-wise multiply-xor hash, for detecting a corrupt or truncated file */
static void io_hash(case_io *io, const void *p, size_t n) {
  const unsigned char *b = p;
  uint64_t h = io->h;
  size_t q = 0;
  for (; q + 8 <= n; q += 8) {
    uint64_t w;
    memcpy(&w, b + q, 8);
    h = (h ^ w) * 0x100000001b3ULL;
    h ^= h >> 29;
  }
  for (; q < n; ++q) h = (h ^ b[q]) * 0x100000001b3ULL;
  io->h = h;
}

static void io_write(case_io *io, const void *p, size_t n) {
  if (io->err) return;
  if (fwrite(p, 1, n, io->f) != n) io->err = 1;
  io_hash(io, p, n);
}

static void io_read(case_io *io, void *p, size_t n) {
  if (io->err) return;
  if (fread(p, 1, n, io->f) != n) {
    io->err = 1;
    return;
  }
  io_hash(io, p, n);
}

static void case_header(nsx_case *c, int32_t *iv, double *dv) {
  int32_t *const ip[17] = {&c->nx, &c->ny, &c->NN, &c->NZ, &c->N_CELL0, &c->N_VERT, &c->N_CELL, &c->N_CELLexact,
                           &c->N_CELLghost, &c->bctop, &c->bcbot, &c->bed, &c->XPB, &c->YPB, &c->XBC, &c->YBC,
                           &c->n_arrays};
  double *const dp[12] = {&c->Lx, &c->Ly, &c->H, &c->stretch, &c->jitter, &c->Ubulk, &c->noise, &c->TotalVolume,
                          &c->XDIni, &c->XDFin, &c->YDIni, &c->YDFin};
  if (iv) {
    for (int q = 0; q < 17; ++q) iv[q] = *ip[q];
    for (int q = 0; q < 12; ++q) dv[q] = *dp[q];
  }
}

int nsx_case_save(const nsx_case *c, const char *path) {
  if (!c || c->status != 0 || !path) return -1;
  case_io io = {fopen(path, "wb"), 0xcbf29ce484222325ULL, 0};
  if (!io.f) return -1;
  int32_t iv[17];
  double dv[12];
  case_header((nsx_case *)c, iv, dv);
  const int32_t version = 1;
  io_write(&io, CASE_MAGIC, 8);
  io_write(&io, &version, 4);
  io_write(&io, iv, sizeof iv);
  io_write(&io, dv, sizeof dv);
  io_write(&io, &c->seed, 8);
  for (int a = 0; a < c->n_arrays; ++a) {
    const nsx_case_array *A = &c->arrays[a];
    const int32_t type = A->type;
    io_write(&io, A->name, sizeof A->name);
    io_write(&io, &type, 4);
    io_write(&io, &A->n, 8);
    io_write(&io, A->ptr, (size_t)A->n * (A->type == NSX_T_F64 ? sizeof(double) : sizeof(int)));
  }
  const uint64_t h = io.h;
  if (!io.err && fwrite(&h, 1, 8, io.f) != 8) io.err = 1;
  if (fclose(io.f) != 0) io.err = 1;
  return io.err ? -1 : 0;
}

nsx_case *nsx_case_load(const char *path) {
  nsx_case *c = xcalloc(1, sizeof(nsx_case));
  case_io io = {path ? fopen(path, "rb") : NULL, 0xcbf29ce484222325ULL, 0};
  c->status = -1;
  if (!io.f) {
    snprintf(c->err, sizeof c->err, "cannot open %s", path ? path : "(null)");
    return c;
  }
  char magic[8];
  int32_t version = 0, iv[17];
  double dv[12];
  io_read(&io, magic, 8);
  io_read(&io, &version, 4);
  io_read(&io, iv, sizeof iv);
  io_read(&io, dv, sizeof dv);
  io_read(&io, &c->seed, 8);
  if (io.err || memcmp(magic, CASE_MAGIC, 8) != 0 || version != 1) {
    snprintf(c->err, sizeof c->err, "%s is not a cached set-up of this version", path);
    fclose(io.f);
    return c;
  }
  c->nx = iv[0]; c->ny = iv[1]; c->NN = iv[2]; c->NZ = iv[3]; c->N_CELL0 = iv[4]; c->N_VERT = iv[5];
  c->N_CELL = iv[6]; c->N_CELLexact = iv[7]; c->N_CELLghost = iv[8]; c->bctop = iv[9]; c->bcbot = iv[10];
  c->bed = iv[11]; c->XPB = iv[12]; c->YPB = iv[13]; c->XBC = iv[14]; c->YBC = iv[15];
  c->Lx = dv[0]; c->Ly = dv[1]; c->H = dv[2]; c->stretch = dv[3]; c->jitter = dv[4]; c->Ubulk = dv[5];
  c->noise = dv[6]; c->TotalVolume = dv[7]; c->XDIni = dv[8]; c->XDFin = dv[9]; c->YDIni = dv[10];
  c->YDFin = dv[11];
  if (c->N_CELL0 < 1 || c->N_VERT < 3 || c->NN < 2 || c->NZ != c->NN + 2 || c->N_CELL < c->N_CELL0 ||
      (int64_t)c->N_CELL * c->NZ > ((int64_t)1 << 34)) {
    snprintf(c->err, sizeof c->err, "%s: sizes out of range", path);
    fclose(io.f);
    return c;
  }
  alloc_case(c);
  if (iv[16] != c->n_arrays) {
    snprintf(c->err, sizeof c->err, "%s holds %d arrays, this build %d", path, iv[16], c->n_arrays);
    fclose(io.f);
    return c;
  }
  for (int a = 0; a < c->n_arrays && !io.err; ++a) {
    nsx_case_array *A = &c->arrays[a];
    char name[sizeof A->name];
    int32_t type = 0;
    int64_t n = 0;
    io_read(&io, name, sizeof name);
    io_read(&io, &type, 4);
    io_read(&io, &n, 8);
    if (io.err || memcmp(name, A->name, sizeof name) != 0 || type != A->type || n != A->n) {
      snprintf(c->err, sizeof c->err, "%s: array %d is not %s of this build", path, a, A->name);
      fclose(io.f);
      return c;
    }
    io_read(&io, A->ptr, (size_t)n * (type == NSX_T_F64 ? sizeof(double) : sizeof(int)));
  }
  const uint64_t h = io.h;
  uint64_t stored = 0;
  if (!io.err && fread(&stored, 1, 8, io.f) != 8) io.err = 1;
  fclose(io.f);
  if (io.err || stored != h) {
    snprintf(c->err, sizeof c->err, "%s is truncated or corrupt (checksum)", path);
    return c;
  }
  c->status = 0;
  return c;
}

int nsx_case_status(const nsx_case *c, char *msg, int len) {
  if (msg && len > 0) snprintf(msg, (size_t)len, "%s", c->err);
  return c->status;
}

int nsx_case_info(const nsx_case *c, int *out /* 8 ints */) {
  out[0] = c->N_CELL0;
  out[1] = c->N_CELL;
  out[2] = c->N_VERT;
  out[3] = c->NZ;
  out[4] = c->N_CELLexact;
  out[5] = c->N_CELLghost;
  out[6] = c->bctop;
  out[7] = c->bcbot;
  return 0;
}

double nsx_case_total_volume(const nsx_case *c) { return c->TotalVolume; }

int nsx_case_num_arrays(const nsx_case *c) { return c->n_arrays; }

const char *nsx_case_array_name(const nsx_case *c, int n) {
  return (n >= 0 && n < c->n_arrays) ? c->arrays[n].name : NULL;
}

void *nsx_case_get(const nsx_case *c, const char *name, int64_t *n, int *type) {
  for (int a = 0; a < c->n_arrays; ++a)
    if (strcmp(c->arrays[a].name, name) == 0) {
      if (n) *n = c->arrays[a].n;
      if (type) *type = c->arrays[a].type;
      return c->arrays[a].ptr;
    }
  return NULL;
}
