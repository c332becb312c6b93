/*
 * nsx_oracle.c -- CPU restatement of the reference's per-step hot path.
 *
 * TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and bench.py's
 * CPU-baseline leg load this library, and only as the checker or as the timed
 * CPU baseline; the product (libnsx.so) never calls it.
 *
 * Every routine follows the reference routine named in its comment loop by
 * loop, in the reference's own array layout ((N_CELL,NZ) column-major with
 * the triangle index fastest, ghost cells after N_CELL0, 1-based No_cp/No_vp)
 * and in the reference's operation order, so that, built with
 * -ffp-contract=off on x86-64, it reproduces the compiled Fortran bit for bit
 * (checked against oracle/_ref and the fixtures in tests/golden).
 *
 * Work the reference does but whose result is multiplied by zero is left out:
 * the least-squares gradients feeding the dle3 (=0) terms of diffusion3DNEW,
 * update_rhsp and newu3D (calcul_geometry.F90:1493 sets dle3 = 0), the first
 * gradientP(pf) of hydro (hydro.F90:381, feeds only those terms) and the
 * vertex values that interpolateU computes but never reads
 * (interpolationU.F90:99-102; its ghost fill is kept).
 *
 * Supported boundary set: horizontal ghosts periodic (flag 3), walls (0),
 * free slip (1) and inflow (2, the constant inflow() of Functions.F90), as
 * update_bc.F90 assigns them for XPB/YPB and XBC/YBC; sigma walls TopBC in
 * {0,1,2} (2: the lid, u = 1 at the top) and BotBC in {0,1}.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef struct orc {
  int N0, NC, NV, NZ, NEX;
  /* geometry, borrowed from the caller (reference layout) */
  const int *No_cp, *No_vp, *nbe, *nbev;
  const double *areaCell, *dlVV, *xnn, *ynn, *dxCC, *dyCC, *dlCC, *doe, *dle1, *dle2;
  const double *sum_xc2, *sum_yc2, *sum_xcyc;
  const int *surrounding, *Dimsurrounding, *vertex_pair, *vertex_edge;
  const double *weight, *dlVsum;
  const int *bcuc, *bcvc, *bcwc, *bcpc, *bcuv, *bcvv, *bcwv, *bcpv;
  const double *AmC0, *AmC1, *AmC2, *AmC3, *AmCT, *AmCB;
  const double *sig, *dsig, *sigv, *dsigv, *walld;
  int bctop, bcbot, TopBC, BotBC;
  double TotalVolume, sigIni, sigFin;
  double dt, Re, bdx, bdy, bdz, eps;
  int MaxIters, FinRK, Iturbu, flag_ab;
  /* state (owned) */
  double *ufn, *vfn, *wfn, *pfn, *uf, *vf, *wf, *pf, *ufnp, *vfnp, *wfnp, *pfnp;
  double *ufv, *vfv, *wfv, *pfv;
  double *U1, *U2, *U3, *UT, *UB; /* U1FACE..UBFACE (N_CELL0,NZ) */
  double *rhsuf, *rhsvf, *rhswf, *eddy;
  /* channel statistics of module stats (module_stats.F90), (N_CELL,NZ) */
  double *sm_uf, *sm_vf, *sm_wf, *sm_u2f, *sm_v2f, *sm_w2f, *sm_uvf, *sm_vwf, *sm_wuf, *sm_pf, *sm_p2f, *sm_ef;
  int sm_n;
  /* solver record */
  int last_iter, n_solves, iter_total;
  double last_e1, last_e2;
  int last_exit; /* 0 converged, 1 no solve needed, 2 breakdown, 3 max iterations */
  double t_solve, t_hydro; /* wall seconds spent in bicgstab / hydro */
  char err[256];
} orc;

#define NCL (o->NC)
#define CI(i, k) ((size_t)(i) + (size_t)(k) * (size_t)NCL)        /* (N_CELL,NZ) */
#define C0(i, k) ((size_t)(i) + (size_t)(k) * (size_t)o->N0)      /* (N_CELL0,NZ) */
#define VI(v, k) ((size_t)(v) + (size_t)(k) * (size_t)o->NV)      /* (N_VERT,NZ-1) */
#define G3(a, i, j) (a)[(size_t)(i) + (size_t)(j) * (size_t)o->N0] /* (N_CELL0,3) */
#define CP(i, j) (o->No_cp[(size_t)(i) + (size_t)(j) * (size_t)NCL] - 1)
#define VP(i, j) (o->No_vp[(size_t)(i) + (size_t)(j) * (size_t)o->N0] - 1)

enum { TAG_P = 0, TAG_U = 1, TAG_V = 2, TAG_W = 3 };

/* The timing build (-fopenmp, liboracle_omp.so) splits the solver's loops
 * over threads; its sums then run in another order, so only the serial
 * build (no -fopenmp: these pragmas vanish) is the bit-exact checker. */
#ifdef _OPENMP
#include <omp.h>
#define ORC_PAR _Pragma("omp parallel for collapse(2) schedule(static)")
#else
#define ORC_PAR
#endif

int orc_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

static double *dalloc(size_t n) {
  double *p = calloc(n ? n : 1, sizeof(double));
  if (!p) abort();
  return p;
}

/* ------------------------------------------------------------ BC cells --- */
/* inflow() of Functions.F90:109-121: 1 for tagBC 2 (v), 0 otherwise, at any
 * point (the reference's comment says u; its test is flag 2) */
static double inflow_value(int tag) { return tag == TAG_V ? 1.0 : 0.0; }

/* BCvelcenter3D, BC.F90:367-594 (ZPB = 0) */
void orc_bcvelcenter(orc *o, double *phi, int tag) {
  const int N0 = o->N0, NZ = o->NZ;
  if (o->bcbot == 1) {
    int anti = (o->BotBC == 0) || (o->BotBC == 1 && tag == TAG_W);
    for (int i = 0; i < N0; ++i) phi[CI(i, 0)] = anti ? -phi[CI(i, 1)] : phi[CI(i, 1)];
  }
  if (o->bctop == 1) {
    if (o->TopBC == 2 && tag == TAG_U) {
      /* the lid (BC.F90:512-521) */
      for (int i = 0; i < N0; ++i) phi[CI(i, NZ - 1)] = 2.0 - phi[CI(i, NZ - 2)];
    } else {
      int anti = (o->TopBC == 0) || (o->TopBC == 1 && tag == TAG_W) || o->TopBC == 2;
      for (int i = 0; i < N0; ++i) phi[CI(i, NZ - 1)] = anti ? -phi[CI(i, NZ - 2)] : phi[CI(i, NZ - 2)];
    }
  }
  const int *bc = tag == TAG_U ? o->bcuc : (tag == TAG_V ? o->bcvc : o->bcwc);
  for (int k = 0; k < NZ; ++k)
    for (int i = 0; i < N0; ++i) {
      if (o->nbe[i] == 0) continue;
      for (int j = 0; j < 3; ++j) {
        int nc = CP(i, j);
        int flag = bc[nc];
        double fB = 0.;
        if (flag == 0) fB = 0. - phi[CI(i, k)];
        if (flag == 1) fB = phi[CI(i, k)];
        if (flag == 2) fB = 2.0 * inflow_value(tag) - phi[CI(i, k)];
        if (flag == 3) fB = phi[CI(o->No_cp[(size_t)nc + 2 * (size_t)NCL] - 1, k)];
        if (flag != -1) phi[CI(nc, k)] = fB;
      }
    }
}

/* BCpcenter, BC.F90:812-977 (ZPB = 0) */
void orc_bcpcenter(orc *o, double *phi) {
  const int N0 = o->N0, NZ = o->NZ;
  if (o->bctop == 1)
    for (int i = 0; i < N0; ++i) phi[CI(i, NZ - 1)] = phi[CI(i, NZ - 2)];
  if (o->bcbot == 1)
    for (int i = 0; i < N0; ++i) phi[CI(i, 0)] = phi[CI(i, 1)];
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (int k = 0; k < NZ; ++k)
    for (int i = 0; i < N0; ++i) {
      if (o->nbe[i] == 0) continue;
      for (int j = 0; j < 3; ++j) {
        int nc = CP(i, j);
        int flag = o->bcpc[nc];
        double fB = 0.;
        if (flag == 0) fB = -phi[CI(i, k)];
        if (flag == 1) fB = phi[CI(i, k)];
        if (flag == 3) fB = phi[CI(o->No_cp[(size_t)nc + 2 * (size_t)NCL] - 1, k)];
        if (flag != -1) phi[CI(nc, k)] = fB;
      }
    }
}

/* ----------------------------------------------------- cell -> vertex ---- */
/* interpolation2D (interpolation2D.F90:8-150) for every layer, then
 * interpolation3D / interpolation3DP (interpolation3D.F90:8-220, 236-444) */
void orc_interp3d(orc *o, double *phiv, const double *phi, int tagdir) {
  const int NV = o->NV, NZ = o->NZ;
  double *dz0 = dalloc((size_t)NV * NZ);
  for (int k = 0; k < NZ; ++k)
    for (int nv = 0; nv < NV; ++nv) {
      double s = 0.0;
      for (int j = 0; j < o->Dimsurrounding[nv]; ++j) {
        int nc = o->surrounding[(size_t)nv + (size_t)j * NV] - 1;
        s = s + o->weight[(size_t)nv + (size_t)j * NV] * phi[CI(nc, k)];
      }
      dz0[(size_t)nv + (size_t)k * NV] = s / o->dlVsum[nv];
    }
  if (tagdir == 3) {
    for (int k = 0; k < NZ - 1; ++k) {
      double dzT = fabs(o->sigv[k] - o->sig[k + 1]);
      double dzB = fabs(o->sigv[k] - o->sig[k]);
      for (int nv = 0; nv < NV; ++nv)
        phiv[VI(nv, k)] = (dzB * dz0[(size_t)nv + (size_t)(k + 1) * NV] + dzT * dz0[(size_t)nv + (size_t)k * NV]) /
                          (dzT + dzB);
    }
  } else {
    for (int k = 0; k < NZ - 1; ++k)
      for (int nv = 0; nv < NV; ++nv) phiv[VI(nv, k)] = dz0[(size_t)nv + (size_t)k * NV];
  }
  free(dz0);
}

/* bc_periodic, BC.F90:1141-1408 (sequential branch) */
void orc_bc_periodic(orc *o, double *phiv, int tag) {
  const int NV = o->NV, NZ = o->NZ;
  const int *bc = tag == TAG_U ? o->bcuv : tag == TAG_V ? o->bcvv : tag == TAG_W ? o->bcwv : o->bcpv;
  int vpair[4], ncount = 0;
  for (int nv = 0; nv < NV && ncount < 4; ++nv)
    if (o->vertex_edge[nv] == 1) vpair[ncount++] = nv;
  double *old = dalloc((size_t)NV);
  for (int k = 0; k < NZ - 1; ++k) {
    for (int nv = 0; nv < NV; ++nv) old[nv] = 0.0;
    for (int nv = 0; nv < NV; ++nv) {
      if (o->nbev[nv] == 0) continue;
      if (o->vertex_edge[nv] == 1) {
        double sumdlV = 0.0;
        for (int q = 0; q < ncount; ++q) {
          int pair = vpair[q];
          old[nv] = old[nv] + phiv[VI(pair, k)] * o->dlVsum[pair];
          sumdlV = sumdlV + o->dlVsum[pair];
        }
        old[nv] = old[nv] / sumdlV;
      } else if (o->vertex_pair[nv] > 0) {
        int pair = o->vertex_pair[nv] - 1;
        double sumdlV = o->dlVsum[nv] + o->dlVsum[pair];
        old[nv] = o->dlVsum[nv] * phiv[VI(nv, k)] + o->dlVsum[pair] * phiv[VI(pair, k)];
        old[nv] = old[nv] / sumdlV;
      }
    }
    for (int nv = 0; nv < NV; ++nv)
      if (o->nbev[nv] != 0 && bc[nv] == 3) phiv[VI(nv, k)] = old[nv];
  }
  free(old);
}

/* BCvelvertex3D, BC.F90:603-807 */
void orc_bcvelvertex(orc *o, double *phiv, int tag, int tagdir) {
  const int NV = o->NV, NZ = o->NZ;
  const int *bc = tag == TAG_U ? o->bcuv : (tag == TAG_V ? o->bcvv : o->bcwv);
  for (int nv = 0; nv < NV; ++nv) {
    if (o->nbev[nv] != 0 && bc[nv] == 0)
      for (int k = 0; k < NZ - 1; ++k) phiv[VI(nv, k)] = 0.;
    if (o->nbev[nv] != 0 && bc[nv] == 2)
      for (int k = 0; k < NZ - 1; ++k) phiv[VI(nv, k)] = inflow_value(tag);
  }
  if (tagdir == 3) {
    if (o->bcbot == 1 && (o->BotBC == 0 || (o->BotBC == 1 && tag == TAG_W)))
      for (int nv = 0; nv < NV; ++nv) phiv[VI(nv, 0)] = 0.;
    if (o->bctop == 1 && (o->TopBC == 0 || (o->TopBC == 1 && tag == TAG_W) || (o->TopBC == 2 && tag != TAG_U)))
      for (int nv = 0; nv < NV; ++nv) phiv[VI(nv, NZ - 2)] = 0.;
    /* the lid: u = 1 at the top level of the interior vertices (BC.F90:771-777) */
    if (o->bctop == 1 && o->TopBC == 2 && tag == TAG_U)
      for (int nv = 0; nv < NV; ++nv)
        if (o->nbev[nv] == 0) phiv[VI(nv, NZ - 2)] = 1.0;
  }
  orc_bc_periodic(o, phiv, tag);
}

/* BCpvertex, BC.F90:986-1135 */
void orc_bcpvertex(orc *o, double *phiv) {
  for (int k = 0; k < o->NZ - 1; ++k)
    for (int nv = 0; nv < o->NV; ++nv)
      if (o->nbev[nv] != 0 && o->bcpv[nv] == 0) phiv[VI(nv, k)] = 0.;
  orc_bc_periodic(o, phiv, TAG_P);
}

/* BCglobalVC, BC.F90:202-295 */
void orc_bcglobalvc(orc *o, double *u, double *v, double *w, double *uv, double *vv, double *wv, int tagdir) {
  orc_bcvelcenter(o, u, TAG_U);
  orc_interp3d(o, uv, u, tagdir);
  orc_bcvelvertex(o, uv, TAG_U, tagdir);
  orc_bcvelcenter(o, v, TAG_V);
  orc_interp3d(o, vv, v, tagdir);
  orc_bcvelvertex(o, vv, TAG_V, tagdir);
  orc_bcvelcenter(o, w, TAG_W);
  orc_interp3d(o, wv, w, tagdir);
  orc_bcvelvertex(o, wv, TAG_W, tagdir);
}

/* BCglobalP, BC.F90:299-363 */
void orc_bcglobalp(orc *o, double *p, double *pv, int tagdir) {
  orc_bcpcenter(o, p);
  orc_interp3d(o, pv, p, tagdir);
  orc_bcpvertex(o, pv);
}

/* ---------------------------------------------------------- gradients ---- */
/* grandientLSM, gradientLSM.F90:8-211 */
void orc_grad_lsm(orc *o, double *dx, double *dy, double *ds, const double *f) {
  const int N0 = o->N0, NZ = o->NZ;
  memset(dx, 0, sizeof(double) * (size_t)NCL * NZ);
  memset(dy, 0, sizeof(double) * (size_t)NCL * NZ);
  memset(ds, 0, sizeof(double) * (size_t)NCL * NZ);
  for (int k = 1; k < NZ - 1; ++k)
    for (int i = 0; i < N0; ++i) {
      double sumfx = 0.0, sumfy = 0.0;
      for (int j = 0; j < 3; ++j) {
        int jc = CP(i, j);
        sumfx = sumfx + G3(o->dxCC, i, j) * (f[CI(jc, k)] - f[CI(i, k)]);
        sumfy = sumfy + G3(o->dyCC, i, j) * (f[CI(jc, k)] - f[CI(i, k)]);
      }
      double deter = o->sum_xc2[i] * o->sum_yc2[i] - o->sum_xcyc[i] * o->sum_xcyc[i];
      dx[CI(i, k)] = (o->sum_yc2[i] * sumfx - o->sum_xcyc[i] * sumfy) / deter;
      dy[CI(i, k)] = (o->sum_xc2[i] * sumfy - o->sum_xcyc[i] * sumfx) / deter;
    }
  /* vertical: the reference loops i = 1..N_CELL; slots past N_CELLexact hold 0 */
  for (int k = 1; k < NZ - 1; ++k) {
    double dzp = o->sig[k + 1] - o->sig[k];
    double dzm = o->sig[k - 1] - o->sig[k];
    double s2 = dzp * dzp + dzm * dzm;
    for (int i = 0; i < o->NEX; ++i)
      ds[CI(i, k)] = ((f[CI(i, k + 1)] - f[CI(i, k)]) * dzp + (f[CI(i, k - 1)] - f[CI(i, k)]) * dzm) / s2;
  }
  orc_bcpcenter(o, dx);
  orc_bcpcenter(o, dy);
  orc_bcpcenter(o, ds);
}

/* --------------------------------------------------------- face values --- */
/* calcul_face, interpolationFace.F90:85-111 (KeyKimChoi; Option ignored) */
void orc_calcul_face(orc *o, double *f1, double *f2, double *f3, double *fT, double *fB, const double *fun,
                     const double *funv) {
  const int N0 = o->N0, NZ = o->NZ;
  for (int k = 1; k < NZ - 1; ++k)
    for (int i = 0; i < N0; ++i) {
      double phi[3];
      for (int j = 0; j < 3; ++j) {
        int jc = CP(i, j), nv1 = VP(i, j), nv2 = VP(i, (j + 1) % 3);
        double phi1 = fun[CI(i, k)], phi2 = fun[CI(jc, k)];
        double extra = G3(o->doe, i, j) * (funv[VI(nv2, k)] - funv[VI(nv1, k)]) / G3(o->dlVV, i, j);
        phi[j] = phi2 * 0.5 + phi1 * 0.5 + extra;
      }
      f1[C0(i, k)] = phi[0];
      f2[C0(i, k)] = phi[1];
      f3[C0(i, k)] = phi[2];
      fT[C0(i, k)] = 0.5 * (fun[CI(i, k)] + fun[CI(i, k + 1)]);
      fB[C0(i, k)] = 0.5 * (fun[CI(i, k)] + fun[CI(i, k - 1)]);
    }
}

/* interpolateU, interpolationU.F90:5-190: face-normal velocities */
void orc_interpolate_u(orc *o, double *u, double *v, double *w) {
  const int N0 = o->N0, NZ = o->NZ;
  /* BCglobalVC(..,2): only its ghost fill is used */
  orc_bcvelcenter(o, u, TAG_U);
  orc_bcvelcenter(o, v, TAG_V);
  orc_bcvelcenter(o, w, TAG_W);
  for (int k = 1; k < NZ - 1; ++k)
    for (int i = 0; i < N0; ++i) {
      double Ui[3], Vi[3];
      for (int j = 0; j < 3; ++j) {
        int jc = CP(i, j);
        Ui[j] = 0.5 * (u[CI(i, k)] + u[CI(jc, k)]);
        Vi[j] = 0.5 * (v[CI(i, k)] + v[CI(jc, k)]);
      }
      double WiT = 0.5 * (w[CI(i, k)] + w[CI(i, k + 1)]);
      double WiB = 0.5 * (w[CI(i, k)] + w[CI(i, k - 1)]);
      o->U1[C0(i, k)] = Ui[0] * G3(o->xnn, i, 0) + Vi[0] * G3(o->ynn, i, 0);
      o->U2[C0(i, k)] = Ui[1] * G3(o->xnn, i, 1) + Vi[1] * G3(o->ynn, i, 1);
      o->U3[C0(i, k)] = Ui[2] * G3(o->xnn, i, 2) + Vi[2] * G3(o->ynn, i, 2);
      o->UT[C0(i, k)] = WiT;
      o->UB[C0(i, k)] = -1.0 * WiB;
    }
}

/* ---------------------------------------------------- momentum fluxes ---- */
/* diffusion3DNEW Kim-Choi branch, diffusion3D.F90:110-191 */
static void diffusion(orc *o, double *rhs, const double *phi, const double *phiv, const double *Gam_eddy) {
  const int N0 = o->N0, NZ = o->NZ;
  const double nu = 1. / o->Re;
  for (int k = 1; k < NZ - 1; ++k)
    for (int i = 0; i < N0; ++i) {
      double Vol = o->areaCell[i] * o->dsigv[k - 1];
      double dn[3];
      for (int j = 0; j < 3; ++j) {
        int jc = CP(i, j), jv1 = VP(i, j), jv2 = VP(i, (j + 1) % 3);
        double diff1 = (phi[CI(jc, k)] - phi[CI(i, k)]) / G3(o->dlCC, i, j);
        double diff2 = (phiv[VI(jv2, k)] - phiv[VI(jv1, k)]) / G3(o->dlVV, i, j);
        dn[j] = diff1 * G3(o->dle1, i, j) + diff2 * G3(o->dle2, i, j);
      }
      double dT = (phi[CI(i, k + 1)] - phi[CI(i, k)]) / o->dsig[k];
      double dB = (phi[CI(i, k - 1)] - phi[CI(i, k)]) / o->dsig[k - 1];
      double src = dn[0] * G3(o->dlVV, i, 0) * o->dsigv[k - 1] + dn[1] * G3(o->dlVV, i, 1) * o->dsigv[k - 1] +
                   dn[2] * G3(o->dlVV, i, 2) * o->dsigv[k - 1] + (dT + dB) * o->areaCell[i];
      double Gam = Gam_eddy[CI(i, k)] + nu;
      rhs[CI(i, k)] = rhs[CI(i, k)] + Gam * src / Vol;
    }
}

/* advection3DNEW, advection3D.F90:8-312 (central, explicit) */
static void advection(orc *o, double *ru, double *rv, double *rw, const double *u, const double *v, const double *w,
                      const double *uv, const double *vv, const double *wv) {
  const int N0 = o->N0, NZ = o->NZ;
  size_t n0z = (size_t)N0 * NZ;
  double *F = dalloc(15 * n0z);
  double *Uf[5], *Vf[5], *Wf[5];
  for (int q = 0; q < 5; ++q) {
    Uf[q] = F + q * n0z;
    Vf[q] = F + (5 + q) * n0z;
    Wf[q] = F + (10 + q) * n0z;
  }
  orc_calcul_face(o, Uf[0], Uf[1], Uf[2], Uf[3], Uf[4], u, uv);
  orc_calcul_face(o, Vf[0], Vf[1], Vf[2], Vf[3], Vf[4], v, vv);
  orc_calcul_face(o, Wf[0], Wf[1], Wf[2], Wf[3], Wf[4], w, wv);
  for (int k = 1; k < NZ - 1; ++k)
    for (int i = 0; i < N0; ++i) {
      size_t c = C0(i, k);
      double UF1 = o->U1[c] * G3(o->dlVV, i, 0) * o->dsigv[k - 1];
      double UF2 = o->U2[c] * G3(o->dlVV, i, 1) * o->dsigv[k - 1];
      double UF3 = o->U3[c] * G3(o->dlVV, i, 2) * o->dsigv[k - 1];
      double UFT = o->UT[c] * o->areaCell[i];
      double UFB = o->UB[c] * o->areaCell[i];
      double uflux = Uf[0][c] * UF1 + Uf[1][c] * UF2 + Uf[2][c] * UF3 + Uf[3][c] * UFT + Uf[4][c] * UFB;
      double vflux = Vf[0][c] * UF1 + Vf[1][c] * UF2 + Vf[2][c] * UF3 + Vf[3][c] * UFT + Vf[4][c] * UFB;
      double wflux = Wf[0][c] * UF1 + Wf[1][c] * UF2 + Wf[2][c] * UF3 + Wf[3][c] * UFT + Wf[4][c] * UFB;
      double Vol = o->areaCell[i] * o->dsigv[k - 1];
      ru[CI(i, k)] = ru[CI(i, k)] - uflux / Vol;
      rv[CI(i, k)] = rv[CI(i, k)] - vflux / Vol;
      rw[CI(i, k)] = rw[CI(i, k)] - wflux / Vol;
    }
  free(F);
}

/* AdvDiffNS, AdvDiffNS.F90:8-140 (DoThis = 3); Gam = eddy + 1/Re */
void orc_advdiff(orc *o, double *ru, double *rv, double *rw, double *u, double *v, double *w, const double *eddy) {
  size_t nvz = (size_t)o->NV * (o->NZ - 1);
  double *uv = dalloc(nvz), *vv = dalloc(nvz), *wv = dalloc(nvz);
  orc_bcglobalvc(o, u, v, w, uv, vv, wv, 2);
  diffusion(o, ru, u, uv, eddy);
  diffusion(o, rv, v, vv, eddy);
  diffusion(o, rw, w, wv, eddy);
  advection(o, ru, rv, rw, u, v, w, uv, vv, wv);
  free(uv);
  free(vv);
  free(wv);
}

/* ------------------------------------------------------- Smagorinsky ----- */
/* sgdamp, sgm_les.F90:1247-1281 (Mason blending, KeyLESVan off), for the
 * Mason power n = 2 of module_les.F90:31.  The compiled reference inlines it
 * into sgles with n constant, so (1/x)**n is evaluated as (1/x)*(1/x) and
 * dn2**(1/n) as sqrt(dn2); those are the bits reproduced here (a libm pow
 * differs from them in the last place for some cells). */
static double sgdamp(double del, double dist, double karman) {
  double lwall = karman * dist;
  double dn1 = (1.0 / del) * (1.0 / del) + (1.0 / lwall) * (1.0 / lwall);
  double dn2 = 1.0 / dn1;
  del = sqrt(dn2);
  return del * del;
}

/* sgles + fvm_gradient, sgm_les.F90:24-275, 284-502 (KeyTESTChannel: utau = 1) */
void orc_sgles(orc *o, double *u, double *v, double *w) {
  const int N0 = o->N0, NZ = o->NZ;
  size_t nvz = (size_t)o->NV * (NZ - 1), n0z = (size_t)N0 * NZ;
  double *uv = dalloc(nvz), *vv = dalloc(nvz), *wv = dalloc(nvz);
  orc_bcglobalvc(o, u, v, w, uv, vv, wv, 2);
  double *F = dalloc(15 * n0z);
  double *Uf[5], *Vf[5], *Wf[5];
  for (int q = 0; q < 5; ++q) {
    Uf[q] = F + q * n0z;
    Vf[q] = F + (5 + q) * n0z;
    Wf[q] = F + (10 + q) * n0z;
  }
  orc_calcul_face(o, Uf[0], Uf[1], Uf[2], Uf[3], Uf[4], u, uv);
  orc_calcul_face(o, Vf[0], Vf[1], Vf[2], Vf[3], Vf[4], v, vv);
  orc_calcul_face(o, Wf[0], Wf[1], Wf[2], Wf[3], Wf[4], w, wv);
  const double Cs = 0.1, karman = (double)0.415f; /* 0.415 is a REAL(4) literal, module_les.F90:33 */
  for (int k = 1; k < NZ - 1; ++k)
    for (int i = 0; i < N0; ++i) {
      size_t c = C0(i, k);
      double x1 = G3(o->xnn, i, 0), x2 = G3(o->xnn, i, 1), x3 = G3(o->xnn, i, 2);
      double y1 = G3(o->ynn, i, 0), y2 = G3(o->ynn, i, 1), y3 = G3(o->ynn, i, 2);
      double l1 = G3(o->dlVV, i, 0), l2 = G3(o->dlVV, i, 1), l3 = G3(o->dlVV, i, 2);
      double a = o->areaCell[i];
      double dudx = Uf[0][c] * x1 * l1 + Uf[1][c] * x2 * l2 + Uf[2][c] * x3 * l3;
      double dvdx = Vf[0][c] * x1 * l1 + Vf[1][c] * x2 * l3 + Vf[2][c] * x3 * l2; /* sgm_les.F90:444-446 */
      double dwdx = Wf[0][c] * x1 * l1 + Wf[1][c] * x2 * l2 + Wf[2][c] * x3 * l3;
      double dudy = Uf[0][c] * y1 * l1 + Uf[1][c] * y2 * l2 + Uf[2][c] * y3 * l3;
      double dvdy = Vf[0][c] * y1 * l1 + Vf[1][c] * y2 * l3 + Vf[2][c] * y3 * l2; /* sgm_les.F90:453-455 */
      double dwdy = Wf[0][c] * y1 * l1 + Wf[1][c] * y2 * l2 + Wf[2][c] * y3 * l3;
      dudx = dudx / a;
      dvdx = dvdx / a;
      dwdx = dwdx / a;
      dudy = dudy / a;
      dvdy = dvdy / a;
      dwdy = dwdy / a;
      double dudz = (Uf[3][c] - Uf[4][c]) / o->dsig[k];
      double dvdz = (Vf[3][c] - Vf[4][c]) / o->dsig[k];
      double dwdz = (Wf[3][c] - Wf[4][c]) / o->dsig[k];
      double DD = Cs * pow(a * o->dsig[k] * 1.0, 1.0 / 3.0);
      DD = sgdamp(DD, o->walld[k], karman);
      double nut = 1.0 * DD;
      double sxx = (dudx + dudx) / 2.0, sxy = (dudy + dvdx) / 2.0, sxz = (dudz + dwdx) / 2.0;
      double syy = (dvdy + dvdy) / 2.0, syz = (dvdz + dwdy) / 2.0, szz = (dwdz + dwdz) / 2.0;
      double Sij = sqrt(2.0 * (sxx * sxx + syy * syy + szz * szz) + 4.0 * (sxy * sxy + sxz * sxz + syz * syz));
      o->eddy[CI(i, k)] = nut * Sij;
    }
  orc_bcpcenter(o, o->eddy);
  free(F);
  free(uv);
  free(vv);
  free(wv);
}

/* ------------------------------------------------------- PPE source ------ */
/* Source3D (Option 1), source3D.F90:8-174 */
void orc_source3d(orc *o, double *src, double *u, double *v, double *w) {
  const int N0 = o->N0, NZ = o->NZ;
  orc_interpolate_u(o, u, v, w);
  memset(src, 0, sizeof(double) * (size_t)NCL * NZ);
  for (int k = 1; k < NZ - 1; ++k)
    for (int i = 0; i < N0; ++i) {
      size_t c = C0(i, k);
      double phi[3] = {o->U1[c], o->U2[c], o->U3[c]};
      double resid = 0.0;
      for (int j = 0; j < 3; ++j) resid = resid + phi[j] * G3(o->dlVV, i, j) * o->dsigv[k - 1];
      resid = resid + o->UT[c] * o->areaCell[i];
      resid = resid + o->UB[c] * o->areaCell[i];
      src[CI(i, k)] = resid;
    }
}

/* update_rhsp Kim-Choi branch, source3D.F90:255-288 */
void orc_update_rhsp(orc *o, double *rhs, double *p, double *pv) {
  const int N0 = o->N0, NZ = o->NZ;
  orc_bcglobalp(o, p, pv, 2);
  for (int k = 1; k < NZ - 1; ++k)
    for (int i = 0; i < N0; ++i) {
      double dn[3];
      for (int j = 0; j < 3; ++j) {
        int jv1 = VP(i, j), jv2 = VP(i, (j + 1) % 3);
        double diff2 = (pv[VI(jv2, k)] - pv[VI(jv1, k)]) / G3(o->dlVV, i, j);
        dn[j] = diff2 * G3(o->dle2, i, j);
      }
      double s = dn[0] * G3(o->dlVV, i, 0) * o->dsigv[k - 1] + dn[1] * G3(o->dlVV, i, 1) * o->dsigv[k - 1] +
                 dn[2] * G3(o->dlVV, i, 2) * o->dsigv[k - 1];
      rhs[CI(i, k)] = rhs[CI(i, k)] - s;
    }
}

/* -------------------------------------------------------------- solver --- */
/* Glob_Norm, bicgstab.F90:598-666 */
static void glob_norm(orc *o, double *f, int tags) {
  const int N0 = o->N0, NZ = o->NZ;
  double local_sum = 0.0;
#ifdef _OPENMP
#pragma omp parallel for collapse(2) schedule(static) reduction(+ : local_sum)
#endif
  for (int k = 1; k < NZ - 1; ++k)
    for (int i = 0; i < N0; ++i) {
      double Vol = o->areaCell[i] * o->dsig[k];
      if (tags == 1) local_sum = f[CI(i, k)] + local_sum;
      else local_sum = f[CI(i, k)] * Vol + local_sum;
    }
  double avec = local_sum / (o->TotalVolume * (o->sigFin - o->sigIni));
  ORC_PAR
  for (int k = 1; k < NZ - 1; ++k)
    for (int i = 0; i < N0; ++i) {
      if (tags == 1) f[CI(i, k)] = f[CI(i, k)] - avec * o->areaCell[i] * o->dsig[k];
      else f[CI(i, k)] = f[CI(i, k)] - avec;
    }
}

static double innerprod(orc *o, const double *a, const double *b) {
  double s = 0.0;
#ifdef _OPENMP
#pragma omp parallel for collapse(2) schedule(static) reduction(+ : s)
#endif
  for (int k = 1; k < o->NZ - 1; ++k)
    for (int i = 0; i < o->N0; ++i) s = s + a[CI(i, k)] * b[CI(i, k)];
  return s;
}

/* get_Apk, bicgstab.F90:931-1001 (TagBC = TagBCp) */
static void get_apk(orc *o, const double *A[6], double *out, double *x) {
  const int N0 = o->N0, NZ = o->NZ;
  orc_bcpcenter(o, x);
  ORC_PAR
  for (int k = 1; k < NZ - 1; ++k)
    for (int i = 0; i < N0; ++i) {
      size_t c = C0(i, k);
      int jc1 = CP(i, 0), jc2 = CP(i, 1), jc3 = CP(i, 2);
      out[CI(i, k)] = A[0][c] * x[CI(i, k)] + A[1][c] * x[CI(jc1, k)] + A[2][c] * x[CI(jc2, k)] +
                      A[3][c] * x[CI(jc3, k)] + A[4][c] * x[CI(i, k + 1)] + A[5][c] * x[CI(i, k - 1)];
    }
}

/* BICGSTAB_SOLVE, bicgstab.F90:303-591 (KeyPPECenter, diag_precdt Opt 1,
 * I3D = 1).  Returns the iteration count; o->last_* record the exit. */
int orc_bicgstab_solve(orc *o, const double *A[6], double *rhs, double *phi) {
  const int N0 = o->N0, NZ = o->NZ;
  size_t nz = (size_t)NCL * NZ;
  double *W = dalloc(8 * nz);
  double *r0bar = W, *vk = W + nz, *sk = W + 2 * nz, *tk = W + 3 * nz, *yk = W + 4 * nz, *rk = W + 5 * nz,
         *pk = W + 6 * nz, *zk = W + 7 * nz;
  double b2 = 0.0, r2 = 0.0, b1 = 0.0, r1 = 0.0;
#ifdef _OPENMP
#pragma omp parallel for collapse(2) schedule(static) reduction(+ : b2, r2, b1, r1)
#endif
  for (int k = 1; k < NZ - 1; ++k)
    for (int i = 0; i < N0; ++i) {
      size_t c = C0(i, k);
      int jc1 = CP(i, 0), jc2 = CP(i, 1), jc3 = CP(i, 2);
      double resid = rhs[CI(i, k)] - (A[0][c] * phi[CI(i, k)] + A[1][c] * phi[CI(jc1, k)] + A[2][c] * phi[CI(jc2, k)] +
                                      A[3][c] * phi[CI(jc3, k)] + A[4][c] * phi[CI(i, k + 1)] +
                                      A[5][c] * phi[CI(i, k - 1)]);
      b2 = b2 + rhs[CI(i, k)] * rhs[CI(i, k)];
      r2 = r2 + resid * resid;
      b1 = b1 + fabs(rhs[CI(i, k)]);
      r1 = r1 + fabs(resid);
      rk[CI(i, k)] = resid;
      r0bar[CI(i, k)] = resid;
    }
  b2 = sqrt(b2);
  r2 = sqrt(r2);
  if (b2 == 0) b2 = 1.0;
  if (b1 == 0) b1 = 1.0;
  double e2 = r2 / b2, e1 = r1 / b1;
  int iter = 0;
  o->last_exit = 0;
  if (e1 < o->eps || e2 < o->eps) {
    o->last_exit = 1;
    goto done;
  }
  double rhokm = 1.0, alpha = 1.0, omega = 1.0, rhok;
  const double tiny = (double)1.0e-30f;
  for (;;) {
    if (fabs(rhokm) < tiny || fabs(omega) < tiny) {
      o->last_exit = 2;
      break;
    }
    rhok = innerprod(o, r0bar, rk);
    if (iter == 0) {
      ORC_PAR
      for (int k = 1; k < NZ - 1; ++k)
        for (int i = 0; i < N0; ++i) pk[CI(i, k)] = rk[CI(i, k)];
    } else {
      double beta = (rhok / rhokm) * (alpha / omega);
      ORC_PAR
      for (int k = 1; k < NZ - 1; ++k)
        for (int i = 0; i < N0; ++i) pk[CI(i, k)] = rk[CI(i, k)] + beta * (pk[CI(i, k)] - omega * vk[CI(i, k)]);
    }
    ORC_PAR
    for (int k = 1; k < NZ - 1; ++k)
      for (int i = 0; i < N0; ++i) yk[CI(i, k)] = pk[CI(i, k)] / A[0][C0(i, k)];
    get_apk(o, A, vk, yk);
    double r0v = innerprod(o, r0bar, vk);
    if (fabs(r0v) < tiny) {
      o->last_exit = 2;
      break;
    }
    alpha = rhok / r0v;
    ORC_PAR
    for (int k = 1; k < NZ - 1; ++k)
      for (int i = 0; i < N0; ++i) sk[CI(i, k)] = rk[CI(i, k)] - alpha * vk[CI(i, k)];
    ORC_PAR
    for (int k = 1; k < NZ - 1; ++k)
      for (int i = 0; i < N0; ++i) zk[CI(i, k)] = sk[CI(i, k)] / A[0][C0(i, k)];
    get_apk(o, A, tk, zk);
    double var1 = innerprod(o, tk, sk), var2 = innerprod(o, tk, tk);
    omega = var1 / var2;
    /* update_xkrk, bicgstab.F90:1047-1111 */
    r2 = 0.0;
    r1 = 0.0;
#ifdef _OPENMP
#pragma omp parallel for collapse(2) schedule(static) reduction(+ : r2, r1)
#endif
    for (int k = 1; k < NZ - 1; ++k)
      for (int i = 0; i < N0; ++i) {
        size_t c = CI(i, k);
        double xnew = phi[c] + alpha * yk[c] + omega * zk[c];
        double resid = sk[c] - omega * tk[c];
        r2 = r2 + resid * resid;
        r1 = r1 + fabs(resid);
        phi[c] = xnew;
        rk[c] = resid;
      }
    r2 = sqrt(r2);
    e2 = r2 / b2;
    e1 = r1 / b1;
    rhokm = rhok;
    iter = iter + 1;
    double enorm = (double)1.0e10f;
    if (enorm > e1) enorm = e1;
    if (enorm > e2) enorm = e2;
    if (enorm < o->eps) break;
    if (iter > o->MaxIters) {
      o->last_exit = 3;
      break;
    }
  }
done:
  free(W);
  o->last_iter = iter;
  o->last_e1 = e1;
  o->last_e2 = e2;
  o->n_solves += 1;
  o->iter_total += iter;
  return iter;
}

static double now(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* bicgstab, bicgstab.F90:8-168 (flg = 1 under KeyTESTChannel) */
int orc_bicgstab(orc *o, double *phi, double *phiv, double *rhs) {
  const double *A[6] = {o->AmC0, o->AmC1, o->AmC2, o->AmC3, o->AmCT, o->AmCB};
  double t0 = now();
  glob_norm(o, rhs, 1);
  int it = orc_bicgstab_solve(o, A, rhs, phi);
  glob_norm(o, phi, 0);
  orc_bcglobalp(o, phi, phiv, 3);
  o->t_solve += now() - t0;
  return it;
}

/* gradientP (Option 2 = LSM), gradientP.F90:8-150 */
void orc_gradient_p(orc *o, double *fun, double *dx, double *dy, double *ds) { orc_grad_lsm(o, dx, dy, ds, fun); }

/* newu3D Kim-Choi branch, newu3D.F90:119-239 */
void orc_newu3d(orc *o, double *fun, double *funv) {
  const int N0 = o->N0, NZ = o->NZ;
  const double c1 = o->dt;
  orc_bcglobalp(o, fun, funv, 2);
  for (int k = 1; k < NZ - 1; ++k)
    for (int i = 0; i < N0; ++i) {
      double dn[3];
      for (int j = 0; j < 3; ++j) {
        int jc = CP(i, j), jv1 = VP(i, j), jv2 = VP(i, (j + 1) % 3);
        if (o->bcpc[jc] == 1) {
          dn[j] = 0.;
          continue;
        }
        double diff1 = (fun[CI(jc, k)] - fun[CI(i, k)]) / G3(o->dlCC, i, j);
        double diff2 = (funv[VI(jv2, k)] - funv[VI(jv1, k)]) / G3(o->dlVV, i, j);
        dn[j] = diff1 * G3(o->dle1, i, j) + diff2 * G3(o->dle2, i, j);
      }
      double dT = (fun[CI(i, k + 1)] - fun[CI(i, k)]) / o->dsig[k];
      double dB = (fun[CI(i, k - 1)] - fun[CI(i, k)]) / o->dsig[k]; /* dsig(k), newu3D.F90:164 */
      size_t c = C0(i, k);
      o->U1[c] = o->U1[c] - dn[0] * c1;
      o->U2[c] = o->U2[c] - dn[1] * c1;
      o->U3[c] = o->U3[c] - dn[2] * c1;
      o->UT[c] = o->UT[c] - dT * c1;
      o->UB[c] = o->UB[c] - dB * c1;
    }
}

/* ------------------------------------------------------------- stage ----- */
/* hydro, hydro.F90:8-501: one projection stage from (uf,vf,wf,pf) into
 * (ufnp,vfnp,wfnp,pfnp), vertex values, faces, rhs history and eddy. */
/* ------------------------------------------------------------ glob_cfl --- */
/* cfl.F90:84-200 (sequential branch) on velocities u, v, w and the module
 * face velocities U1..UB.  out[13]: umax, vmax, wmax (the signed value of
 * largest magnitude, first in k-then-i order), cfl_max, aveu, avev, avew,
 * flux_total, flux_max (signed, largest magnitude), flux_bc, flux_bot,
 * flux_top, flux_side.  A face counts as a side face when its neighbour
 * No_cp(i,j) is not an interior cell (cfl.F90:150-159). */
void orc_glob_cfl(orc *o, const double *u, const double *v, const double *w, double *out) {
  double umax = 0., vmax = 0., wmax = 0., sumu = 0., sumv = 0., sumw = 0.;
  for (int k = 1; k < o->NZ - 1; ++k)
    for (int i = 0; i < o->N0; ++i) {
      const size_t s = CI(i, k);
      if (fabs(u[s]) > fabs(umax)) umax = u[s];
      if (fabs(v[s]) > fabs(vmax)) vmax = v[s];
      if (fabs(w[s]) > fabs(wmax)) wmax = w[s];
      sumu = u[s] * o->areaCell[i] * o->dsig[k] + sumu;
      sumv = v[s] * o->areaCell[i] * o->dsig[k] + sumv;
      sumw = w[s] * o->areaCell[i] * o->dsig[k] + sumw;
    }
  double cfl = 0., flux_l = 0., flux_m = 0., flux_i = 0., flux_o = 0., flux_w = 0.;
  for (int k = 1; k < o->NZ - 1; ++k)
    for (int i = 0; i < o->N0; ++i) {
      double flux = 0., flux_cfl = 0.;
      const double uface[3] = {o->U1[C0(i, k)], o->U2[C0(i, k)], o->U3[C0(i, k)]};
      for (int j = 0; j < 3; ++j) {
        flux = flux + uface[j] * G3(o->dlVV, i, j) * o->dsig[k - 1];
        flux_cfl = flux_cfl + fabs(uface[j] * G3(o->dlVV, i, j) * o->dsig[k]);
        const int jc = CP(i, j);
        if (jc < 0 || jc >= o->N0) flux_w = flux_w + uface[j] * G3(o->dlVV, i, j) * o->dsig[k];
      }
      flux = flux + o->UT[C0(i, k)] * o->areaCell[i];
      flux = flux + o->UB[C0(i, k)] * o->areaCell[i];
      flux_cfl = flux_cfl + fabs(o->UB[C0(i, k)] * o->areaCell[i]) + fabs(o->UT[C0(i, k)] * o->areaCell[i]);
      flux_cfl = 0.5 * o->dt * flux_cfl / (o->areaCell[i] * o->dsig[k]);
      if (flux_cfl > cfl) cfl = flux_cfl;
      flux_l = flux_l + flux;
      if (fabs(flux_m) < fabs(flux)) flux_m = flux;
      if (k == 1 && o->bcbot == 1) flux_i = flux_i + o->UB[C0(i, k)] * o->areaCell[i];
      if (k == o->NZ - 2 && o->bctop == 1) flux_o = flux_o + o->UT[C0(i, k)] * o->areaCell[i];
    }
  const double Vol = o->TotalVolume * (o->sigFin - o->sigIni);
  out[0] = umax;
  out[1] = vmax;
  out[2] = wmax;
  out[3] = cfl;
  out[4] = sumu / Vol;
  out[5] = sumv / Vol;
  out[6] = sumw / Vol;
  out[7] = flux_l;
  out[8] = flux_m;
  out[9] = flux_i + flux_o + flux_w;
  out[10] = flux_i;
  out[11] = flux_o;
  out[12] = flux_w;
}

int orc_hydro(orc *o) {
  const int N0 = o->N0, NZ = o->NZ;
  const double t0 = now();
  size_t nz = (size_t)NCL * NZ;
  double c1 = 1.0, c2 = 0.0;
  if (o->flag_ab == 1) {
    c1 = 1.55;
    c2 = -0.55;
  }
  if (o->Iturbu == 1) orc_sgles(o, o->uf, o->vf, o->wf);
  double *T = dalloc(8 * nz);
  double *us = T, *vs = T + nz, *ws = T + 2 * nz, *ru = T + 3 * nz, *rv = T + 4 * nz, *rw = T + 5 * nz, *rp = T + 6 * nz;
  for (size_t n = 0; n < nz; ++n) {
    ru[n] = o->bdx;
    rv[n] = o->bdy;
    rw[n] = o->bdz;
  }
  memcpy(us, o->uf, sizeof(double) * nz);
  memcpy(vs, o->vf, sizeof(double) * nz);
  memcpy(ws, o->wf, sizeof(double) * nz);
  orc_advdiff(o, ru, rv, rw, o->uf, o->vf, o->wf, o->eddy);
  for (int k = 1; k < NZ - 1; ++k)
    for (int i = 0; i < N0; ++i) {
      size_t c = CI(i, k);
      us[c] = us[c] + o->dt * (c1 * ru[c] + c2 * o->rhsuf[c]);
      vs[c] = vs[c] + o->dt * (c1 * rv[c] + c2 * o->rhsvf[c]);
      o->rhsuf[c] = ru[c];
      o->rhsvf[c] = rv[c];
      ws[c] = ws[c] + o->dt * (c1 * rw[c] + c2 * o->rhswf[c]);
      o->rhswf[c] = rw[c];
    }
  orc_bcglobalvc(o, us, vs, ws, o->ufv, o->vfv, o->wfv, 3);
  orc_source3d(o, rp, us, vs, ws);
  for (int k = 1; k < NZ - 1; ++k)
    for (int i = 0; i < N0; ++i) rp[CI(i, k)] = rp[CI(i, k)] / o->dt;
  memcpy(o->pfnp, o->pf, sizeof(double) * nz);
  orc_update_rhsp(o, rp, o->pf, o->pfv);
  int it = orc_bicgstab(o, o->pfnp, o->pfv, rp);
  double *dpdx = ru, *dpdy = rv, *dpds = rw;
  orc_gradient_p(o, o->pfnp, dpdx, dpdy, dpds);
  for (int k = 1; k < NZ - 1; ++k)
    for (int i = 0; i < N0; ++i) {
      size_t c = CI(i, k);
      o->ufnp[c] = us[c] - o->dt * dpdx[c];
      o->vfnp[c] = vs[c] - o->dt * dpdy[c];
      o->wfnp[c] = ws[c] - o->dt * dpds[c];
    }
  orc_bcglobalvc(o, o->ufnp, o->vfnp, o->wfnp, o->ufv, o->vfv, o->wfv, 3);
  orc_newu3d(o, o->pfnp, o->pfv);
  if (o->FinRK == 1) o->flag_ab = 1;
  free(T);
  o->t_hydro += now() - t0;
  return it;
}

/* one time step, nsmp3D.F90:404-564 */
int orc_step(orc *o) {
  size_t nz = (size_t)NCL * o->NZ;
  int it = 0;
  memcpy(o->uf, o->ufn, sizeof(double) * nz);
  memcpy(o->vf, o->vfn, sizeof(double) * nz);
  memcpy(o->wf, o->wfn, sizeof(double) * nz);
  memcpy(o->pf, o->pfn, sizeof(double) * nz);
  for (int s = 0; s < o->FinRK; ++s) {
    it += orc_hydro(o);
    if (o->FinRK == 2) {
      memcpy(o->uf, o->ufnp, sizeof(double) * nz);
      memcpy(o->vf, o->vfnp, sizeof(double) * nz);
      memcpy(o->wf, o->wfnp, sizeof(double) * nz);
      memcpy(o->pf, o->pfnp, sizeof(double) * nz);
    }
  }
  if (o->FinRK != 1)
    for (size_t n = 0; n < nz; ++n) {
      o->ufnp[n] = 0.5 * (o->ufn[n] + o->ufnp[n]);
      o->vfnp[n] = 0.5 * (o->vfn[n] + o->vfnp[n]);
      o->wfnp[n] = 0.5 * (o->wfn[n] + o->wfnp[n]);
    }
  memcpy(o->ufn, o->ufnp, sizeof(double) * nz);
  memcpy(o->vfn, o->vfnp, sizeof(double) * nz);
  memcpy(o->wfn, o->wfnp, sizeof(double) * nz);
  memcpy(o->pfn, o->pfnp, sizeof(double) * nz);
  return it;
}

/* ----------------------------------------------------------- set-up ------ */
/* ------------------------------------------------------- statistics ---- */
/* sm_init, stats.F90:34-84 (rows 1..N_CELL0, every layer) */
void orc_sm_init(orc *o) {
  double *a[12] = {o->sm_uf, o->sm_vf, o->sm_wf, o->sm_u2f, o->sm_v2f, o->sm_w2f,
                   o->sm_uvf, o->sm_vwf, o->sm_wuf, o->sm_pf, o->sm_p2f, o->sm_ef};
  for (int q = 0; q < 12; ++q)
    for (int k = 0; k < o->NZ; ++k)
      for (int i = 0; i < o->N0; ++i) a[q][CI(i, k)] = 0.0;
  o->sm_n = 0;
}

/* glob_sample, stats.F90:91-197 (KeyLES: the eddy term from o->eddy) */
void orc_glob_sample(orc *o, const double *phiu, const double *phiv, const double *phiw, const double *phip) {
  for (int k = 0; k < o->NZ; ++k)
    for (int i = 0; i < o->N0; ++i) {
      size_t c = CI(i, k);
      double tu = phiu[c], tv = phiv[c], tw = phiw[c], tp = phip[c];
      double tp2 = tp * tp, tu2 = tu * tu, tv2 = tv * tv, tw2 = tw * tw;
      double tuv = tu * tv, tvw = tv * tw, twu = tw * tu;
      o->sm_uf[c] = o->sm_uf[c] + tu;
      o->sm_vf[c] = o->sm_vf[c] + tv;
      o->sm_wf[c] = o->sm_wf[c] + tw;
      o->sm_pf[c] = o->sm_pf[c] + tp;
      o->sm_u2f[c] = o->sm_u2f[c] + tu2;
      o->sm_v2f[c] = o->sm_v2f[c] + tv2;
      o->sm_w2f[c] = o->sm_w2f[c] + tw2;
      o->sm_uvf[c] = o->sm_uvf[c] + tuv;
      o->sm_vwf[c] = o->sm_vwf[c] + tvw;
      o->sm_wuf[c] = o->sm_wuf[c] + twu;
      o->sm_p2f[c] = o->sm_p2f[c] + tp2;
      o->sm_ef[c] = o->sm_ef[c] + o->eddy[c];
    }
  o->sm_n = o->sm_n + 1;
}

/* the plane sums of glob_put_plane_mm, stats.F90:266-279: planes[q*NZ + k]
 * (k 0-based) for the nine velocity moments, interior layers; returns sm_n */
int orc_sm_planes(orc *o, double *planes) {
  const double *a[9] = {o->sm_uf, o->sm_vf, o->sm_wf, o->sm_u2f, o->sm_v2f, o->sm_w2f, o->sm_uvf, o->sm_vwf, o->sm_wuf};
  const int ns = o->sm_n;
  for (int q = 0; q < 9; ++q)
    for (int k = 0; k < o->NZ; ++k) planes[(size_t)q * o->NZ + k] = 0.0;
  for (int k = 1; k < o->NZ - 1; ++k)
    for (int i = 0; i < o->N0; ++i)
      for (int q = 0; q < 9; ++q) planes[(size_t)q * o->NZ + k] = planes[(size_t)q * o->NZ + k] + a[q][CI(i, k)] / ns;
  return ns;
}

orc *orc_create(int N0, int NC, int NV, int NZ, int NEX) {
  orc *o = calloc(1, sizeof(orc));
  o->N0 = N0;
  o->NC = NC;
  o->NV = NV;
  o->NZ = NZ;
  o->NEX = NEX;
  size_t nz = (size_t)NC * NZ, nvz = (size_t)NV * (NZ - 1), n0z = (size_t)N0 * NZ;
  o->ufn = dalloc(nz); o->vfn = dalloc(nz); o->wfn = dalloc(nz); o->pfn = dalloc(nz);
  o->uf = dalloc(nz); o->vf = dalloc(nz); o->wf = dalloc(nz); o->pf = dalloc(nz);
  o->ufnp = dalloc(nz); o->vfnp = dalloc(nz); o->wfnp = dalloc(nz); o->pfnp = dalloc(nz);
  o->ufv = dalloc(nvz); o->vfv = dalloc(nvz); o->wfv = dalloc(nvz); o->pfv = dalloc(nvz);
  o->U1 = dalloc(n0z); o->U2 = dalloc(n0z); o->U3 = dalloc(n0z); o->UT = dalloc(n0z); o->UB = dalloc(n0z);
  o->rhsuf = dalloc(nz); o->rhsvf = dalloc(nz); o->rhswf = dalloc(nz); o->eddy = dalloc(nz);
  o->sm_uf = dalloc(nz); o->sm_vf = dalloc(nz); o->sm_wf = dalloc(nz); o->sm_u2f = dalloc(nz);
  o->sm_v2f = dalloc(nz); o->sm_w2f = dalloc(nz); o->sm_uvf = dalloc(nz); o->sm_vwf = dalloc(nz);
  o->sm_wuf = dalloc(nz); o->sm_pf = dalloc(nz); o->sm_p2f = dalloc(nz); o->sm_ef = dalloc(nz);
  o->FinRK = 2;
  o->Iturbu = 1;
  o->MaxIters = 20000;
  o->eps = 1e-10;
  return o;
}

void orc_free(orc *o) {
  if (!o) return;
  double *own[] = {o->ufn, o->vfn, o->wfn, o->pfn, o->uf, o->vf, o->wf, o->pf, o->ufnp, o->vfnp, o->wfnp, o->pfnp,
                   o->ufv, o->vfv, o->wfv, o->pfv, o->U1, o->U2, o->U3, o->UT, o->UB,
                   o->rhsuf, o->rhsvf, o->rhswf, o->eddy, o->sm_uf, o->sm_vf, o->sm_wf, o->sm_u2f, o->sm_v2f,
                   o->sm_w2f, o->sm_uvf, o->sm_vwf, o->sm_wuf, o->sm_pf, o->sm_p2f, o->sm_ef};
  for (size_t n = 0; n < sizeof own / sizeof own[0]; ++n) free(own[n]);
  free(o);
}

#define BORROW(X)                                                                                                     \
  X(No_cp) X(No_vp) X(nbe) X(nbev) X(areaCell) X(dlVV) X(xnn) X(ynn) X(dxCC) X(dyCC) X(dlCC) X(doe) X(dle1) X(dle2) \
  X(sum_xc2) X(sum_yc2) X(sum_xcyc) X(surrounding) X(Dimsurrounding) X(vertex_pair) X(vertex_edge) X(weight)         \
  X(dlVsum) X(bcuc) X(bcvc) X(bcwc) X(bcpc) X(bcuv) X(bcvv) X(bcwv) X(bcpv) X(AmC0) X(AmC1) X(AmC2) X(AmC3) X(AmCT) \
  X(AmCB) X(sig) X(dsig) X(sigv) X(dsigv) X(walld)

int orc_set_array(orc *o, const char *name, const void *ptr) {
#define SETP(n)                        \
  if (strcmp(name, #n) == 0) {         \
    *(const void **)&o->n = ptr;       \
    return 0;                          \
  }
  BORROW(SETP)
#undef SETP
  return -1;
}

#define STATE(X)                                                                                            \
  X(ufn) X(vfn) X(wfn) X(pfn) X(uf) X(vf) X(wf) X(pf) X(ufnp) X(vfnp) X(wfnp) X(pfnp) X(ufv) X(vfv) X(wfv) \
  X(pfv) X(U1) X(U2) X(U3) X(UT) X(UB) X(rhsuf) X(rhsvf) X(rhswf) X(eddy) X(sm_uf) X(sm_vf) X(sm_wf)     \
  X(sm_u2f) X(sm_v2f) X(sm_w2f) X(sm_uvf) X(sm_vwf) X(sm_wuf) X(sm_pf) X(sm_p2f) X(sm_ef)

double *orc_state(orc *o, const char *name) {
#define GETS(n) \
  if (strcmp(name, #n) == 0) return o->n;
  STATE(GETS)
#undef GETS
  return NULL;
}

void orc_set_params(orc *o, double dt, double Re, double bdx, double bdy, double bdz, double eps, int MaxIters,
                    int FinRK, int Iturbu, int TopBC, int BotBC, int bctop, int bcbot, double TotalVolume,
                    double sigIni, double sigFin) {
  o->dt = dt; o->Re = Re; o->bdx = bdx; o->bdy = bdy; o->bdz = bdz; o->eps = eps;
  o->MaxIters = MaxIters; o->FinRK = FinRK; o->Iturbu = Iturbu; o->TopBC = TopBC; o->BotBC = BotBC;
  o->bctop = bctop; o->bcbot = bcbot; o->TotalVolume = TotalVolume; o->sigIni = sigIni; o->sigFin = sigFin;
}

/* user_initial.F90:229-253 after the case's initial field is in ufn..pfn */
void orc_init_state(orc *o) {
  orc_bcglobalvc(o, o->ufn, o->vfn, o->wfn, o->ufv, o->vfv, o->wfv, 3);
  orc_bcglobalp(o, o->pfn, o->pfv, 3);
  orc_interpolate_u(o, o->ufn, o->vfn, o->wfn);
  size_t nz = (size_t)NCL * o->NZ;
  memset(o->rhsuf, 0, nz * sizeof(double));
  memset(o->rhsvf, 0, nz * sizeof(double));
  memset(o->rhswf, 0, nz * sizeof(double));
  /* LoopTimeInitial: (.)np <- (.)n */
  memcpy(o->ufnp, o->ufn, nz * sizeof(double));
  memcpy(o->vfnp, o->vfn, nz * sizeof(double));
  memcpy(o->wfnp, o->wfn, nz * sizeof(double));
  memcpy(o->pfnp, o->pfn, nz * sizeof(double));
  o->flag_ab = 0;
}

void orc_timers(orc *o, double *dv) {
  dv[0] = o->t_solve;
  dv[1] = o->t_hydro;
}

void orc_solver_record(orc *o, int *iv, double *dv) {
  iv[0] = o->last_iter;
  iv[1] = o->n_solves;
  iv[2] = o->iter_total;
  iv[3] = o->last_exit;
  dv[0] = o->last_e1;
  dv[1] = o->last_e2;
}
