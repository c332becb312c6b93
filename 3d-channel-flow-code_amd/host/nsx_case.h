/*
 * nsx_case.h -- the case structure behind the set-up API of nsx_setup.h:
 * mesh, geometry, flags, PPE matrix and initial field of a channel case in
 * the reference's own array layout.
 *
 * Arrays are column-major like the Fortran modules they mirror, with the
 * reference's 1-based connectivity; names match the Fortran variables
 * (module geometry / module variables of the reference).
 */
#ifndef NSX_CASE_H
#define NSX_CASE_H
#include <stdint.h>

#include "nsx_setup.h"

#ifdef __cplusplus
extern "C" {
#endif

#define NSX_MAX_SURROUND 10 /* surrounding(N_VERT,10), module_geometry.F90 */
#define NSX_T_F64 1
#define NSX_T_I32 2

/* counter-based RNG streams */
#define NSX_STREAM_DIAG 1u
#define NSX_STREAM_JITTER_X 2u
#define NSX_STREAM_JITTER_Y 3u
#define NSX_STREAM_NOISE_U 4u
#define NSX_STREAM_NOISE_V 5u
#define NSX_STREAM_NOISE_W 6u
#define NSX_STREAM_FORCE 7u

typedef struct {
  char name[32];
  int type;
  int64_t n;
  void *ptr;
} nsx_case_array;

typedef struct nsx_case {
  int nx, ny, NN, NZ, N_CELL0, N_VERT, N_CELL, N_CELLexact, N_CELLghost;
  int bctop, bcbot, bed, status;
  double Lx, Ly, H, stretch, jitter, Ubulk, noise, TotalVolume;
  double XDIni, XDFin, YDIni, YDFin; /* the box (common.mpf:66-69) */
  int XPB, YPB, XBC, YBC;             /* periodicity and wall kinds (common.mpf:72-80) */
  uint64_t seed;
  char err[256];
  int n_arrays;
  nsx_case_array arrays[96];
  int *No_vp, *No_cp, *nbe, *nbev;
  double *xv, *yv, *zbv;
  double *sig, *dsig, *sigv, *dsigv;
  double *xc, *yc, *areaCell;
  double *dxVV, *dyVV, *dlVV, *xnn, *ynn, *dxCC, *dyCC, *dlCC;
  double *xme, *yme, *xmo, *ymo, *doe, *dle1, *dle2, *dle3;
  double *sum_xc2, *sum_yc2, *sum_xcyc;
  int *surrounding, *Dimsurrounding, *vertex_pair, *vertex_edge, *tagXV, *tagYV, *tagsp;
  double *weight, *dlVsum;
  int *bcuc, *bcvc, *bcwc, *bcpc, *bcuv, *bcvv, *bcwv, *bcpv;
  double *AmC0, *AmC1, *AmC2, *AmC3, *AmCT, *AmCB;
  double *walld;
  double *u0, *v0, *w0, *p0;
  int *ic1tec, *ic2tec, *ic3tec, *Numictec; /* Tecplot cell-centre elements (5 N_VERT), their count */
  int *grid;                                /* nx, ny, NN of a generated mesh (0, 0, NN for a mesh file) */
} nsx_case;

#ifdef __cplusplus
}
#endif
#endif
