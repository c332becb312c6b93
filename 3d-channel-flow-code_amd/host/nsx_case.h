/*
 * nsx_case.h -- host-side set-up of a channel case (mesh, geometry, flags,
 * PPE matrix, initial field) in the reference's own array layout.
 *
 * Arrays are column-major like the Fortran modules they mirror, with the
 * reference's 1-based connectivity; names match the Fortran variables
 * (module geometry / module variables of the reference).
 */
#ifndef NSX_CASE_H
#define NSX_CASE_H
#include <stdint.h>

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
} nsx_case;

double nsx_rng_u01(uint64_t seed, uint32_t stream, uint64_t index);

nsx_case *nsx_case_create(int nx, int ny, int nn, double Lx, double Ly, double H, double stretch,
                          double jitter, int bed, double Ubulk, double noise, uint64_t seed);
void nsx_case_free(nsx_case *c);
int nsx_case_status(const nsx_case *c, char *msg, int len);
int nsx_case_info(const nsx_case *c, int *out);
double nsx_case_total_volume(const nsx_case *c);
int nsx_case_num_arrays(const nsx_case *c);
const char *nsx_case_array_name(const nsx_case *c, int n);
void *nsx_case_get(const nsx_case *c, const char *name, int64_t *n, int *type);

#ifdef __cplusplus
}
#endif
#endif
