/*
 * nsx_setup.h -- the set-up path on the host (libnsx_host.so): the mesh file
 * of the reference, its geometry, boundary flags and PPE matrix, in the
 * reference's own array layout (column-major, 1-based connectivity, ghost
 * cells after N_CELL0, names of module geometry / module variables).
 *
 * What each entry point replaces:
 *   nsx_read_mesh       input.F90:147-194 (data.txt)
 *   nsx_case_from_mesh  calcul_geometry.F90:290-1517 (centroids, areas,
 *                       normals, periodic ghost cells, IDW weights, Kim-Choi
 *                       and LSM coefficients; not the Tecplot connectivity
 *                       ic1tec..ic3tec of :1595-1670), update_bc.F90:4-848
 *                       (adjust_tagv, adjust_bc) and PDE.F90:69-121
 *                       (PDE_center), for the doubly periodic box with sigma
 *                       walls; the periodic searches are indexed, O(N log N)
 *                       instead of the reference's O(ghosts (N_VERT + N0))
 *   nsx_case_create     the same on a seeded synthetic channel mesh
 * A case carries its arrays by name (nsx_case_get); they are the inputs of
 * nsx_set_mesh .. nsx_set_les of nsx.h.  All entry points run on the CPU.
 */
#ifndef NSX_SETUP_H
#define NSX_SETUP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nsx_case nsx_case;

/* input.F90:147-194: the mesh file (three title records, then after a title
 * record each No_vp 3I8 per cell, xv yv 2F15.5 per vertex, No_cp 3I8 per
 * cell, zbv 10F8.3, nbe 20I7, nbev 20I7), read as Fortran's formatted input
 * reads it (blanks ignored, all-blank fields zero, short records padded, a
 * real without a decimal point scaled by 10^-d).  Arrays No_vp, No_cp
 * (N_CELL0,3), nbe (N_CELL0), xv, yv, zbv, nbev (N_VERT).  Returns 0, -1 when
 * the file cannot be opened, or -(record number) of a record that cannot be
 * read. */
int nsx_read_mesh(const char *path, int N_CELL0, int N_VERT, int *No_vp, double *xv, double *yv, int *No_cp,
                  double *zbv, int *nbe, int *nbev);
/* the same format written (a field that does not fit is asterisks) */
int nsx_write_mesh(const char *path, int N_CELL0, int N_VERT, const int *No_vp, const double *xv, const double *yv,
                   const int *No_cp, const double *zbv, const int *nbe, const int *nbev);

/* the set-up of a given mesh (nsx_read_mesh's arrays; zbv may be NULL) in the
 * periodic box [XDIni,XDFin] x [YDIni,YDFin] (common.mpf:66-69), nn layers
 * between 0 and H clustered by stretch (tanh), and a seeded initial field
 * (laminar profile of bulk velocity Ubulk plus noise) */
nsx_case *nsx_case_from_mesh(int N_CELL0, int N_VERT, int nn, const int *No_vp, const double *xv, const double *yv,
                             const int *No_cp, const double *zbv, const int *nbe, const int *nbev, double XDIni,
                             double XDFin, double YDIni, double YDFin, double H, double stretch, double Ubulk,
                             double noise, uint64_t seed);
/* the set-up of a seeded nx x ny quad mesh (each quad split into two
 * triangles, optional jitter of interior vertices, an inert bed) */
nsx_case *nsx_case_create(int nx, int ny, int nn, double Lx, double Ly, double H, double stretch, double jitter,
                          int bed, double Ubulk, double noise, uint64_t seed);
/* the same with the horizontal boundary conditions of common.mpf:72-80: XPB,
 * YPB 1 periodic in x (y), 0 walled; XBC 0 no slip, 1 free slip, 2 inflow
 * and outflow; YBC 0 no slip, 1 free slip.  Wall ghost cells keep the centre
 * mirrored across their edge and get update_bc.F90's flags (nsx_case_create
 * is XPB = YPB = 1). */
nsx_case *nsx_case_create_walls(int nx, int ny, int nn, double Lx, double Ly, double H, double stretch,
                                double jitter, int bed, double Ubulk, double noise, uint64_t seed, int XPB, int YPB,
                                int XBC, int YBC);
void nsx_case_free(nsx_case *c);
/* the cached binary set-up: nsx_case_save writes a finished case (sizes,
 * scalars, every array, a checksum) to path, 0 or -1; nsx_case_load reads it
 * back without rebuilding the geometry, flags and matrix (check
 * nsx_case_status: a missing, foreign, truncated or corrupt file is refused) */
int nsx_case_save(const nsx_case *c, const char *path);
nsx_case *nsx_case_load(const char *path);
/* 0, or -1 with the set-up's message in msg */
int nsx_case_status(const nsx_case *c, char *msg, int len);
/* N_CELL0, N_CELL, N_VERT, NZ, N_CELLexact, N_CELLghost, bctop, bcbot */
int nsx_case_info(const nsx_case *c, int *out);
double nsx_case_total_volume(const nsx_case *c);
/* the arrays by name: type 1 double, 2 int32; n elements */
int nsx_case_num_arrays(const nsx_case *c);
const char *nsx_case_array_name(const nsx_case *c, int n);
void *nsx_case_get(const nsx_case *c, const char *name, int64_t *n, int *type);
/* the seeded counter-based uniform numbers of the synthetic cases */
double nsx_rng_u01(uint64_t seed, uint32_t stream, uint64_t index);

#ifdef __cplusplus
}
#endif
#endif
