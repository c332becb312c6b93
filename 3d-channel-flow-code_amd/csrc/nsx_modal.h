// nsx_modal.h -- tables of the vertical-mode preconditioner (NSX_REFINE_MODAL):
// built on the host by modal_build_host (nsx_modal.hip), uploaded by the
// context, read by the kernels in nsx_bicgstab.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#define NSX_MODAL_MAXLEV 12
// vertical modes solved by the 2D multigrid (the others by their diagonal)
#define NSX_MODAL_M0 16
// A cycle of the FP32 CG starts from a true FP64 residual (beta = 0) and ends
// when rho = <r, M^-1 r> has fallen below NSX_MODAL_DROP times its first
// value (about 1e-5 in the residual, what FP32 recursion holds to), or after
// the iterations the host launched for it: NSX_MODAL_FIRST for the first
// cycle, then the last cycle's count + 4 within [8, NSX_MODAL_CYCLE]
#define NSX_MODAL_DROP 1e-10
#define NSX_MODAL_FIRST 30
#define NSX_MODAL_CYCLE 60
// the damping of the V-cycle's Jacobi sweeps
#define NSX_MODAL_OMEGA 0.8f

// host tables
struct ModalHost {
  int N0 = 0, nzi = 0, m0 = 0, nc = 0;
  std::vector<float> phi;  // nzi x nzi, phi[k*nzi+m]: component k (interior layer k+1) of mode m
  std::vector<float> mu;   // nzi eigenvalues, ascending
  struct Level {
    int n = 0;
    std::vector<int> ptr, col;   // CSR of the off-diagonal weights
    std::vector<float> w, diag, area;
    std::vector<int> agg;        // coarse node of each node (all levels but the coarsest)
    std::vector<int> eagg;       // agg of each edge's far node
    // all but level 0: the restriction of the pre-smoothed residual, one CSR
    // over the level below: row J holds its children i (db = 1 - omega on b_i)
    // and the fine neighbours j of its children (dz = omega x the sum of w_ij
    // over those children, on z_j), merged per fine node
    std::vector<int> dptr, dcol;
    std::vector<float> db, dz;
  };
  std::vector<Level> levels;  // level 0: the N0 columns
  std::vector<float> cinv;    // m0 dense nc x nc inverses of the coarsest level
};

// tlay[q+1]: the vertical coupling per unit area between interior layers q
// and q+1 (q = 0..nzi-2); tlay[0] and tlay[nzi]: the couplings of the bottom
// and top interior layers to their ghost layers where those are not Neumann
// (a diagonal term), else 0; svl[q]: the horizontal couplings' layer factor
// of interior layer q; nbr/ppec: the columns' neighbours (a value >= N0 or < 0
// is a wall) and edge coefficients, N0 x 3.  Returns 0, or -1 with err set.
int modal_build_host(int N0, int nzi, const int *nbr, const double *ppec, const double *area, const double *tlay,
                     const double *svl, int m0, ModalHost &H, std::string &err);

// device view (pointers into device memory)
struct ModalLevelDev {
  int n;
  const int *ptr, *col, *agg, *eagg, *dptr, *dcol;
  const float *w, *db, *dz, *diag, *area;
};

struct ModalDev {
  int m0, nzi, nlev, nc;
  int mp;     // modes padded to a multiple of 16 (the row length of xh)
  int mt;     // mp / 16; above 8 the vector-unit transforms
  int agrid;  // blocks of the transform kernels (the partial rows of <r, u>)
  int lds_optin;  // the most dynamic LDS a block may opt into (sharedMemPerBlockOptin)
  int nzp;    // nzi rounded up to a multiple of 4 (the vector-unit tables)
  const float *phi;   // nzp x nzp, [k][m]
  const float *phiT;  // nzp x nzp, [m][k]
  const float4 *tfwd;  // matrix-core operands of the forward transform, [layer group][mode tile][lane] float4
  const float4 *tbwd;  // and of the backward one, [mode group][layer tile][lane]
  const float *mu;    // mp (padding modes: 1)
  const float *cinv;  // the coarsest level's inverses, [J][K][m]
  ModalLevelDev lev[NSX_MODAL_MAXLEV];
  float *xh;                                            // N0 x mp modal coefficients
  float *b[NSX_MODAL_MAXLEV], *z[NSX_MODAL_MAXLEV], *x[NSX_MODAL_MAXLEV];  // n_l x m0 per level
};
