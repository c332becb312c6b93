// nsx_context.hip -- the C ABI of include/nsx.h: one device context holding
// geometry, matrix and every field resident in HBM, the stage/step drivers
// and the single-routine entry points.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "nsx.h"
#include "nsx_kernels.h"

namespace {

enum FieldKind { FK_NONE = 0, FK_CELL = 1, FK_VERT = 2, FK_FACE = 3 };
constexpr int NSLOT = 128;
constexpr int MAXGRID = 2048;

struct Timer {
  double ms = 0.0;
  long long n = 0;
};

// Work fields of the differentiation form, allocated on its first use: the
// checkpoints, the stage inputs of one segment, the bars of the adjoint
// sweep and the tangents of the tangent-linear run.  A state is nine fields:
// u, v, w, p and the five face velocities.
struct AdjWork {
  std::vector<double *> ckpt;  // checkpoints, 9 fields each
  std::vector<double *> seg;   // per step of a segment: S^n then S^a, 9 fields each
  double *fin[9] = {};         // the resident state at the end of the forward run
  double *Sb[9] = {};          // bars of the state at the current level
  double *Yb[4] = {}, *Fb[5] = {};
  double *Xb[2][4] = {}, *Fib[2][5] = {};
  double *usb[3] = {}, *eb = nullptr, *lam = nullptr, *E = nullptr;
  double *vb[3] = {}, *pvb = nullptr, *pvb2 = nullptr, *valb = nullptr;  // vertex fields
  double *V[3] = {};                                                     // vertex fields
  // tangents: u, v, w, p by pool set, the faces, and the stage's work fields
  double *T[3][4] = {}, *TF[5] = {}, *TV[3] = {}, *Tpv = nullptr, *TE = nullptr, *Tus[3] = {},
         *Trhs = nullptr, *zero = nullptr;
  double *J = nullptr, *Jpart = nullptr;  // two accumulators, 64 partials
  bool ready = false;
  bool unconverged = false;  // a solve of the current run reached MaxIters
};

}  // namespace

struct nsx_ctx {
  int N0 = 0, NC = 0, NV = 0, NZ = 0, NS = 0, NEX = 0;
  int NGW = 0;                 // wall ghost columns after the N0 interior ones
  bool has_inflow = false;     // a ghost or vertex flagged inflow (2)
  std::vector<int> gw_ghost;   // the reference's ghost row (0-based) of each wall ghost column
  int dev = 0;
  hipStream_t stream = nullptr;
  std::string err;
  bool committed = false;
  // set-up arrays (host, reference layout) kept until commit
  std::vector<int> No_cp, No_vp, nbe, nbev;
  std::vector<double> areaCell, dlVV, dlCC, dxCC, dyCC, xnn, ynn, dle1, dle2, doe, sum_xc2, sum_yc2, sum_xcyc;
  double TotalVolume = 0.0;
  std::vector<double> sig, dsig, sigv, dsigv;
  double sigIni = 0.0, sigFin = 1.0;
  std::vector<int> surrounding, Dimsurrounding, vertex_pair, vertex_edge;
  std::vector<double> weight, dlVsum;
  std::vector<int> bcc[4], bcv[4];  // u,v,w,p
  int bctop = 1, bcbot = 1, TopBC = 0, BotBC = 0;
  std::vector<double> AmC[6];
  std::vector<double> walld;
  double dt = 0, Re = 1, bdx = 0, bdy = 0, bdz = 0, eps = 1e-10;
  int MaxIters = 20000, FinRK = 2, Iturbu = 1;
  unsigned have = 0;
  // device
  Geo g{};
  std::vector<void *> allocs;
  double *slot[NSLOT] = {};
  int kind[NSLOT] = {};
  double *pool_cell[3][4] = {};  // three time-level sets of (u,v,w,p)
  int set_n = 0;                 // pool set holding time level n
  double *ustar[3] = {}, *uv2[3] = {}, *rhsp = nullptr;
  // FP64 BiCGSTAB vectors: rt = r/A0, rhat, q = (p - omega v)/A0, vt = v/A0, t = A z
  double *sr = nullptr, *srhat = nullptr, *sq = nullptr, *svt = nullptr, *stv = nullptr;
  // NSX_PREC_FP64_CG: the single-reduction CG's second buffer set, with the
  // BiCGSTAB vectors above as its first (a solve uses one solver at a time)
  double *scg[4] = {};
  // differentiation form: the body-force switch and its work fields
  bool body_force = false;
  AdjWork adj;
  SolverState *st_dev = nullptr, *st_host = nullptr;
  Partials part{};
  nsx_solve_info info{};
  int iter_hint = 0;
  // mixed precision: FP32 CG vectors and settings
  float *fr = nullptr, *frhat = nullptr, *fp = nullptr, *fv = nullptr, *ft = nullptr, *fe = nullptr;
  float *fe2 = nullptr, *fr2 = nullptr, *fp2 = nullptr;  // FP32 CG: the second set of e, rt, p
  double *cfl_part = nullptr, *cfl_out = nullptr;  // glob_cfl partials and result
  // channel statistics (stats.F90), allocated by nsx_sm_init
  double *sm[SM_N] = {};
  double *sm_planes = nullptr;  // SM_NPLANE x NZ
  int sm_n = 0;
  int precision = NSX_PREC_FP64;
  int refine = NSX_REFINE_CGSR_BLOCK;
  double delta_cg = 0.1;  // reliable-update fraction of the CG forms
  // initial guess of nsx_step's pressure solves in the precision forms
  // (nsx_set_initial_guess): the last two solutions of each RK stage, which
  // ring slot takes the next one, how many are stored, the extrapolation,
  // and the stage nsx_step is running (-1 outside it: the stage's input
  // pressure, as the reference starts)
  int guess_mode = NSX_GUESS_PREVIOUS;
  double *phist[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  int hslot[2] = {0, 0}, nhist[2] = {0, 0};
  double *pguess = nullptr;
  int guess_stage = -1;
  // the vertical-mode preconditioner (NSX_REFINE_MODAL): the column graph
  // kept from nsx_commit, the device tables once built
  std::vector<int> h_nbr;
  std::vector<double> h_ppec;
  ModalDev modal{};
  bool modal_ok = false;
  int iter_hint_cg = 0;
  int iter_hint_cgs = 0;
  int iter_hint_cgs64 = 0;
  // timing
  bool timing = false;
  std::vector<hipEvent_t> evpool;
  size_t evused = 0;
  Timer t_k[3], t_stage, t_cg[2];  // FP64 A, B, C; stage kernels; FP32 CG W, X
  Timer t_cgs, t_cgs64;              // single-reduction CG, FP32 and FP64
  Timer t_mod[5];                    // vertical-mode CG: forward transform, V-cycle, backward, product, update
  // (event index, timer: 0..2 FP64 solver, 5 stage, 12..13 FP32 CG, 14 single-reduction CG, 15 its FP64 instance)
  std::vector<std::pair<size_t, int>> pending;
};

static int nsx_fail(nsx_ctx *ctx, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (ctx) ctx->err = buf;
  return code;
}

static int nsx_fail_hip(nsx_ctx *ctx, hipError_t e, const char *what, int line) {
  return nsx_fail(ctx, -2, "HIP error %d (%s) at nsx_context.hip:%d: %s", (int)e, hipGetErrorString(e), line, what);
}

#define LAUNCH(x)                                                                     \
  do {                                                                                \
    int e_ = (x);                                                                     \
    if (e_ != 0) return nsx_fail_hip(ctx, (hipError_t)e_, #x, __LINE__);              \
  } while (0)

#define NEED_COMMIT(ctx)                                                              \
  do {                                                                                \
    if (!(ctx)) return -1;                                                            \
    if (!(ctx)->committed) return nsx_fail(ctx, -1, "nsx_commit has not been called"); \
  } while (0)

template <class T>
static int dev_alloc(nsx_ctx *ctx, T **p, size_t n) {
  void *q = nullptr;
  hipError_t e = hipMalloc(&q, (n ? n : 1) * sizeof(T));
  if (e != hipSuccess) return nsx_fail(ctx, -3, "hipMalloc of %zu bytes failed: %s", n * sizeof(T), hipGetErrorString(e));
  ctx->allocs.push_back(q);
  *p = (T *)q;
  return 0;
}

template <class T>
static int dev_upload(nsx_ctx *ctx, T **p, const std::vector<T> &h) {
  int rc = dev_alloc(ctx, p, h.size());
  if (rc) return rc;
  HIPCHECK(hipMemcpy(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

// cell fields hold the N0 interior columns and the wall ghost columns; face
// fields the interior columns
static size_t field_elems(const nsx_ctx *c, int kind) {
  if (kind == FK_VERT) return (size_t)c->NV * (c->NZ - 1);
  if (kind == FK_CELL) return (size_t)(c->N0 + c->NGW) * c->NS;
  return (size_t)c->N0 * c->NS;
}

static size_t cell_bytes(const nsx_ctx *ctx) { return field_elems(ctx, FK_CELL) * sizeof(double); }

// the stored pressure solutions of nsx_set_initial_guess no longer lead to
// the resident state: start again from two new ones
static void reset_guess(nsx_ctx *ctx) { ctx->nhist[0] = ctx->nhist[1] = ctx->hslot[0] = ctx->hslot[1] = 0; }

// ------------------------------------------------------------- lifecycle --
extern "C" int nsx_create(int n_cell0, int n_cell, int n_vert, int nz, nsx_ctx **out) {
  if (!out) return -1;
  *out = nullptr;
  if (n_cell0 <= 0 || n_cell < n_cell0 || n_vert <= 0 || nz < 4) return -1;
  nsx_ctx *ctx = new nsx_ctx();
  ctx->N0 = n_cell0;
  ctx->NC = n_cell;
  ctx->NV = n_vert;
  ctx->NZ = nz;
  ctx->NS = (nz + 3) / 4 * 4;
  hipError_t e = hipGetDevice(&ctx->dev);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete ctx;
    return -2;
  }
  *out = ctx;
  return 0;
}

extern "C" int nsx_destroy(nsx_ctx *ctx) {
  if (!ctx) return 0;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (void *p : ctx->allocs) (void)hipFree(p);
  if (ctx->st_host) (void)hipHostFree(ctx->st_host);
  for (hipEvent_t ev : ctx->evpool) (void)hipEventDestroy(ev);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return 0;
}

extern "C" const char *nsx_last_error(const nsx_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }
extern "C" int nsx_device(const nsx_ctx *ctx) { return ctx ? ctx->dev : -1; }

// ---------------------------------------------------------------- set-up --
template <class T>
static void take(std::vector<T> &v, const T *p, size_t n) {
  v.assign(p, p + n);
}

extern "C" int nsx_set_mesh(nsx_ctx *ctx, const int *No_cp, const int *No_vp, const int *nbe, const int *nbev) {
  if (!ctx || !No_cp || !No_vp || !nbe || !nbev) return -1;
  take(ctx->No_cp, No_cp, (size_t)ctx->NC * 3);
  take(ctx->No_vp, No_vp, (size_t)ctx->N0 * 3);
  take(ctx->nbe, nbe, (size_t)ctx->N0);
  take(ctx->nbev, nbev, (size_t)ctx->NV);
  ctx->have |= 1;
  return 0;
}

extern "C" int nsx_set_geometry(nsx_ctx *ctx, const double *areaCell, const double *dlVV, const double *dlCC,
                                const double *dxCC, const double *dyCC, const double *xnn, const double *ynn,
                                const double *dle1, const double *dle2, const double *doe, const double *sum_xc2,
                                const double *sum_yc2, const double *sum_xcyc, double TotalVolume) {
  if (!ctx) return -1;
  const size_t n3 = (size_t)ctx->N0 * 3;
  take(ctx->areaCell, areaCell, (size_t)ctx->NC);
  take(ctx->dlVV, dlVV, n3);
  take(ctx->dlCC, dlCC, n3);
  take(ctx->dxCC, dxCC, n3);
  take(ctx->dyCC, dyCC, n3);
  take(ctx->xnn, xnn, n3);
  take(ctx->ynn, ynn, n3);
  take(ctx->dle1, dle1, n3);
  take(ctx->dle2, dle2, n3);
  take(ctx->doe, doe, n3);
  take(ctx->sum_xc2, sum_xc2, (size_t)ctx->N0);
  take(ctx->sum_yc2, sum_yc2, (size_t)ctx->N0);
  take(ctx->sum_xcyc, sum_xcyc, (size_t)ctx->N0);
  ctx->TotalVolume = TotalVolume;
  ctx->have |= 2;
  return 0;
}

extern "C" int nsx_set_sigma(nsx_ctx *ctx, const double *sig, const double *dsig, const double *sigv,
                             const double *dsigv, double sigIni, double sigFin) {
  if (!ctx) return -1;
  take(ctx->sig, sig, (size_t)ctx->NZ);
  take(ctx->dsig, dsig, (size_t)ctx->NZ);
  take(ctx->sigv, sigv, (size_t)ctx->NZ - 1);
  take(ctx->dsigv, dsigv, (size_t)ctx->NZ - 1);
  ctx->sigIni = sigIni;
  ctx->sigFin = sigFin;
  ctx->have |= 4;
  return 0;
}

extern "C" int nsx_set_vertex_weights(nsx_ctx *ctx, const int *surrounding, const double *weight,
                                      const int *Dimsurrounding, const double *dlVsum, const int *vertex_pair,
                                      const int *vertex_edge) {
  if (!ctx) return -1;
  take(ctx->surrounding, surrounding, (size_t)ctx->NV * 10);
  take(ctx->weight, weight, (size_t)ctx->NV * 10);
  take(ctx->Dimsurrounding, Dimsurrounding, (size_t)ctx->NV);
  take(ctx->dlVsum, dlVsum, (size_t)ctx->NV);
  take(ctx->vertex_pair, vertex_pair, (size_t)ctx->NV);
  take(ctx->vertex_edge, vertex_edge, (size_t)ctx->NV);
  ctx->have |= 8;
  return 0;
}

extern "C" int nsx_set_bc(nsx_ctx *ctx, const int *bcuc, const int *bcvc, const int *bcwc, const int *bcpc,
                          const int *bcuv, const int *bcvv, const int *bcwv, const int *bcpv, int bctop, int bcbot,
                          int TopBC, int BotBC) {
  if (!ctx) return -1;
  const int *c[4] = {bcuc, bcvc, bcwc, bcpc}, *v[4] = {bcuv, bcvv, bcwv, bcpv};
  for (int q = 0; q < 4; ++q) {
    take(ctx->bcc[q], c[q], (size_t)ctx->NC);
    take(ctx->bcv[q], v[q], (size_t)ctx->NV);
  }
  if (TopBC < 0 || TopBC > 2 || BotBC < 0 || BotBC > 1)
    return nsx_fail(ctx, -1, "TopBC must be 0 (no slip), 1 (free slip) or 2 (the lid), BotBC 0 or 1, got %d/%d",
                    TopBC, BotBC);
  ctx->bctop = bctop;
  ctx->bcbot = bcbot;
  ctx->TopBC = TopBC;
  ctx->BotBC = BotBC;
  ctx->have |= 16;
  return 0;
}

extern "C" int nsx_set_ppe_matrix(nsx_ctx *ctx, const double *A0, const double *A1, const double *A2,
                                  const double *A3, const double *AT, const double *AB) {
  if (!ctx) return -1;
  const double *a[6] = {A0, A1, A2, A3, AT, AB};
  for (int q = 0; q < 6; ++q) take(ctx->AmC[q], a[q], (size_t)ctx->N0 * ctx->NZ);
  ctx->have |= 32;
  return 0;
}

extern "C" int nsx_set_les(nsx_ctx *ctx, const double *walld_level) {
  if (!ctx) return -1;
  take(ctx->walld, walld_level, (size_t)ctx->NZ);
  ctx->have |= 64;
  return 0;
}

extern "C" int nsx_set_params(nsx_ctx *ctx, double dt, double Re, double bdx, double bdy, double bdz, double eps,
                              int MaxIters, int FinRK, int Iturbu) {
  if (!ctx) return -1;
  if (FinRK != 1 && FinRK != 2) return nsx_fail(ctx, -1, "FinRK must be 1 or 2");
  ctx->dt = dt;
  ctx->Re = Re;
  ctx->bdx = bdx;
  ctx->bdy = bdy;
  ctx->bdz = bdz;
  ctx->eps = eps;
  ctx->MaxIters = MaxIters;
  ctx->FinRK = FinRK;
  ctx->Iturbu = Iturbu;
  ctx->g.dt = dt;
  ctx->g.nu = 1. / Re;
  ctx->g.bdx = bdx;
  ctx->g.bdy = bdy;
  ctx->g.bdz = bdz;
  ctx->have |= 128;
  return 0;
}

// (N0,3) column-major -> [i*3+j]
static std::vector<double> edge_major(const std::vector<double> &a, int N0) {
  std::vector<double> o((size_t)N0 * 3);
  for (int i = 0; i < N0; ++i)
    for (int j = 0; j < 3; ++j) o[(size_t)i * 3 + j] = a[(size_t)i + (size_t)j * N0];
  return o;
}

extern "C" int nsx_commit(nsx_ctx *ctx) {
  if (!ctx) return -1;
  if (ctx->committed) return nsx_fail(ctx, -1, "context already committed");
  if ((ctx->have & 255) != 255)
    return nsx_fail(ctx, -1, "set-up incomplete (mask %u): mesh, geometry, sigma, vertex weights, bc, ppe matrix, les, params",
                    ctx->have);
  const int N0 = ctx->N0, NC = ctx->NC, NV = ctx->NV, NZ = ctx->NZ, NS = ctx->NS;
  Geo &g = ctx->g;
  g.N0 = N0;
  g.NV = NV;
  g.NZ = NZ;
  g.NS = NS;
  g.NZi = NZ - 2;
  g.NL = NZ - 1;
  g.ncell = (uint32_t)N0 * (uint32_t)(NZ - 2);
  if ((uint64_t)N0 * NS >= (1ull << 31) || (uint64_t)NV * (NZ - 1) >= (1ull << 31))
    return nsx_fail(ctx, -1, "mesh too large for 32-bit cell numbering");
  g.divNZi = make_fastdiv((uint32_t)g.NZi);
  g.divNL = make_fastdiv((uint32_t)g.NL);
  g.NQ = NS / 4;
  g.nquad = (uint32_t)N0 * (uint32_t)g.NQ;
  g.divNQ = make_fastdiv((uint32_t)g.NQ);
  {
    // one grid for the solver's partial-sum chain: every block resident at
    // once (CUs x the lowest occupancy of those kernels), at most MAXGRID
    hipDeviceProp_t prop;
    HIPCHECK(hipGetDeviceProperties(&prop, ctx->dev));
    const int per_cu = solver_blocks_per_cu(g);
    if (per_cu < 1) return nsx_fail(ctx, -2, "occupancy query of the solver kernels failed");
    long G = (long)prop.multiProcessorCount * per_cu;
    const long need = ((long)g.nquad + NSX_BLOCK - 1) / NSX_BLOCK;
    if (G > need) G = need;
    if (G > MAXGRID) G = MAXGRID;
    G = G / NSX_NXCD * NSX_NXCD;
    if (G < NSX_NXCD) G = NSX_NXCD;
    g.grid = (int)G;
    const int per_cu_cg = cg_blocks_per_cu(g);
    if (per_cu_cg < 1) return nsx_fail(ctx, -2, "occupancy query of the CG kernels failed");
    long Gc = (long)prop.multiProcessorCount * per_cu_cg;
    if (Gc > need) Gc = need;
    if (Gc > MAXGRID) Gc = MAXGRID;
    Gc = Gc / NSX_NXCD * NSX_NXCD;
    if (Gc < NSX_NXCD) Gc = NSX_NXCD;
    g.grid_cg = (int)Gc;
    // the single-reduction CG kernel runs 4 blocks a CU, below its occupancy
    // (5): at 6.3M prisms 1024 blocks sweep 6 windows each, a band per XCD
    // that its L2 holds, and took 50.5 us a launch against 52.2 us for 1280
    // (896: 52.5, 1152: 51.8)
    const int per_cu_cgs = std::min(cgs_blocks_per_cu(g), 4);
    if (per_cu_cgs < 1) return nsx_fail(ctx, -2, "occupancy query of the single-reduction CG kernel failed");
    long Gs = (long)prop.multiProcessorCount * per_cu_cgs;
    if (Gs > need) Gs = need;
    if (Gs > MAXGRID) Gs = MAXGRID;
    Gs = Gs / NSX_NXCD * NSX_NXCD;
    if (Gs < NSX_NXCD) Gs = NSX_NXCD;
    g.grid_cgs = (int)Gs;
    // its FP64 instance: every block its occupancy allows, at most 4 a CU
    const int per_cu_64 = std::min(cgs64_blocks_per_cu(g), 4);
    if (per_cu_64 < 1) return nsx_fail(ctx, -2, "occupancy query of the FP64 single-reduction CG kernel failed");
    long G6 = (long)prop.multiProcessorCount * per_cu_64;
    if (G6 > need) G6 = need;
    if (G6 > MAXGRID) G6 = MAXGRID;
    G6 = G6 / NSX_NXCD * NSX_NXCD;
    if (G6 < NSX_NXCD) G6 = NSX_NXCD;
    g.grid_cgs64 = (int)G6;
  }
  if (NS > 4096) return nsx_fail(ctx, -1, "NZ = %d exceeds the 4094 layers the solver's LDS level table holds", NZ);
  g.bctop = ctx->bctop;
  g.bcbot = ctx->bcbot;
  g.TopBC = ctx->TopBC;
  g.BotBC = ctx->BotBC;
  g.TotalVolume = ctx->TotalVolume;
  g.sigIni = ctx->sigIni;
  g.sigFin = ctx->sigFin;

  // neighbours: periodic ghosts resolved to their partner columns, wall
  // ghosts (flags 0 and 1 of update_bc.F90) given columns of their own after
  // the interior ones
  std::vector<int> nbr((size_t)N0 * 3), vtx((size_t)N0 * 3);
  std::vector<signed char> eside((size_t)N0 * 3, 0);
  std::vector<int> gw_owner;
  std::vector<unsigned char> gw_neu, gw_inf;
  ctx->gw_ghost.clear();
  int nex = N0;
  for (int i = 0; i < N0; ++i)
    for (int j = 0; j < 3; ++j) {
      int jc = ctx->No_cp[(size_t)i + (size_t)j * NC] - 1;
      if (jc < 0 || jc >= NC) return nsx_fail(ctx, -1, "No_cp(%d,%d) = %d out of range", i + 1, j + 1, jc + 1);
      if (jc >= N0) {
        eside[(size_t)i * 3 + j] = 1;
        if (jc + 1 > nex) nex = jc + 1;
        const int fu = ctx->bcc[0][jc], fv = ctx->bcc[1][jc], fw = ctx->bcc[2][jc], fp = ctx->bcc[3][jc];
        if (fu == 3 && fv == 3 && fw == 3 && fp == 3) {
          int partner = ctx->No_cp[(size_t)jc + 2 * (size_t)NC] - 1;
          if (partner < 0 || partner >= N0) return nsx_fail(ctx, -1, "ghost cell %d has no periodic partner", jc + 1);
          jc = partner;
        } else if (fu >= 0 && fu <= 2 && fv >= 0 && fv <= 2 && fw >= 0 && fw <= 2 && fp == 1) {
          // bit q: flag 1 (copy), resp. 2 (inflow), for kind q (0 p, 1 u, 2 v, 3 w)
          gw_neu.push_back((unsigned char)(1 | ((fu == 1) << 1) | ((fv == 1) << 2) | ((fw == 1) << 3)));
          gw_inf.push_back((unsigned char)(((fu == 2) << 1) | ((fv == 2) << 2) | ((fw == 2) << 3)));
          gw_owner.push_back(i);
          ctx->gw_ghost.push_back(jc);
          jc = N0 + (int)gw_owner.size() - 1;
        } else {
          return nsx_fail(ctx, -1,
                          "ghost cell %d has flags u %d v %d w %d p %d: supported are periodic ghosts (all 3) and "
                          "walls (u, v, w 0, 1 or 2, p 1); parallel (-1) ghosts are not",
                          jc + 1, fu, fv, fw, fp);
        }
      }
      nbr[(size_t)i * 3 + j] = jc;
      int v = ctx->No_vp[(size_t)i + (size_t)j * N0] - 1;
      if (v < 0 || v >= NV) return nsx_fail(ctx, -1, "No_vp(%d,%d) out of range", i + 1, j + 1);
      vtx[(size_t)i * 3 + j] = v;
    }
  ctx->NEX = nex;
  ctx->NGW = (int)gw_owner.size();
  ctx->has_inflow = false;
  for (unsigned char b : gw_inf) ctx->has_inflow = ctx->has_inflow || b != 0;
  for (int q = 0; q < 3; ++q)
    for (int v = 0; v < NV; ++v) ctx->has_inflow = ctx->has_inflow || (ctx->nbev[v] != 0 && ctx->bcv[q][v] == 2);
  if ((uint64_t)(N0 + ctx->NGW) * NS >= (1ull << 31))
    return nsx_fail(ctx, -1, "the mesh with its wall ghost columns exceeds 32-bit cell numbering");

  std::vector<double> ppec((size_t)N0 * 3);
  for (int i = 0; i < N0; ++i)
    for (int j = 0; j < 3; ++j) {
      const size_t f = (size_t)i + (size_t)j * N0;
      const double co1 = ctx->dle1[f] / ctx->dlCC[f];
      int jc = ctx->No_cp[(size_t)i + (size_t)j * NC] - 1;
      ppec[(size_t)i * 3 + j] = (ctx->bcc[3][jc] == 1) ? 0.0 : co1 * ctx->dlVV[f];
    }
  ctx->h_nbr = nbr;
  ctx->h_ppec = ppec;
  // A^T: the lateral coefficient of row i towards neighbour n is row n's
  // coefficient towards i (the edge of n whose neighbour is i; where several
  // edges of n lead to i, the one with the closest coefficient)
  std::vector<double> ppec_t((size_t)N0 * 3);
  for (int i = 0; i < N0; ++i)
    for (int j = 0; j < 3; ++j) {
      const int n = nbr[(size_t)i * 3 + j];
      if (n >= N0) {
        // a wall: the coupling is zero both ways (Neumann p, flag 1)
        ppec_t[(size_t)i * 3 + j] = ppec[(size_t)i * 3 + j];
        continue;
      }
      int best = -1;
      for (int jj = 0; jj < 3; ++jj)
        if (nbr[(size_t)n * 3 + jj] == i &&
            (best < 0 || std::fabs(ppec[(size_t)n * 3 + jj] - ppec[(size_t)i * 3 + j]) <
                             std::fabs(ppec[(size_t)n * 3 + best] - ppec[(size_t)i * 3 + j])))
          best = jj;
      if (best < 0) return nsx_fail(ctx, -1, "cell %d is not a neighbour of its neighbour %d", i + 1, n + 1);
      ppec_t[(size_t)i * 3 + j] = ppec[(size_t)n * 3 + best];
    }
  // the compact matrix must reproduce PDE_center's coefficients bit for bit
  for (int k = 1; k < NZ - 1; ++k)
    for (int i = 0; i < N0; ++i) {
      const size_t n = (size_t)i + (size_t)k * N0;
      const double sv = ctx->dsigv[k - 1];
      const double a1 = ppec[(size_t)i * 3] * sv, a2 = ppec[(size_t)i * 3 + 1] * sv, a3 = ppec[(size_t)i * 3 + 2] * sv;
      double AT = ctx->areaCell[i] / ctx->dsig[k], AB = ctx->areaCell[i] / ctx->dsig[k - 1];
      if (k == 1 && ctx->bcbot == 1) AB = 0.0;
      if (k == NZ - 2 && ctx->bctop == 1) AT = 0.0;
      const double A0 = -1.0 * (AT + AB + a1 + a2 + a3);
      const double ref[6] = {ctx->AmC[0][n], ctx->AmC[1][n], ctx->AmC[2][n], ctx->AmC[3][n], ctx->AmC[4][n],
                             ctx->AmC[5][n]};
      const double mine[6] = {A0, a1, a2, a3, AT, AB};
      for (int q = 0; q < 6; ++q)
        if (!(ref[q] == mine[q]))
          return nsx_fail(ctx, -1, "PPE coefficient %d at (%d,%d) is %.17g, PDE_center's form gives %.17g", q, i + 1,
                          k + 1, ref[q], mine[q]);
    }

  // vertex stencil (CSR) and periodic data
  std::vector<int> vptr(NV + 1), vcell;
  std::vector<double> vw;
  std::vector<signed char> vkind(NV), vact[4];
  for (int q = 0; q < 4; ++q) vact[q].assign(NV, VA_NONE);
  vptr[0] = 0;
  g.ncorner = 0;
  for (int v = 0; v < NV; ++v) {
    const int d = ctx->Dimsurrounding[v];
    if (d < 0 || d > 10) return nsx_fail(ctx, -1, "Dimsurrounding(%d) = %d", v + 1, d);
    for (int j = 0; j < d; ++j) {
      int c = ctx->surrounding[(size_t)v + (size_t)j * NV] - 1;
      if (c < 0 || c >= N0) return nsx_fail(ctx, -1, "surrounding(%d,%d) is not an interior cell", v + 1, j + 1);
      vcell.push_back(c);
      vw.push_back(ctx->weight[(size_t)v + (size_t)j * NV]);
    }
    vptr[v + 1] = (int)vcell.size();
    if (ctx->vertex_edge[v] == 1 && g.ncorner < 4) g.corner[g.ncorner++] = v;
    if (ctx->nbev[v] == 0) {
      vkind[v] = VK_INTERIOR;
      continue;
    }
    vkind[v] = ctx->vertex_edge[v] == 1 ? VK_CORNER : (ctx->vertex_pair[v] > 0 ? VK_PAIR : VK_OTHER);
    for (int q = 0; q < 4; ++q) {
      const int fl = ctx->bcv[q][v];
      vact[q][v] = fl == 0 ? VA_ZERO : (fl == 2 ? VA_INFLOW : (fl == 3 ? VA_PERIODIC : VA_NONE));
      if (fl == 2 && q == 3) return nsx_fail(ctx, -1, "vertex %d: a pressure vertex flagged inflow", v + 1);
    }
  }
  std::vector<int> vpair(NV);
  for (int v = 0; v < NV; ++v) {
    vpair[v] = ctx->vertex_pair[v] > 0 ? ctx->vertex_pair[v] - 1 : -1;
    if (vpair[v] >= NV) return nsx_fail(ctx, -1, "vertex_pair(%d) out of range", v + 1);
    if (vkind[v] == VK_PAIR && vpair[v] < 0) return nsx_fail(ctx, -1, "vertex %d has no pair", v + 1);
  }
  if (g.ncorner != 0 && g.ncorner != 4) return nsx_fail(ctx, -1, "%d corner vertices (expected 0 or 4)", g.ncorner);

  // upload
  int rc = 0;
#define UP(dst, vec)                                              \
  do {                                                            \
    rc = dev_upload(ctx, &dst, vec);                              \
    if (rc) return rc;                                            \
  } while (0)
  int *d_nbr, *d_vtx, *d_vptr, *d_vcell, *d_vpair;
  double *d[16];
  UP(d_nbr, nbr);
  UP(d_vtx, vtx);
  {
    signed char *d_es;
    UP(d_es, eside);
    g.eside = d_es;
  }
  g.NGW = ctx->NGW;
  g.gw_owner = nullptr;
  g.gw_neu = nullptr;
  g.gw_inf = nullptr;
  if (ctx->NGW > 0) {
    int *d_go;
    unsigned char *d_gn, *d_gi;
    UP(d_go, gw_owner);
    UP(d_gn, gw_neu);
    UP(d_gi, gw_inf);
    g.gw_owner = d_go;
    g.gw_neu = d_gn;
    g.gw_inf = d_gi;
  }
  const std::vector<double> *edge[9] = {&ctx->dlVV, &ctx->dlCC, &ctx->dle1, &ctx->dle2, &ctx->doe,
                                        &ctx->xnn, &ctx->ynn, &ctx->dxCC, &ctx->dyCC};
  for (int q = 0; q < 9; ++q) UP(d[q], edge_major(*edge[q], N0));
  UP(d[9], ppec);
  UP(d[10], ppec_t);
  g.nbr = d_nbr;
  g.vtx = d_vtx;
  g.dlVV = d[0];
  g.dlCC = d[1];
  g.dle1 = d[2];
  g.dle2 = d[3];
  g.doe = d[4];
  g.xnn = d[5];
  g.ynn = d[6];
  g.dxCC = d[7];
  g.dyCC = d[8];
  g.ppec = d[9];
  g.ppec_t = d[10];
  g.trans = 0;
  double *a_area, *a_sx, *a_sy, *a_sxy, *a_sig, *a_dsig, *a_sigv, *a_dsigv, *a_walld, *a_vw, *a_dl;
  UP(a_area, std::vector<double>(ctx->areaCell.begin(), ctx->areaCell.begin() + N0));
  UP(a_sx, ctx->sum_xc2);
  UP(a_sy, ctx->sum_yc2);
  UP(a_sxy, ctx->sum_xcyc);
  UP(a_sig, ctx->sig);
  UP(a_dsig, ctx->dsig);
  UP(a_sigv, ctx->sigv);
  UP(a_dsigv, ctx->dsigv);
  UP(a_walld, ctx->walld);
  UP(d_vptr, vptr);
  UP(d_vcell, vcell);
  UP(a_vw, vw);
  UP(a_dl, ctx->dlVsum);
  UP(d_vpair, vpair);
  g.area = a_area;
  g.sum_xc2 = a_sx;
  g.sum_yc2 = a_sy;
  g.sum_xcyc = a_sxy;
  g.sig = a_sig;
  g.dsig = a_dsig;
  g.sigv = a_sigv;
  g.dsigv = a_dsigv;
  g.walld = a_walld;
  {
    // FP32 copies for the inner solver of the mixed-precision form
    auto f32 = [](const std::vector<double> &v, size_t n) {
      std::vector<float> o(n);
      for (size_t q = 0; q < n; ++q) o[q] = (float)v[q];
      return o;
    };
    float *fa, *fpc, *fds, *fdv, *fpt;
    UP(fa, f32(ctx->areaCell, (size_t)N0));
    UP(fpc, f32(ppec, ppec.size()));
    UP(fpt, f32(ppec_t, ppec_t.size()));
    g.ppec_t_f = fpt;
    UP(fds, f32(ctx->dsig, (size_t)NZ));
    UP(fdv, f32(ctx->dsigv, (size_t)NZ - 1));
    g.area_f = fa;
    g.ppec_f = fpc;
    g.dsig_f = fds;
    g.dsigv_f = fdv;
  }
  g.vptr = d_vptr;
  g.vcell = d_vcell;
  g.vw = a_vw;
  g.dlVsum = a_dl;
  g.vpair = d_vpair;
  signed char *d_vkind, *d_act[4];
  UP(d_vkind, vkind);
  for (int q = 0; q < 4; ++q) UP(d_act[q], vact[q]);
  g.vkind = d_vkind;
  g.vact_u = d_act[0];
  g.vact_v = d_act[1];
  g.vact_w = d_act[2];
  g.vact_p = d_act[3];
#undef UP

  // fields (cell fields and solver vectors with the wall ghost columns)
  const size_t nc = (size_t)(N0 + ctx->NGW) * NS, nv = (size_t)NV * (NZ - 1);
  auto cell = [&](double **p) { return dev_alloc(ctx, p, nc); };
  for (int s = 0; s < 3; ++s)
    for (int q = 0; q < 4; ++q)
      if ((rc = cell(&ctx->pool_cell[s][q]))) return rc;
  int ids_cell[] = {NSX_F_RHSUF, NSX_F_RHSVF, NSX_F_RHSWF, NSX_F_EDDY, NSX_F_FORCE};
  for (int id : ids_cell) {
    if ((rc = cell(&ctx->slot[id]))) return rc;
    ctx->kind[id] = FK_CELL;
  }
  for (int id = NSX_F_UFV; id <= NSX_F_PFV; ++id) {
    if ((rc = dev_alloc(ctx, &ctx->slot[id], nv))) return rc;
    ctx->kind[id] = FK_VERT;
  }
  for (int id = NSX_F_U1FACE; id <= NSX_F_UBFACE; ++id) {
    if ((rc = cell(&ctx->slot[id]))) return rc;
    ctx->kind[id] = FK_FACE;
  }
  for (int q = 0; q < NSX_N_CELL_SCRATCH; ++q) {
    if ((rc = cell(&ctx->slot[NSX_F_CELL_SCRATCH + q]))) return rc;
    ctx->kind[NSX_F_CELL_SCRATCH + q] = FK_CELL;
  }
  for (int q = 0; q < NSX_N_VERT_SCRATCH; ++q) {
    if ((rc = dev_alloc(ctx, &ctx->slot[NSX_F_VERT_SCRATCH + q], nv))) return rc;
    ctx->kind[NSX_F_VERT_SCRATCH + q] = FK_VERT;
  }
  for (int q = 0; q < NSX_N_FACE_SCRATCH; ++q) {
    if ((rc = cell(&ctx->slot[NSX_F_FACE_SCRATCH + q]))) return rc;
    ctx->kind[NSX_F_FACE_SCRATCH + q] = FK_FACE;
  }
  for (int q = 0; q < 3; ++q) {
    if ((rc = cell(&ctx->ustar[q]))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->uv2[q], nv))) return rc;
  }
  if ((rc = cell(&ctx->rhsp)) || (rc = cell(&ctx->sr)) || (rc = cell(&ctx->srhat)) || (rc = cell(&ctx->sq)) ||
      (rc = cell(&ctx->svt)) || (rc = cell(&ctx->stv)))
    return rc;
  for (int q = 0; q < 4; ++q)
    if ((rc = cell(&ctx->scg[q]))) return rc;
  // the time-level slots point into pool set 0 (n), 1 (stage), 2 (new)
  ctx->set_n = 0;
  for (int q = 0; q < 4; ++q) {
    ctx->slot[NSX_F_UFN + q] = ctx->pool_cell[0][q];
    ctx->slot[NSX_F_UF + q] = ctx->pool_cell[1][q];
    ctx->slot[NSX_F_UFNP + q] = ctx->pool_cell[2][q];
    ctx->kind[NSX_F_UFN + q] = ctx->kind[NSX_F_UF + q] = ctx->kind[NSX_F_UFNP + q] = FK_CELL;
  }
  // every field starts at zero, like the zero-filled work arrays of the reference
  for (int st = 0; st < 3; ++st)
    for (int q = 0; q < 4; ++q) HIPCHECK(hipMemsetAsync(ctx->pool_cell[st][q], 0, nc * sizeof(double), ctx->stream));
  for (int id = 0; id < NSLOT; ++id)
    if (ctx->slot[id] && ctx->kind[id] != FK_NONE && !(id >= NSX_F_UFN && id <= NSX_F_PFNP))
      HIPCHECK(hipMemsetAsync(ctx->slot[id], 0, field_elems(ctx, ctx->kind[id]) * sizeof(double), ctx->stream));
  double *zs[] = {ctx->ustar[0], ctx->ustar[1], ctx->ustar[2], ctx->rhsp,      ctx->sr,        ctx->srhat, ctx->sq,
                  ctx->svt,      ctx->stv,      ctx->scg[0],    ctx->scg[1], ctx->scg[2], ctx->scg[3]};
  for (double *p : zs) HIPCHECK(hipMemsetAsync(p, 0, nc * sizeof(double), ctx->stream));
  for (int q = 0; q < 3; ++q) HIPCHECK(hipMemsetAsync(ctx->uv2[q], 0, nv * sizeof(double), ctx->stream));
  float **fvec[9] = {&ctx->fr, &ctx->frhat, &ctx->fp, &ctx->fv, &ctx->ft, &ctx->fe, &ctx->fe2, &ctx->fr2, &ctx->fp2};
  for (float **f : fvec) {
    if ((rc = dev_alloc(ctx, f, nc))) return rc;
    HIPCHECK(hipMemsetAsync(*f, 0, nc * sizeof(float), ctx->stream));
  }

  // solver state and reduction scratch
  if ((rc = dev_alloc(ctx, &ctx->st_dev, 1))) return rc;
  HIPCHECK(hipMemsetAsync(ctx->st_dev, 0, sizeof(SolverState), ctx->stream));
  HIPCHECK(hipHostMalloc((void **)&ctx->st_host, sizeof(SolverState), hipHostMallocDefault));
  // one buffer of [rows][MAXGRID] per reduction site; every member of
  // Partials must be listed here (checked below)
  double **rows[6] = {&ctx->part.src, &ctx->part.init, &ctx->part.a, &ctx->part.b, &ctx->part.c, &ctx->part.u};
  const int nrow[6] = {1, 5, 1, 2, 4, 4};
  static_assert(sizeof(Partials) == 6 * sizeof(double *), "a Partials member has no buffer");
  for (int q = 0; q < 6; ++q) {
    if ((rc = dev_alloc(ctx, rows[q], (size_t)nrow[q] * MAXGRID))) return rc;
    HIPCHECK(hipMemsetAsync(*rows[q], 0, (size_t)nrow[q] * MAXGRID * sizeof(double), ctx->stream));
  }
  for (int q = 0; q < 6; ++q)
    if (!*rows[q]) return nsx_fail(ctx, -1, "reduction buffer %d is not allocated", q);
  if ((rc = dev_alloc(ctx, &ctx->cfl_part, (size_t)glob_cfl_rows() * MAXGRID))) return rc;
  if ((rc = dev_alloc(ctx, &ctx->cfl_out, 16))) return rc;
  // the static scratch of nsx_glob_cfl_noalloc on this device
  {
    const int e = glob_cfl_scratch_init(ctx->dev);
    if (e != 0) return nsx_fail_hip(ctx, (hipError_t)e, "hipGetSymbolAddress of the glob_cfl scratch", __LINE__);
  }

  // the eddy viscosity's geometric factor (Cs Delta)^2 per prism
  {
    double *nutg = nullptr;
    if ((rc = dev_alloc(ctx, &nutg, (size_t)N0 * NS))) return rc;
    g.nutg = nutg;
    LAUNCH(launch_les_length(g, ctx->stream, nutg));
  }
  // compact PPE operator, and a device-side check of its diagonal
  double *A0chk = nullptr;
  if ((rc = cell(&A0chk))) return rc;
  LAUNCH(launch_ppe_diag(g, ctx->stream, A0chk));
  std::vector<double> a0((size_t)nc);
  HIPCHECK(hipMemcpyAsync(a0.data(), A0chk, nc * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHECK(hipStreamSynchronize(ctx->stream));
  for (int k = 1; k < NZ - 1; ++k)
    for (int i = 0; i < N0; ++i)
      if (!(a0[(size_t)i * NS + k] == ctx->AmC[0][(size_t)i + (size_t)k * N0]))
        return nsx_fail(ctx, -1, "device PPE diagonal differs from AmC0 at (%d,%d)", i + 1, k + 1);
  ctx->committed = true;
  // the host copies are no longer needed except the ghost map of nsx_get_field
  for (int q = 0; q < 6; ++q) std::vector<double>().swap(ctx->AmC[q]);
  return 0;
}

// ---------------------------------------------------------------- fields --
static int check_field(nsx_ctx *ctx, int f) {
  if (f < 0 || f >= NSLOT || !ctx->slot[f]) return nsx_fail(ctx, -1, "unknown field id %d", f);
  return 0;
}

extern "C" int nsx_put_field(nsx_ctx *ctx, int f, const double *host) {
  NEED_COMMIT(ctx);
  if (check_field(ctx, f)) return -1;
  reset_guess(ctx);  // a new state: the stored solutions no longer lead to it
  const int N0 = ctx->N0, NZ = ctx->NZ, NS = ctx->NS, NV = ctx->NV;
  std::vector<double> buf(field_elems(ctx, ctx->kind[f]), 0.0);
  if (ctx->kind[f] == FK_VERT) {
    for (int l = 0; l < NZ - 1; ++l)
      for (int v = 0; v < NV; ++v) buf[(size_t)v * (NZ - 1) + l] = host[(size_t)v + (size_t)l * NV];
  } else {
    const size_t ld = ctx->kind[f] == FK_FACE ? (size_t)N0 : (size_t)ctx->NC;
    for (int k = 0; k < NZ; ++k)
      for (int i = 0; i < N0; ++i) buf[(size_t)i * NS + k] = host[(size_t)i + (size_t)k * ld];
    // wall ghost rows are state the caller's BC calls wrote; periodic ghost
    // rows are not stored (the neighbour index points at the partner)
    if (ctx->kind[f] == FK_CELL)
      for (int w = 0; w < ctx->NGW; ++w)
        for (int k = 0; k < NZ; ++k)
          buf[(size_t)(N0 + w) * NS + k] = host[(size_t)ctx->gw_ghost[w] + (size_t)k * ld];
  }
  HIPCHECK(hipMemcpyAsync(ctx->slot[f], buf.data(), buf.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  HIPCHECK(hipStreamSynchronize(ctx->stream));
  return 0;
}

extern "C" int nsx_get_field(nsx_ctx *ctx, int f, double *host, int fill_ghost_rows) {
  NEED_COMMIT(ctx);
  if (check_field(ctx, f)) return -1;
  const int N0 = ctx->N0, NZ = ctx->NZ, NS = ctx->NS, NV = ctx->NV, NC = ctx->NC;
  std::vector<double> buf(field_elems(ctx, ctx->kind[f]));
  HIPCHECK(hipMemcpyAsync(buf.data(), ctx->slot[f], buf.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHECK(hipStreamSynchronize(ctx->stream));
  if (ctx->kind[f] == FK_VERT) {
    for (int l = 0; l < NZ - 1; ++l)
      for (int v = 0; v < NV; ++v) host[(size_t)v + (size_t)l * NV] = buf[(size_t)v * (NZ - 1) + l];
    return 0;
  }
  const size_t ld = ctx->kind[f] == FK_FACE ? (size_t)N0 : (size_t)NC;
  for (int k = 0; k < NZ; ++k)
    for (int i = 0; i < N0; ++i) host[(size_t)i + (size_t)k * ld] = buf[(size_t)i * NS + k];
  if (fill_ghost_rows && ctx->kind[f] == FK_CELL) {
    for (int gh = N0; gh < ctx->NEX; ++gh) {
      const int partner = ctx->No_cp[(size_t)gh + 2 * (size_t)NC] - 1;
      if (partner < 0 || partner >= N0) continue;
      for (int k = 0; k < NZ; ++k) host[(size_t)gh + (size_t)k * NC] = buf[(size_t)partner * NS + k];
    }
    for (int w = 0; w < ctx->NGW; ++w)
      for (int k = 0; k < NZ; ++k)
        host[(size_t)ctx->gw_ghost[w] + (size_t)k * NC] = buf[(size_t)(N0 + w) * NS + k];
  }
  return 0;
}

extern "C" int nsx_copy_field(nsx_ctx *ctx, int dst, int src) {
  NEED_COMMIT(ctx);
  if (check_field(ctx, dst) || check_field(ctx, src)) return -1;
  if (ctx->kind[dst] != ctx->kind[src]) return nsx_fail(ctx, -1, "field kinds differ");
  if (dst == src) return 0;
  LAUNCH(launch_copy(ctx->stream, field_elems(ctx, ctx->kind[src]), ctx->slot[src], ctx->slot[dst]));
  HIPCHECK(hipStreamSynchronize(ctx->stream));
  return 0;
}

// -------------------------------------------------------------- timing ----
static hipEvent_t *events(nsx_ctx *ctx, int n) {
  while (ctx->evpool.size() < ctx->evused + n) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    ctx->evpool.push_back(e);
  }
  hipEvent_t *p = &ctx->evpool[ctx->evused];
  ctx->evused += n;
  return p;
}

static void collect_timing(nsx_ctx *ctx) {
  if (!ctx->timing) return;
  (void)hipStreamSynchronize(ctx->stream);
  Timer *t[21] = {&ctx->t_k[0],   &ctx->t_k[1],   &ctx->t_k[2],   nullptr,        nullptr,      &ctx->t_stage,
                  nullptr,         nullptr,         nullptr,         nullptr,         nullptr,      nullptr,
                  &ctx->t_cg[0],   &ctx->t_cg[1],   &ctx->t_cgs,     &ctx->t_cgs64,   &ctx->t_mod[0],
                  &ctx->t_mod[1],  &ctx->t_mod[2],  &ctx->t_mod[3],  &ctx->t_mod[4]};
  for (auto &pr : ctx->pending) {
    float ms = 0.f;
    if (t[pr.second] && hipEventElapsedTime(&ms, ctx->evpool[pr.first], ctx->evpool[pr.first + 1]) == hipSuccess) {
      t[pr.second]->ms += ms;
      t[pr.second]->n += 1;
    }
  }
  ctx->pending.clear();
  ctx->evused = 0;
}

// ---------------------------------------------------------- the solver ----
// bicgstab.F90:8-168: Glob_Norm(rhs,1) [sum done by the caller into
// st->avec_rhs], BICGSTAB_SOLVE, Glob_Norm(phi,0) and the ghost layers of p.
static int solve_fp64(nsx_ctx *ctx, const double *rhs, const double *phi_in, double *x, bool wrapper = true) {
  const Geo &g = ctx->g;
  hipStream_t s = ctx->stream;
  LAUNCH(launch_bicg_init(g, s, rhs, phi_in, x, ctx->sr, ctx->srhat, ctx->part, ctx->st_dev, ctx->eps,
                          ctx->MaxIters));
  int issued = 0;
  const size_t pend0 = ctx->pending.size();
  // queue about as many iterations as the previous solve needed, then top up
  // in small chunks: few host round trips and few no-op launches
  const int hint = ctx->iter_hint;
  int chunk = hint > 0 ? (hint - hint / 16 > 4 ? hint - hint / 16 : 4) : 32;
  for (;;) {
    for (int c = 0; c < chunk; ++c) {
      hipEvent_t *ev = nullptr;
      if (ctx->timing) {
        ev = events(ctx, 6);
        if (!ev) return nsx_fail(ctx, -2, "hipEventCreate failed");
      }
      LAUNCH(launch_bicg_iter(g, s, x, ctx->sr, ctx->srhat, ctx->sq, ctx->svt, ctx->stv,
                              ctx->part, ctx->st_dev, ev));
      if (ev) {
        const size_t base = (size_t)(ev - &ctx->evpool[0]);
        for (int q = 0; q < 3; ++q) ctx->pending.push_back({base + 2 * q, q});
      }
      ++issued;
    }
    HIPCHECK(hipMemcpyAsync(ctx->st_host, ctx->st_dev, sizeof(SolverState), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    if (ctx->st_host->done) break;
    if (issued > ctx->MaxIters + 4) return nsx_fail(ctx, -4, "solver did not stop after %d iterations", issued);
    chunk = hint > 0 ? (hint / 32 > 4 ? hint / 32 : 4) : 16;
  }
  // bicgstab's Glob_Norm(phi,0) and ghost layers; BICGSTAB_SOLVE alone leaves x as solved
  if (wrapper) LAUNCH(launch_gnorm_apply(g, s, x, ctx->st_dev));
  const SolverState &h = *ctx->st_host;
  ctx->info.iters = h.a_iter;
  ctx->info.exit_code = h.exit_code;
  ctx->info.e1norm = h.e1;
  ctx->info.e2norm = h.e2;
  ctx->info.n_solves += 1;
  ctx->info.iters_total += h.a_iter;
  ctx->info.outer_passes = 0;
  ctx->info.iters_fp32 = 0;
  ctx->iter_hint = h.a_iter > 0 ? h.a_iter : ctx->iter_hint;
  // time only the iterations that ran, not the queued launches that found
  // the solver finished
  if (ctx->timing && ctx->pending.size() > pend0 + 3 * (size_t)h.a_iter)
    ctx->pending.resize(pend0 + 3 * (size_t)h.a_iter);
  collect_timing(ctx);
  return 0;
}

// The FP64 solve (the reference's BiCGSTAB) from the x an FP32 phase has
// reached; the solve record counts both phases.
static int finish_fp64(nsx_ctx *ctx, const double *rhs, double *x, int its32, int passes) {
  const int before = ctx->info.iters_total;
  int rc = solve_fp64(ctx, rhs, x, x);
  if (rc) return rc;
  ctx->info.iters_fp32 = its32;
  ctx->info.iters += its32;
  ctx->info.iters_total = before + ctx->info.iters;
  ctx->info.outer_passes = passes;
  return 0;
}

// Mixed precision, NSX_REFINE_CG: FP32 conjugate gradients on the correction
// with reliable updates against the FP64 solution (k_cg_start, k_cg_w/x); the solve
// stops when the true FP64 residual meets the reference's rule.  A run that
// reaches the iteration cap or breaks down leaves x at its last update, and
// the reference's FP64 BiCGSTAB finishes from there.
static int solve_cg(nsx_ctx *ctx, const double *rhs, const double *phi_in, double *x) {
  const Geo &g = ctx->g;
  hipStream_t s = ctx->stream;
  // CG does one product with A per iteration, BiCGSTAB two
  const int cap = 2 * ctx->MaxIters + 100;
  LAUNCH(launch_cg_init(g, s, rhs, phi_in, x, ctx->fr, ctx->frhat, ctx->fe, ctx->part, ctx->st_dev, ctx->eps,
                        ctx->delta_cg, cap));
  float *const v[7] = {ctx->fe, ctx->fr, ctx->fp, ctx->fv, ctx->fe2, ctx->fr2, ctx->fp2};
  int issued = 0;
  const size_t pend0 = ctx->pending.size();
  const int hint = ctx->iter_hint_cg;
  int chunk = hint > 0 ? (hint - hint / 16 > 4 ? hint - hint / 16 : 4) : 32;
  for (;;) {
    for (int c = 0; c < chunk; ++c) {
      hipEvent_t *ev = nullptr;
      if (ctx->timing) {
        ev = events(ctx, 4);
        if (!ev) return nsx_fail(ctx, -2, "hipEventCreate failed");
      }
      LAUNCH(launch_cg_iter(g, s, rhs, x, v, ctx->part, ctx->st_dev, ev));
      if (ev) {
        const size_t base = (size_t)(ev - &ctx->evpool[0]);
        for (int q = 0; q < 2; ++q) ctx->pending.push_back({base + 2 * q, 12 + q});
      }
      ++issued;
    }
    HIPCHECK(hipMemcpyAsync(ctx->st_host, ctx->st_dev, sizeof(SolverState), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    if (ctx->st_host->done) break;
    if (issued > cap + 4) return nsx_fail(ctx, -4, "solver did not stop after %d iterations", issued);
    chunk = hint > 0 ? (hint / 32 > 4 ? hint / 32 : 4) : 16;
  }
  const SolverState &h = *ctx->st_host;
  // launches queued after the stop returned at once: keep the timing of
  // the iterations that ran (W of the stopping iteration included)
  const size_t keep = pend0 + 2 * (size_t)h.a_iter + 1;
  if (ctx->timing && ctx->pending.size() > keep) ctx->pending.resize(keep);
  if (h.exit_code == NSX_SOLVE_MAX_ITERS || h.exit_code == NSX_SOLVE_BREAKDOWN) {
    collect_timing(ctx);
    const int its32 = h.a_iter, ups = h.mp_pass;
    if (its32 > 0) ctx->iter_hint_cg = its32;
    return finish_fp64(ctx, rhs, x, its32, ups);
  }
  LAUNCH(launch_gnorm_apply(g, s, x, ctx->st_dev));
  ctx->info.iters = h.a_iter;
  ctx->info.exit_code = h.exit_code;
  ctx->info.e1norm = h.e1;
  ctx->info.e2norm = h.e2;
  ctx->info.n_solves += 1;
  ctx->info.iters_total += h.a_iter;
  ctx->info.outer_passes = h.mp_pass;
  ctx->info.iters_fp32 = h.a_iter;
  if (h.a_iter > 0) ctx->iter_hint_cg = h.a_iter;
  collect_timing(ctx);
  return 0;
}

// Mixed precision, NSX_REFINE_CGSR (_BLOCK, _LINE): the same FP32 CG with reliable updates
// in the single-reduction form, one kernel per iteration (k_cgs); a launch
// is an iteration, a true residual or the product after one, as the device
// decides.  Fallback as in solve_cg.
static int solve_cgs(nsx_ctx *ctx, const double *rhs, const double *phi_in, double *x, bool fp64 = false) {
  const Geo &g = ctx->g;
  hipStream_t s = ctx->stream;
  // the preconditioner: point Jacobi, 4-layer blocks or lines (a window of
  // the line preconditioner holds whole columns; nsx_set_precision checks
  // this too, but may run before the mesh is committed)
  const int pc = fp64 ? 1 : ctx->refine == NSX_REFINE_CGSR_LINE ? 2 : ctx->refine == NSX_REFINE_CGSR_BLOCK ? 1 : 0;
  if (pc == 2 && g.NQ > NSX_BLOCK)
    return nsx_fail(ctx, -1, "line preconditioner: %d layers per column exceed %d", g.NS, 4 * NSX_BLOCK);
  const int cap = 2 * ctx->MaxIters + 100;
  // FP64: e = stv, p = sq, rt = (sr, scg0), om = (svt, scg1), sg = (scg2, scg3)
  const CgsVecsT<double> v64{ctx->stv, ctx->sq, {ctx->sr, ctx->scg[0]}, {ctx->svt, ctx->scg[1]},
                             {ctx->scg[2], ctx->scg[3]}};
  if (fp64)
    LAUNCH(launch_cg_init64(g, s, rhs, phi_in, x, v64.rt[0], ctx->srhat, v64.e, ctx->part, ctx->st_dev, ctx->eps,
                            ctx->delta_cg, cap));
  else
    LAUNCH(launch_cg_init(g, s, rhs, phi_in, x, ctx->fr, ctx->frhat, ctx->fe, ctx->part, ctx->st_dev, ctx->eps,
                          ctx->delta_cg, cap));
  const CgsVecs v{ctx->fe, ctx->fp, {ctx->fr, ctx->fr2}, {ctx->fv, ctx->fp2}, {ctx->ft, ctx->fe2}};
  int issued = 0;
  const size_t pend0 = ctx->pending.size();
  int &hint_ref = fp64 ? ctx->iter_hint_cgs64 : ctx->iter_hint_cgs;
  const int hint = hint_ref;
  int chunk = hint > 0 ? (hint - hint / 16 > 4 ? hint - hint / 16 : 4) : 32;
  for (;;) {
    for (int c = 0; c < chunk; ++c) {
      hipEvent_t *ev = nullptr;
      if (ctx->timing) {
        ev = events(ctx, 2);
        if (!ev) return nsx_fail(ctx, -2, "hipEventCreate failed");
      }
      if (fp64)
        LAUNCH(launch_cgs64(g, s, issued, rhs, x, v64, ctx->part, ctx->st_dev, ev));
      else
        LAUNCH(launch_cgs(g, s, issued, rhs, x, v, ctx->part, ctx->st_dev, ev, pc));
      if (ev) ctx->pending.push_back({(size_t)(ev - &ctx->evpool[0]), fp64 ? 15 : 14});
      ++issued;
    }
    HIPCHECK(hipMemcpyAsync(ctx->st_host, ctx->st_dev, sizeof(SolverState), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    if (ctx->st_host->done) break;
    // an iteration is one launch, a true residual two more
    if (issued > 3 * cap + 8) return nsx_fail(ctx, -4, "solver did not stop after %d launches", issued);
    chunk = hint > 0 ? (hint / 32 > 4 ? hint / 32 : 4) : 16;
  }
  const SolverState &h = *ctx->st_host;
  // launches queued after the stop returned at once: keep those that ran
  const size_t keep = pend0 + (size_t)h.cgs_runs;
  if (ctx->timing && ctx->pending.size() > keep) ctx->pending.resize(keep);
  if (h.exit_code == NSX_SOLVE_MAX_ITERS || h.exit_code == NSX_SOLVE_BREAKDOWN) {
    collect_timing(ctx);
    const int its32 = h.a_iter, ups = h.mp_pass;
    if (its32 > 0) hint_ref = its32;
    return finish_fp64(ctx, rhs, x, its32, ups);
  }
  LAUNCH(launch_gnorm_apply(g, s, x, ctx->st_dev));
  ctx->info.iters = h.a_iter;
  ctx->info.exit_code = h.exit_code;
  ctx->info.e1norm = h.e1;
  ctx->info.e2norm = h.e2;
  ctx->info.n_solves += 1;
  ctx->info.iters_total += h.a_iter;
  ctx->info.outer_passes = h.mp_pass;
  ctx->info.iters_fp32 = fp64 ? 0 : h.a_iter;
  if (h.a_iter > 0) hint_ref = h.a_iter;
  collect_timing(ctx);
  return 0;
}

// the vertical-mode preconditioner's tables (nsx_modal.hip), built on the
// host and uploaded at the first solve that uses them
static int modal_setup(nsx_ctx *ctx) {
  if (ctx->modal_ok) return 0;
  const Geo &g = ctx->g;
  const int N0 = ctx->N0, NZ = ctx->NZ, nzi = NZ - 2;
  // vertical couplings per unit area: between interior layers k and k+1,
  // 1/dsig(k) (the quotient area/dsig(k) of the rows); a coupling to a ghost
  // layer that is not Neumann stays on the diagonal
  std::vector<double> tlay((size_t)nzi + 1, 0.0), svl((size_t)nzi);
  for (int k = 1; k <= NZ - 3; ++k) tlay[k] = 1.0 / ctx->dsig[k];
  tlay[0] = g.bcbot == 1 ? 0.0 : 1.0 / ctx->dsig[0];
  tlay[nzi] = g.bctop == 1 ? 0.0 : 1.0 / ctx->dsig[NZ - 2];
  for (int k = 1; k <= NZ - 2; ++k) svl[k - 1] = ctx->dsigv[k - 1];
  ModalHost H;
  std::string err;
  const int m0 = std::min(NSX_MODAL_M0, nzi);
  if (modal_build_host(N0, nzi, ctx->h_nbr.data(), ctx->h_ppec.data(), ctx->areaCell.data(), tlay.data(),
                       svl.data(), m0, H, err))
    return nsx_fail(ctx, -1, "%s", err.c_str());
  ModalDev &M = ctx->modal;
  M = ModalDev{};
  const int nzp = (nzi + 3) / 4 * 4, MP = (nzi + 15) / 16 * 16, MT = MP / 16;
  M.m0 = H.m0;
  M.nzi = nzi;
  M.nzp = nzp;
  M.mp = MP;
  M.mt = MT;
  M.nlev = (int)H.levels.size();
  M.nc = H.nc;
  {
    // transform blocks: two per CU, four waves of 16 columns each (the
    // vector-unit transforms take the same grid)
    hipDeviceProp_t prop;
    HIPCHECK(hipGetDeviceProperties(&prop, ctx->dev));
    const int waves = (N0 + 15) / 16, want = (waves + 3) / 4;
    M.agrid = std::max(1, std::min(want, 2 * prop.multiProcessorCount));
    M.lds_optin = (int)std::max(prop.sharedMemPerBlockOptin, (size_t)65536);
  }
  std::vector<float> ph((size_t)nzp * nzp, 0.f), pt((size_t)nzp * nzp, 0.f), mu((size_t)MP, 1.f);
  for (int k = 0; k < nzi; ++k)
    for (int m = 0; m < nzi; ++m) {
      ph[(size_t)k * nzp + m] = H.phi[(size_t)k * nzi + m];
      pt[(size_t)m * nzp + k] = H.phi[(size_t)k * nzi + m];
    }
  for (int m = 0; m < nzi; ++m) mu[m] = H.mu[m];
  // matrix-core operands: forward, lane (j, q4) of (layer group t, mode tile
  // mt) holds phi(16 t + 4 q4 + q, 16 mt + j); backward, of (mode group t,
  // layer tile nt), phi(16 nt + j, 16 t + 4 q4 + q)
  auto phik = [&](int k, int m) -> float { return (k < nzi && m < nzi) ? H.phi[(size_t)k * nzi + m] : 0.f; };
  std::vector<float> tf((size_t)MT * MT * 256, 0.f), tb((size_t)MT * MT * 256, 0.f);
  for (int t = 0; t < MT; ++t)
    for (int mt = 0; mt < MT; ++mt)
      for (int lane = 0; lane < 64; ++lane)
        for (int q = 0; q < 4; ++q) {
          const int j = lane & 15, q4 = lane >> 4;
          tf[(((size_t)t * MT + mt) * 64 + lane) * 4 + q] = phik(16 * t + 4 * q4 + q, 16 * mt + j);
          tb[(((size_t)t * MT + mt) * 64 + lane) * 4 + q] = phik(16 * mt + j, 16 * t + 4 * q4 + q);
        }
  int rc;
  float *d_ph, *d_pt, *d_mu, *d_ci, *d_xh, *d_tf, *d_tb;
  if ((rc = dev_upload(ctx, &d_ph, ph)) || (rc = dev_upload(ctx, &d_pt, pt)) || (rc = dev_upload(ctx, &d_mu, mu)) ||
      (rc = dev_upload(ctx, &d_tf, tf)) || (rc = dev_upload(ctx, &d_tb, tb)) ||
      (rc = dev_alloc(ctx, &d_xh, (size_t)N0 * MP)))
    return rc;
  {
    // the coarsest inverses laid out [J][K][m] for coalesced reads
    const int nc = H.nc, m0v = H.m0;
    std::vector<float> ct((size_t)nc * nc * m0v, 0.f);
    for (int m = 0; m < m0v; ++m)
      for (int J = 0; J < nc; ++J)
        for (int K = 0; K < nc; ++K)
          ct[((size_t)J * nc + K) * m0v + m] = H.cinv[((size_t)m * nc + J) * nc + K];
    if ((rc = dev_upload(ctx, &d_ci, ct))) return rc;
  }
  HIPCHECK(hipMemset(d_xh, 0, (size_t)N0 * MP * sizeof(float)));
  M.phi = d_ph;
  M.phiT = d_pt;
  M.tfwd = reinterpret_cast<const float4 *>(d_tf);
  M.tbwd = reinterpret_cast<const float4 *>(d_tb);
  M.mu = d_mu;
  M.cinv = d_ci;
  M.xh = d_xh;
  for (int l = 0; l < M.nlev; ++l) {
    const ModalHost::Level &h = H.levels[l];
    ModalLevelDev &d = M.lev[l];
    d.n = h.n;
    auto up_i = [&](const std::vector<int> &v, const int *&dst) -> int {
      if (v.empty()) return 0;
      int *q;
      const int r = dev_upload(ctx, &q, v);
      dst = q;
      return r;
    };
    auto up_f = [&](const std::vector<float> &v, const float *&dst) -> int {
      if (v.empty()) return 0;
      float *q;
      const int r = dev_upload(ctx, &q, v);
      dst = q;
      return r;
    };
    if ((rc = up_i(h.ptr, d.ptr)) || (rc = up_i(h.col, d.col)) || (rc = up_f(h.w, d.w)) ||
        (rc = up_f(h.diag, d.diag)) || (rc = up_f(h.area, d.area)) || (rc = up_i(h.agg, d.agg)) ||
        (rc = up_i(h.eagg, d.eagg)) || (rc = up_i(h.dptr, d.dptr)) || (rc = up_i(h.dcol, d.dcol)) ||
        (rc = up_f(h.db, d.db)) || (rc = up_f(h.dz, d.dz)))
      return rc;
    const size_t nm = (size_t)h.n * (size_t)M.m0;
    if ((rc = dev_alloc(ctx, &M.b[l], nm)) || (rc = dev_alloc(ctx, &M.z[l], nm)) ||
        (rc = dev_alloc(ctx, &M.x[l], nm)))
      return rc;
    HIPCHECK(hipMemset(M.x[l], 0, nm * sizeof(float)));
  }
  ctx->modal_ok = true;
  return 0;
}

// Mixed precision, NSX_REFINE_MODAL: FP32 CG preconditioned by the vertical
// modes (launch_modal_apply) in cycles.  e gathers a cycle's correction in
// FP32; a cycle starts from the true FP64 residual with beta = 0 and ends
// when rho has fallen by NSX_MODAL_DROP (FP32 recursion holds the residual to
// about 1e-5 of the cycle's start; keeping p across the update, the next beta
// would divide the true residual's rho by the drifted recursive one) or after
// the iterations launched for it.  Then the true FP64 residual of x + e
// replaces r, e joins x and the reference's stopping rule is tested on it.
// The cap and a breakdown leave x at its last update and the FP64 BiCGSTAB
// finishes, as for the other CG forms.
static int solve_modal(nsx_ctx *ctx, const double *rhs, const double *phi_in, double *x) {
  int rc = modal_setup(ctx);
  if (rc) return rc;
  const Geo &g = ctx->g;
  hipStream_t s = ctx->stream;
  const int cap = 2 * ctx->MaxIters + 100;
  // r0 = b - A x0 (FP32 into frhat), e = 0, x = x0, the norms of b and the
  // reference's no-solve test
  LAUNCH(launch_cg_init(g, s, rhs, phi_in, x, ctx->fr, ctx->frhat, ctx->fe, ctx->part, ctx->st_dev, ctx->eps,
                        ctx->delta_cg, cap));
  float *r = ctx->frhat, *u = ctx->fr, *w = ctx->fv, *e = ctx->fe;
  float *p[2] = {ctx->fp, ctx->fp2};
  HIPCHECK(hipMemcpyAsync(ctx->st_host, ctx->st_dev, sizeof(SolverState), hipMemcpyDeviceToHost, s));
  HIPCHECK(hipStreamSynchronize(s));
  int launched = 0, cycle = NSX_MODAL_FIRST, before = 0;
  if (!ctx->st_host->done) {
    for (;;) {
      const size_t pend0 = ctx->pending.size();
      for (int c = 0; c < cycle; ++c) {
        // with timing: start/stop pairs of the forward transform, the
        // V-cycle, the backward transform, the product and the update
        hipEvent_t *ev = nullptr;
        if (ctx->timing) {
          ev = events(ctx, 10);
          if (!ev) return nsx_fail(ctx, -2, "hipEventCreate failed");
        }
        LAUNCH(launch_modal_apply(g, s, ctx->modal, r, u, ctx->part, ctx->st_dev, ev));
        LAUNCH(launch_modal_spmv(g, s, ctx->modal, u, p[launched & 1], p[(launched + 1) & 1], w, ctx->part,
                                 ctx->st_dev, c == 0, ev ? ev + 6 : nullptr));
        LAUNCH(launch_modal_update(g, s, p[(launched + 1) & 1], w, e, r, ctx->part, ctx->st_dev,
                                   ev ? ev + 8 : nullptr));
        if (ev) {
          const size_t base = (size_t)(ev - &ctx->evpool[0]);
          for (int q = 0; q < 5; ++q) ctx->pending.push_back({base + 2 * q, 16 + q});
        }
        ++launched;
      }
      LAUNCH(launch_modal_true(g, s, rhs, x, e, r, ctx->part, ctx->st_dev, cap));
      HIPCHECK(hipMemcpyAsync(ctx->st_host, ctx->st_dev, sizeof(SolverState), hipMemcpyDeviceToHost, s));
      HIPCHECK(hipStreamSynchronize(s));
      const int used = ctx->st_host->m_iters - before;
      before = ctx->st_host->m_iters;
      // timing: only the iterations that ran (after a drop the launches return at once)
      if (ctx->timing && ctx->pending.size() > pend0 + 5 * (size_t)used) ctx->pending.resize(pend0 + 5 * (size_t)used);
      if (ctx->st_host->done) break;
      // the next cycle: as long as this one ran, a few more
      cycle = std::min(NSX_MODAL_CYCLE, std::max(8, used + 4));
    }
  }
  collect_timing(ctx);
  const SolverState &h = *ctx->st_host;
  const int it = h.m_iters;
  if (h.exit_code == NSX_SOLVE_MAX_ITERS || h.exit_code == NSX_SOLVE_BREAKDOWN)
    return finish_fp64(ctx, rhs, x, it, h.mp_pass);
  LAUNCH(launch_gnorm_apply(g, s, x, ctx->st_dev));
  ctx->info.iters = it;
  ctx->info.exit_code = h.exit_code;
  ctx->info.e1norm = h.e1;
  ctx->info.e2norm = h.e2;
  ctx->info.n_solves += 1;
  ctx->info.iters_total += it;
  ctx->info.outer_passes = h.mp_pass;
  ctx->info.iters_fp32 = it;
  return 0;
}

static int solve(nsx_ctx *ctx, const double *rhs, const double *phi_in, double *x) {
  if (ctx->precision == NSX_PREC_FP64_CG) return solve_cgs(ctx, rhs, phi_in, x, true);
  if (ctx->precision != NSX_PREC_MIXED && ctx->precision != NSX_PREC_MIXED_FLUX32)
    return solve_fp64(ctx, rhs, phi_in, x);
  if (ctx->refine == NSX_REFINE_CG) return solve_cg(ctx, rhs, phi_in, x);
  if (ctx->refine == NSX_REFINE_MODAL) return solve_modal(ctx, rhs, phi_in, x);
  return solve_cgs(ctx, rhs, phi_in, x);
}

// BCpcenter (kind 0) or BCvelcenter3D (kinds 1..3 = u, v, w) on a stored
// field: the sigma ghost layers, then the wall ghost columns
static int bc_center(nsx_ctx *ctx, double *f, int kind) {
  LAUNCH(launch_ghost_layers(ctx->g, ctx->stream, f, kind));
  LAUNCH(launch_wall_ghosts(ctx->g, ctx->stream, f, kind));
  return 0;
}

#define BC_CENTER(f, kind)                 \
  do {                                     \
    const int rc_ = bc_center(ctx, f, kind); \
    if (rc_) return rc_;                   \
  } while (0)

// the wall ghost columns alone, of a field whose sigma ghost layers the
// kernel that wrote it filled
#define BC_WALLS(f, kind) LAUNCH(launch_wall_ghosts(ctx->g, ctx->stream, f, kind))

// vertex values of up to three fields; kinds 0 p, 1 u, 2 v, 3 w
static int vertex_of(nsx_ctx *ctx, int nf, const double *const f[3], double *const fv[3], const int kinds[3],
                     int tagdir) {
  const Geo &g = ctx->g;
  const signed char *act[3];
  int zb[3], zt[3], lid[3] = {0, 0, 0};
  double infl[3] = {0.0, 0.0, 0.0};
  for (int q = 0; q < nf; ++q) {
    const int kq = kinds[q];
    act[q] = kq == 0 ? g.vact_p : (kq == 1 ? g.vact_u : (kq == 2 ? g.vact_v : g.vact_w));
    // BCvelvertex3D zeroes the wall levels for tagdir 3 (BC.F90:737-787); BCpvertex does not
    zb[q] = kq == 0 ? 0 : (g.bcbot == 1 && (g.BotBC == 0 || (g.BotBC == 1 && kq == 3)));
    zt[q] = kq == 0 ? 0
                    : (g.bctop == 1 && (g.TopBC == 0 || (g.TopBC == 1 && kq == 3) || (g.TopBC == 2 && kq != 1)));
    // the lid sets u = 1 on the top level of the interior vertices (BC.F90:771-777)
    lid[q] = kq == 1 && g.bctop == 1 && g.TopBC == 2;
    // inflow() of Functions.F90:109-121 at an inflow vertex: 1 for v, 0 for u and w
    infl[q] = kq == 2 ? 1.0 : 0.0;
  }
  LAUNCH(launch_vertex(g, ctx->stream, nf, f, fv, act, zb, zt, tagdir, lid, infl));
  return 0;
}

// ---------------------------------------------------------------- stage ---
// hydro.F90:8-501 with inputs from pool set `in` and outputs into set `out`
static int stage(nsx_ctx *ctx, int in, int out, int flag_ab) {
  const Geo &g = ctx->g;
  hipStream_t s = ctx->stream;
  const double c1 = flag_ab == 1 ? 1.55 : 1.0, c2 = flag_ab == 1 ? -0.55 : 0.0;
  double *const *X = ctx->pool_cell[in];
  double *const *Y = ctx->pool_cell[out];
  hipEvent_t *ev = nullptr;
  if (ctx->timing) {
    ev = events(ctx, 4);
    if (!ev) return nsx_fail(ctx, -2, "hipEventCreate failed");
    (void)hipEventRecord(ev[0], s);
  }
  // the reference fills the sigma ghost layers of the stage input in place
  // (BCvelcenter3D inside BCglobalVC of sgles/AdvDiffNS, BCpcenter of p);
  // a resident input already holds them, an input put from the host may not
  for (int q = 0; q < 3; ++q) BC_CENTER(X[q], q + 1);
  BC_CENTER(X[3], 0);
  const double *uvw[3] = {X[0], X[1], X[2]};
  const int vk[3] = {1, 2, 3}, pk[3] = {0, 0, 0};
  // vertex values of the stage velocity (tagdir 2), shared by sgles and AdvDiffNS
  int rc = vertex_of(ctx, 3, uvw, ctx->uv2, vk, 2);
  if (rc) return rc;
  // the precision form with FP32 flux and eddy-viscosity arithmetic
  const bool f32 = ctx->precision == NSX_PREC_MIXED_FLUX32;
  if (ctx->Iturbu == 1) {
    if (f32)
      LAUNCH(launch_les32(g, s, X[0], X[1], X[2], ctx->uv2[0], ctx->uv2[1], ctx->uv2[2], ctx->slot[NSX_F_EDDY]));
    else
      LAUNCH(launch_les(g, s, X[0], X[1], X[2], ctx->uv2[0], ctx->uv2[1], ctx->uv2[2], ctx->slot[NSX_F_EDDY]));
    BC_WALLS(ctx->slot[NSX_F_EDDY], 0);  // sgles' BCpcenter(eddy)
  }
  const double *vtx[3] = {ctx->uv2[0], ctx->uv2[1], ctx->uv2[2]};
  const double *face[5] = {ctx->slot[NSX_F_U1FACE], ctx->slot[NSX_F_U2FACE], ctx->slot[NSX_F_U3FACE],
                           ctx->slot[NSX_F_UTFACE], ctx->slot[NSX_F_UBFACE]};
  double *hist[3] = {ctx->slot[NSX_F_RHSUF], ctx->slot[NSX_F_RHSVF], ctx->slot[NSX_F_RHSWF]};
  const double *force = ctx->body_force ? ctx->slot[NSX_F_FORCE] : nullptr;
  if (f32)
    LAUNCH(launch_momentum32(g, s, uvw, vtx, face, ctx->slot[NSX_F_EDDY], hist, ctx->ustar, c1, c2, force));
  else
    LAUNCH(launch_momentum(g, s, uvw, vtx, face, ctx->slot[NSX_F_EDDY], nullptr, hist, nullptr, ctx->ustar, c1, c2,
                           force));
  // BCglobalVC(ufstar, ...) of hydro.F90:342: the momentum kernels fill the
  // sigma ghost layers of u*, then the wall ghost columns
  for (int q = 0; q < 3; ++q) BC_WALLS(ctx->ustar[q], q + 1);
  // update_rhsp's BCglobalP(pf, pfv, 2)
  const double *pin[3] = {X[3], nullptr, nullptr};
  double *pfv[3] = {ctx->slot[NSX_F_PFV], nullptr, nullptr};
  if ((rc = vertex_of(ctx, 1, pin, pfv, pk, 2))) return rc;
  const double *us[3] = {ctx->ustar[0], ctx->ustar[1], ctx->ustar[2]};
  double *facew[5] = {ctx->slot[NSX_F_U1FACE], ctx->slot[NSX_F_U2FACE], ctx->slot[NSX_F_U3FACE],
                      ctx->slot[NSX_F_UTFACE], ctx->slot[NSX_F_UBFACE]};
  LAUNCH(launch_source(g, s, us, facew, nullptr, ctx->rhsp, ctx->slot[NSX_F_PFV], ctx->part));
  if (ev) {
    (void)hipEventRecord(ev[1], s);
    ctx->pending.push_back({(size_t)(ev - &ctx->evpool[0]), 5});
  }
  {
    // the initial guess: the stage's input pressure, or (nsx_step with
    // NSX_GUESS_EXTRAPOLATE) 2 p_k - p_{k-1} from the last two solutions of
    // this RK stage
    const int gs = ctx->guess_stage;
    const double *guess = X[3];
    if (gs >= 0 && ctx->nhist[gs] >= 2) {
      const int nw = ctx->hslot[gs] ^ 1;
      LAUNCH(launch_extrapolate(g, s, ctx->phist[gs][nw], ctx->phist[gs][nw ^ 1], ctx->pguess));
      guess = ctx->pguess;
    }
    if ((rc = solve(ctx, ctx->rhsp, guess, Y[3]))) return rc;
    if (gs >= 0) {
      HIPCHECK(hipMemcpyAsync(ctx->phist[gs][ctx->hslot[gs]], Y[3], cell_bytes(ctx), hipMemcpyDeviceToDevice, s));
      ctx->hslot[gs] ^= 1;
      ctx->nhist[gs] += 1;
    }
  }
  BC_WALLS(Y[3], 0);  // bicgstab's BCpcenter(phi)
  if (ctx->timing) {
    ev = events(ctx, 2);
    (void)hipEventRecord(ev[0], s);
  }
  double *unp[3] = {Y[0], Y[1], Y[2]};
  LAUNCH(launch_corrector(g, s, Y[3], us, unp));
  for (int q = 0; q < 3; ++q) BC_WALLS(Y[q], q + 1);  // BCglobalVC(ufnp, ...) of hydro.F90:452
  const double *un[3] = {Y[0], Y[1], Y[2]};
  double *uvf[3] = {ctx->slot[NSX_F_UFV], ctx->slot[NSX_F_VFV], ctx->slot[NSX_F_WFV]};
  if ((rc = vertex_of(ctx, 3, un, uvf, vk, 3))) return rc;
  const double *pout[3] = {Y[3], nullptr, nullptr};
  if ((rc = vertex_of(ctx, 1, pout, pfv, pk, 2))) return rc;
  LAUNCH(launch_newu(g, s, Y[3], ctx->slot[NSX_F_PFV], facew));
  if (ev) {
    (void)hipEventRecord(ev[1], s);
    ctx->pending.push_back({(size_t)(ev - &ctx->evpool[0]), 5});
    collect_timing(ctx);
  }
  return 0;
}

static void point_slots(nsx_ctx *ctx, int n, int stage_in, int newest) {
  for (int q = 0; q < 4; ++q) {
    ctx->slot[NSX_F_UFN + q] = ctx->pool_cell[n][q];
    ctx->slot[NSX_F_UF + q] = ctx->pool_cell[stage_in][q];
    ctx->slot[NSX_F_UFNP + q] = ctx->pool_cell[newest][q];
  }
}

extern "C" int nsx_init_state(nsx_ctx *ctx) {
  NEED_COMMIT(ctx);
  reset_guess(ctx);
  const Geo &g = ctx->g;
  hipStream_t s = ctx->stream;
  // the UFN..PFN slots hold the initial field; make them pool set 0
  int n = -1;
  for (int st = 0; st < 3; ++st)
    if (ctx->slot[NSX_F_UFN] == ctx->pool_cell[st][0]) n = st;
  if (n < 0) return nsx_fail(ctx, -1, "time-level slots are inconsistent");
  double *const *X = ctx->pool_cell[n];
  for (int q = 0; q < 3; ++q) BC_CENTER(X[q], q + 1);
  BC_CENTER(X[3], 0);
  const double *uvw[3] = {X[0], X[1], X[2]};
  const int vk[3] = {1, 2, 3}, pk[3] = {0, 0, 0};
  double *uvf[3] = {ctx->slot[NSX_F_UFV], ctx->slot[NSX_F_VFV], ctx->slot[NSX_F_WFV]};
  int rc = vertex_of(ctx, 3, uvw, uvf, vk, 3);
  if (rc) return rc;
  const double *pin[3] = {X[3], nullptr, nullptr};
  double *pfv[3] = {ctx->slot[NSX_F_PFV], nullptr, nullptr};
  if ((rc = vertex_of(ctx, 1, pin, pfv, pk, 3))) return rc;
  double *facew[5] = {ctx->slot[NSX_F_U1FACE], ctx->slot[NSX_F_U2FACE], ctx->slot[NSX_F_U3FACE],
                      ctx->slot[NSX_F_UTFACE], ctx->slot[NSX_F_UBFACE]};
  LAUNCH(launch_source(g, s, uvw, facew, nullptr, nullptr, nullptr, ctx->part));
  const size_t nc = field_elems(ctx, FK_CELL);
  for (int id : {NSX_F_RHSUF, NSX_F_RHSVF, NSX_F_RHSWF, NSX_F_EDDY})
    HIPCHECK(hipMemsetAsync(ctx->slot[id], 0, nc * sizeof(double), s));
  // LoopTimeInitial: (.)np <- (.)n
  const int a = (n + 1) % 3, b = (n + 2) % 3;
  for (int q = 0; q < 4; ++q)
    HIPCHECK(hipMemcpyAsync(ctx->pool_cell[b][q], X[q], nc * sizeof(double), hipMemcpyDeviceToDevice, s));
  point_slots(ctx, n, a, b);
  ctx->set_n = n;
  ctx->info = nsx_solve_info{};
  HIPCHECK(hipStreamSynchronize(s));
  return 0;
}

extern "C" int nsx_hydro(nsx_ctx *ctx, int flag_ab, nsx_solve_info *info) {
  NEED_COMMIT(ctx);
  // UF.. -> UFNP..: find their pool sets
  int in = -1, out = -1;
  for (int st = 0; st < 3; ++st) {
    if (ctx->slot[NSX_F_UF] == ctx->pool_cell[st][0]) in = st;
    if (ctx->slot[NSX_F_UFNP] == ctx->pool_cell[st][0]) out = st;
  }
  if (in < 0 || out < 0 || in == out) return nsx_fail(ctx, -1, "stage input and output must be distinct fields");
  int rc = stage(ctx, in, out, flag_ab);
  if (rc) return rc;
  HIPCHECK(hipStreamSynchronize(ctx->stream));
  if (info) *info = ctx->info;
  return ctx->info.exit_code == NSX_SOLVE_MAX_ITERS ? NSX_SOLVE_MAX_ITERS : 0;
}

extern "C" int nsx_step(nsx_ctx *ctx, int nsteps, int precision, nsx_solve_info *info) {
  NEED_COMMIT(ctx);
  if (precision < NSX_PREC_FP64 || precision > NSX_PREC_MIXED_FLUX32)
    return nsx_fail(ctx, -1, "unknown precision %d", precision);
  hipStream_t s = ctx->stream;
  int worst = 0;
  const int saved = ctx->precision;
  ctx->precision = precision;
  struct Restore {
    nsx_ctx *c;
    int p;
    ~Restore() {
      c->precision = p;
      c->guess_stage = -1;
    }
  } restore{ctx, saved};
  // extrapolated initial guesses in the precision forms only; steps without
  // them leave the stored solutions stale, so they are dropped
  const bool extrap = ctx->guess_mode == NSX_GUESS_EXTRAPOLATE && precision != NSX_PREC_FP64;
  if (!extrap) reset_guess(ctx);
  for (int n = 0; n < nsteps; ++n) {
    const int N = ctx->set_n, A = (N + 1) % 3, B = (N + 2) % 3;
    int rc;
    if (ctx->FinRK == 2) {
      ctx->guess_stage = extrap ? 0 : -1;
      if ((rc = stage(ctx, N, A, 0))) return rc;
      if (ctx->info.exit_code == NSX_SOLVE_MAX_ITERS) worst = NSX_SOLVE_MAX_ITERS;
      ctx->guess_stage = extrap ? 1 : -1;
      if ((rc = stage(ctx, A, B, 0))) return rc;
      if (ctx->info.exit_code == NSX_SOLVE_MAX_ITERS) worst = NSX_SOLVE_MAX_ITERS;
      const double *un[3] = {ctx->pool_cell[N][0], ctx->pool_cell[N][1], ctx->pool_cell[N][2]};
      double *unp[3] = {ctx->pool_cell[B][0], ctx->pool_cell[B][1], ctx->pool_cell[B][2]};
      LAUNCH(launch_rk2avg(ctx->g, s, un, unp));
      ctx->set_n = B;
      point_slots(ctx, B, A, B);
    } else {
      // Euler, then AB2 from the second step on (hydro.F90:469-478)
      const int flag_ab = ctx->info.n_solves > 0 ? 1 : 0;
      ctx->guess_stage = extrap ? 0 : -1;
      if ((rc = stage(ctx, N, A, flag_ab))) return rc;
      if (ctx->info.exit_code == NSX_SOLVE_MAX_ITERS) worst = NSX_SOLVE_MAX_ITERS;
      ctx->set_n = A;
      point_slots(ctx, A, N, A);
    }
  }
  HIPCHECK(hipStreamSynchronize(s));
  if (info) *info = ctx->info;
  return worst;
}

extern "C" int nsx_solve_info_get(nsx_ctx *ctx, nsx_solve_info *info) {
  if (!ctx || !info) return -1;
  *info = ctx->info;
  return 0;
}

// ------------------------------------------------------ single routines ---
static int kind_ok(nsx_ctx *ctx, int id, int kind) {
  if (check_field(ctx, id)) return -1;
  if (ctx->kind[id] != kind) return nsx_fail(ctx, -1, "field %d has the wrong kind for this argument", id);
  return 0;
}

extern "C" int nsx_bcglobalvc(nsx_ctx *ctx, int u, int v, int w, int uv, int vv, int wv, int tagdir) {
  NEED_COMMIT(ctx);
  if (kind_ok(ctx, u, FK_CELL) || kind_ok(ctx, v, FK_CELL) || kind_ok(ctx, w, FK_CELL) || kind_ok(ctx, uv, FK_VERT) ||
      kind_ok(ctx, vv, FK_VERT) || kind_ok(ctx, wv, FK_VERT))
    return -1;
  if (tagdir != 2 && tagdir != 3) return nsx_fail(ctx, -1, "tagdir must be 2 or 3");
  const int ids[3] = {u, v, w};
  for (int q = 0; q < 3; ++q) BC_CENTER(ctx->slot[ids[q]], q + 1);
  const double *f[3] = {ctx->slot[u], ctx->slot[v], ctx->slot[w]};
  double *fv[3] = {ctx->slot[uv], ctx->slot[vv], ctx->slot[wv]};
  const int vk[3] = {1, 2, 3};
  int rc = vertex_of(ctx, 3, f, fv, vk, tagdir);
  if (rc) return rc;
  HIPCHECK(hipStreamSynchronize(ctx->stream));
  return 0;
}

static int bcglobalp_(nsx_ctx *ctx, int p, int pv, int tagdir) {
  BC_CENTER(ctx->slot[p], 0);
  const double *f[3] = {ctx->slot[p], nullptr, nullptr};
  double *fv[3] = {ctx->slot[pv], nullptr, nullptr};
  const int pk[3] = {0, 0, 0};
  return vertex_of(ctx, 1, f, fv, pk, tagdir);
}

extern "C" int nsx_bcglobalp(nsx_ctx *ctx, int p, int pv, int tagdir) {
  NEED_COMMIT(ctx);
  if (kind_ok(ctx, p, FK_CELL) || kind_ok(ctx, pv, FK_VERT)) return -1;
  if (tagdir != 2 && tagdir != 3) return nsx_fail(ctx, -1, "tagdir must be 2 or 3");
  int rc = bcglobalp_(ctx, p, pv, tagdir);
  if (rc) return rc;
  HIPCHECK(hipStreamSynchronize(ctx->stream));
  return 0;
}

extern "C" int nsx_grad_lsm(nsx_ctx *ctx, int dfdx, int dfdy, int dfds, int f) {
  NEED_COMMIT(ctx);
  if (kind_ok(ctx, dfdx, FK_CELL) || kind_ok(ctx, dfdy, FK_CELL) || kind_ok(ctx, dfds, FK_CELL) ||
      kind_ok(ctx, f, FK_CELL))
    return -1;
  LAUNCH(launch_grad_lsm(ctx->g, ctx->stream, ctx->slot[f], ctx->slot[dfdx], ctx->slot[dfdy], ctx->slot[dfds]));
  // grandientLSM's BCpcenter on its three outputs (gradientLSM.F90:199-201)
  BC_WALLS(ctx->slot[dfdx], 0);
  BC_WALLS(ctx->slot[dfdy], 0);
  BC_WALLS(ctx->slot[dfds], 0);
  HIPCHECK(hipStreamSynchronize(ctx->stream));
  return 0;
}

extern "C" int nsx_calcul_face(nsx_ctx *ctx, int f1, int f2, int f3, int fT, int fB, int fun, int funv) {
  NEED_COMMIT(ctx);
  const int o[5] = {f1, f2, f3, fT, fB};
  for (int q = 0; q < 5; ++q)
    if (kind_ok(ctx, o[q], FK_FACE)) return -1;
  if (kind_ok(ctx, fun, FK_CELL) || kind_ok(ctx, funv, FK_VERT)) return -1;
  double *out[5] = {ctx->slot[f1], ctx->slot[f2], ctx->slot[f3], ctx->slot[fT], ctx->slot[fB]};
  LAUNCH(launch_calcul_face(ctx->g, ctx->stream, ctx->slot[fun], ctx->slot[funv], out));
  HIPCHECK(hipStreamSynchronize(ctx->stream));
  return 0;
}

extern "C" int nsx_sgles(nsx_ctx *ctx, int u, int v, int w) {
  NEED_COMMIT(ctx);
  if (kind_ok(ctx, u, FK_CELL) || kind_ok(ctx, v, FK_CELL) || kind_ok(ctx, w, FK_CELL)) return -1;
  const int ids[3] = {u, v, w};
  for (int q = 0; q < 3; ++q) BC_CENTER(ctx->slot[ids[q]], q + 1);
  const double *f[3] = {ctx->slot[u], ctx->slot[v], ctx->slot[w]};
  const int vk[3] = {1, 2, 3};
  int rc = vertex_of(ctx, 3, f, ctx->uv2, vk, 2);
  if (rc) return rc;
  LAUNCH(launch_les(ctx->g, ctx->stream, f[0], f[1], f[2], ctx->uv2[0], ctx->uv2[1], ctx->uv2[2],
                    ctx->slot[NSX_F_EDDY]));
  BC_WALLS(ctx->slot[NSX_F_EDDY], 0);
  HIPCHECK(hipStreamSynchronize(ctx->stream));
  return 0;
}

extern "C" int nsx_advdiff(nsx_ctx *ctx, int rhsu, int rhsv, int rhsw, int u, int v, int w, int eddy) {
  NEED_COMMIT(ctx);
  const int ids[7] = {rhsu, rhsv, rhsw, u, v, w, eddy};
  for (int q = 0; q < 7; ++q)
    if (kind_ok(ctx, ids[q], FK_CELL)) return -1;
  for (int q = 0; q < 3; ++q) BC_CENTER(ctx->slot[ids[3 + q]], q + 1);
  const double *f[3] = {ctx->slot[u], ctx->slot[v], ctx->slot[w]};
  const int vk[3] = {1, 2, 3};
  int rc = vertex_of(ctx, 3, f, ctx->uv2, vk, 2);
  if (rc) return rc;
  const double *vtx[3] = {ctx->uv2[0], ctx->uv2[1], ctx->uv2[2]};
  const double *face[5] = {ctx->slot[NSX_F_U1FACE], ctx->slot[NSX_F_U2FACE], ctx->slot[NSX_F_U3FACE],
                           ctx->slot[NSX_F_UTFACE], ctx->slot[NSX_F_UBFACE]};
  const double *rin[3] = {ctx->slot[rhsu], ctx->slot[rhsv], ctx->slot[rhsw]};
  double *rout[3] = {ctx->slot[rhsu], ctx->slot[rhsv], ctx->slot[rhsw]};
  LAUNCH(launch_momentum(ctx->g, ctx->stream, f, vtx, face, ctx->slot[eddy], rin, nullptr, rout, nullptr, 1.0, 0.0));
  HIPCHECK(hipStreamSynchronize(ctx->stream));
  return 0;
}

// AdvDiffNS with the reference's three diffusivity fields (hydro.F90:280-293
// passes Gamu = Gamv = Gamw = eddy + 1/Re): rhs? += Gam? x diffusion - advection
extern "C" int nsx_advdiff_gam(nsx_ctx *ctx, int rhsu, int rhsv, int rhsw, int u, int v, int w, int gamx, int gamy,
                               int gamz) {
  NEED_COMMIT(ctx);
  const int ids[9] = {rhsu, rhsv, rhsw, u, v, w, gamx, gamy, gamz};
  for (int q = 0; q < 9; ++q)
    if (kind_ok(ctx, ids[q], FK_CELL)) return -1;
  for (int q = 0; q < 3; ++q) BC_CENTER(ctx->slot[ids[3 + q]], q + 1);
  const double *f[3] = {ctx->slot[u], ctx->slot[v], ctx->slot[w]};
  const int vk[3] = {1, 2, 3};
  int rc = vertex_of(ctx, 3, f, ctx->uv2, vk, 2);
  if (rc) return rc;
  const double *vtx[3] = {ctx->uv2[0], ctx->uv2[1], ctx->uv2[2]};
  const double *face[5] = {ctx->slot[NSX_F_U1FACE], ctx->slot[NSX_F_U2FACE], ctx->slot[NSX_F_U3FACE],
                           ctx->slot[NSX_F_UTFACE], ctx->slot[NSX_F_UBFACE]};
  const double *rin[3] = {ctx->slot[rhsu], ctx->slot[rhsv], ctx->slot[rhsw]};
  double *rout[3] = {ctx->slot[rhsu], ctx->slot[rhsv], ctx->slot[rhsw]};
  const double *gam[3] = {ctx->slot[gamx], ctx->slot[gamy], ctx->slot[gamz]};
  LAUNCH(launch_momentum(ctx->g, ctx->stream, f, vtx, face, ctx->slot[gamx], rin, nullptr, rout, nullptr, 1.0, 0.0,
                         nullptr, gam));
  HIPCHECK(hipStreamSynchronize(ctx->stream));
  return 0;
}

// interpolateU (interpolationU.F90:5-190): the ghost layers of u, v, w and
// their face-normal velocities into the five face fields f1..fB
extern "C" int nsx_interpolate_u(nsx_ctx *ctx, int f1, int f2, int f3, int fT, int fB, int u, int v, int w) {
  NEED_COMMIT(ctx);
  const int o[5] = {f1, f2, f3, fT, fB};
  for (int q = 0; q < 5; ++q)
    if (kind_ok(ctx, o[q], FK_FACE)) return -1;
  const int ids[3] = {u, v, w};
  for (int q = 0; q < 3; ++q)
    if (kind_ok(ctx, ids[q], FK_CELL)) return -1;
  for (int q = 0; q < 3; ++q) BC_CENTER(ctx->slot[ids[q]], q + 1);
  const double *f[3] = {ctx->slot[u], ctx->slot[v], ctx->slot[w]};
  double *out[5] = {ctx->slot[f1], ctx->slot[f2], ctx->slot[f3], ctx->slot[fT], ctx->slot[fB]};
  LAUNCH(launch_source(ctx->g, ctx->stream, f, out, nullptr, nullptr, nullptr, ctx->part));
  HIPCHECK(hipStreamSynchronize(ctx->stream));
  return 0;
}

extern "C" int nsx_source3d(nsx_ctx *ctx, int src, int u, int v, int w) {
  NEED_COMMIT(ctx);
  const int ids[4] = {src, u, v, w};
  for (int q = 0; q < 4; ++q)
    if (kind_ok(ctx, ids[q], FK_CELL)) return -1;
  for (int q = 0; q < 3; ++q) BC_CENTER(ctx->slot[ids[1 + q]], q + 1);
  const double *f[3] = {ctx->slot[u], ctx->slot[v], ctx->slot[w]};
  double *facew[5] = {ctx->slot[NSX_F_U1FACE], ctx->slot[NSX_F_U2FACE], ctx->slot[NSX_F_U3FACE],
                      ctx->slot[NSX_F_UTFACE], ctx->slot[NSX_F_UBFACE]};
  HIPCHECK(hipMemsetAsync(ctx->slot[src], 0, (size_t)ctx->N0 * ctx->NS * sizeof(double), ctx->stream));
  LAUNCH(launch_source(ctx->g, ctx->stream, f, facew, ctx->slot[src], nullptr, nullptr, ctx->part));
  HIPCHECK(hipStreamSynchronize(ctx->stream));
  return 0;
}

extern "C" int nsx_update_rhsp(nsx_ctx *ctx, int rhs, int p, int pv) {
  NEED_COMMIT(ctx);
  if (kind_ok(ctx, rhs, FK_CELL) || kind_ok(ctx, p, FK_CELL) || kind_ok(ctx, pv, FK_VERT)) return -1;
  int rc = bcglobalp_(ctx, p, pv, 2);
  if (rc) return rc;
  LAUNCH(launch_update_rhsp(ctx->g, ctx->stream, ctx->slot[rhs], ctx->slot[pv]));
  HIPCHECK(hipStreamSynchronize(ctx->stream));
  return 0;
}

extern "C" int nsx_bicgstab(nsx_ctx *ctx, int phi, int phiv, int rhs, nsx_solve_info *info) {
  NEED_COMMIT(ctx);
  if (kind_ok(ctx, phi, FK_CELL) || kind_ok(ctx, phiv, FK_VERT) || kind_ok(ctx, rhs, FK_CELL)) return -1;
  LAUNCH(launch_gnorm_sum(ctx->g, ctx->stream, ctx->slot[rhs], ctx->part));
  int rc = solve(ctx, ctx->slot[rhs], ctx->slot[phi], ctx->slot[phi]);
  if (rc) return rc;
  if ((rc = bcglobalp_(ctx, phi, phiv, 3))) return rc;
  HIPCHECK(hipStreamSynchronize(ctx->stream));
  if (info) *info = ctx->info;
  return ctx->info.exit_code == NSX_SOLVE_CONVERGED ? 0 : ctx->info.exit_code;
}

// BICGSTAB_SOLVE (bicgstab.F90:303-591) alone: A phi = rhs with the PPE
// matrix of the set-up, from the phi given, with the reference's stopping
// rule; neither mean removal of bicgstab's wrapper (Glob_Norm) is applied
extern "C" int nsx_bicgstab_solve(nsx_ctx *ctx, int phi, int rhs, nsx_solve_info *info) {
  NEED_COMMIT(ctx);
  if (kind_ok(ctx, phi, FK_CELL) || kind_ok(ctx, rhs, FK_CELL)) return -1;
  if (phi == rhs) return nsx_fail(ctx, -1, "nsx_bicgstab_solve: phi and rhs must be different fields");
  // no Glob_Norm(rhs,1): the solver's rhs average is zero
  HIPCHECK(hipMemsetAsync(ctx->part.src, 0, (size_t)MAXGRID * sizeof(double), ctx->stream));
  const int rc = solve_fp64(ctx, ctx->slot[rhs], ctx->slot[phi], ctx->slot[phi], false);
  if (rc) return rc;
  HIPCHECK(hipStreamSynchronize(ctx->stream));
  if (info) *info = ctx->info;
  return ctx->info.exit_code == NSX_SOLVE_CONVERGED ? 0 : ctx->info.exit_code;
}

// The adjoint of nsx_bicgstab with respect to its right-hand side.  The
// forward map is phi = N0 A^+ N1 rhs with N1 = Glob_Norm(.,1) (subtract the
// volume-weighted mean) and N0 = Glob_Norm(.,0) (subtract the mean over the
// volume); their transposes are each other's form, so the adjoint is the
// same wrapper on A^T: lam = N0 (A^T)^+ N1 g, solved by the FP64 BiCGSTAB
// from lam = 0 with the lateral couplings of A^T (ppec_t).
extern "C" int nsx_bicgstab_adjoint(nsx_ctx *ctx, int lam, int g, nsx_solve_info *info) {
  NEED_COMMIT(ctx);
  if (kind_ok(ctx, lam, FK_CELL) || kind_ok(ctx, g, FK_CELL)) return -1;
  if (lam == g) return nsx_fail(ctx, -1, "nsx_bicgstab_adjoint: lam and g must be different fields");
  HIPCHECK(hipMemsetAsync(ctx->slot[lam], 0, field_elems(ctx, FK_CELL) * sizeof(double), ctx->stream));
  LAUNCH(launch_gnorm_sum(ctx->g, ctx->stream, ctx->slot[g], ctx->part));
  ctx->g.trans = 1;
  const int rc = solve_fp64(ctx, ctx->slot[g], ctx->slot[lam], ctx->slot[lam]);
  ctx->g.trans = 0;
  if (rc) return rc;
  HIPCHECK(hipStreamSynchronize(ctx->stream));
  if (info) *info = ctx->info;
  return ctx->info.exit_code == NSX_SOLVE_CONVERGED ? 0 : ctx->info.exit_code;
}

extern "C" int nsx_newu3d(nsx_ctx *ctx, int p, int pv) {
  NEED_COMMIT(ctx);
  if (kind_ok(ctx, p, FK_CELL) || kind_ok(ctx, pv, FK_VERT)) return -1;
  int rc = bcglobalp_(ctx, p, pv, 2);
  if (rc) return rc;
  double *facew[5] = {ctx->slot[NSX_F_U1FACE], ctx->slot[NSX_F_U2FACE], ctx->slot[NSX_F_U3FACE],
                      ctx->slot[NSX_F_UTFACE], ctx->slot[NSX_F_UBFACE]};
  LAUNCH(launch_newu(ctx->g, ctx->stream, ctx->slot[p], ctx->slot[pv], facew));
  HIPCHECK(hipStreamSynchronize(ctx->stream));
  return 0;
}

// -------------------------------------------------------------- glob_cfl --
extern "C" int nsx_glob_cfl(nsx_ctx *ctx, int u, int v, int w, nsx_cfl_info *out) {
  NEED_COMMIT(ctx);
  if (!out) return nsx_fail(ctx, -1, "nsx_glob_cfl: out is NULL");
  if (kind_ok(ctx, u, FK_CELL) || kind_ok(ctx, v, FK_CELL) || kind_ok(ctx, w, FK_CELL)) return -1;
  const double *face[5] = {ctx->slot[NSX_F_U1FACE], ctx->slot[NSX_F_U2FACE], ctx->slot[NSX_F_U3FACE],
                           ctx->slot[NSX_F_UTFACE], ctx->slot[NSX_F_UBFACE]};
  LAUNCH(launch_glob_cfl(ctx->g, ctx->stream, ctx->slot[u], ctx->slot[v], ctx->slot[w], face, ctx->cfl_part,
                         ctx->cfl_out));
  double r[13];
  HIPCHECK(hipMemcpyAsync(r, ctx->cfl_out, sizeof r, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHECK(hipStreamSynchronize(ctx->stream));
  out->umax = r[0];
  out->vmax = r[1];
  out->wmax = r[2];
  out->cfl = r[3];
  out->aveu = r[4];
  out->avev = r[5];
  out->avew = r[6];
  out->flux_total = r[7];
  out->flux_max = r[8];
  out->flux_bc = r[9];
  out->flux_bot = r[10];
  out->flux_top = r[11];
  out->flux_side = r[12];
  return 0;
}

// glob_cfl with its partial sums and outputs in the library's static
// per-device scratch instead of the context's buffers: no allocation of any
// kind, for callers that cannot allocate.  Calls that overlap on one GPU race
// on the shared scratch.
extern "C" int nsx_glob_cfl_noalloc(nsx_ctx *ctx, int u, int v, int w, nsx_cfl_info *out) {
  NEED_COMMIT(ctx);
  if (!out) return nsx_fail(ctx, -1, "nsx_glob_cfl_noalloc: out is NULL");
  if (kind_ok(ctx, u, FK_CELL) || kind_ok(ctx, v, FK_CELL) || kind_ok(ctx, w, FK_CELL)) return -1;
  const double *face[5] = {ctx->slot[NSX_F_U1FACE], ctx->slot[NSX_F_U2FACE], ctx->slot[NSX_F_U3FACE],
                           ctx->slot[NSX_F_UTFACE], ctx->slot[NSX_F_UBFACE]};
  double *res = nullptr;
  const int rc = launch_glob_cfl_scratch(ctx->g, ctx->stream, ctx->dev, ctx->slot[u], ctx->slot[v], ctx->slot[w],
                                         face, &res);
  if (rc == -1)
    return nsx_fail(ctx, -1, "nsx_glob_cfl_noalloc: a grid of %d blocks exceeds the scratch's 2048", ctx->g.grid);
  if (rc == -2) return nsx_fail(ctx, -1, "nsx_glob_cfl_noalloc: the scratch of device %d is not set up", ctx->dev);
  if (rc) return nsx_fail_hip(ctx, (hipError_t)rc, "launch_glob_cfl_scratch", __LINE__);
  double r[13];
  HIPCHECK(hipMemcpyAsync(r, res, sizeof r, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHECK(hipStreamSynchronize(ctx->stream));
  out->umax = r[0];
  out->vmax = r[1];
  out->wmax = r[2];
  out->cfl = r[3];
  out->aveu = r[4];
  out->avev = r[5];
  out->avew = r[6];
  out->flux_total = r[7];
  out->flux_max = r[8];
  out->flux_bc = r[9];
  out->flux_bot = r[10];
  out->flux_top = r[11];
  out->flux_side = r[12];
  return 0;
}

// ------------------------------------------------------------ statistics --
// sm_init (stats.F90:34-84): the accumulators of the channel statistics,
// allocated on first use (twelve cell fields), set to zero; sample count 0
extern "C" int nsx_sm_init(nsx_ctx *ctx) {
  NEED_COMMIT(ctx);
  const size_t nc = (size_t)ctx->N0 * ctx->NS;
  int rc;
  for (int q = 0; q < SM_N; ++q) {
    if (!ctx->sm[q] && (rc = dev_alloc(ctx, &ctx->sm[q], nc))) return rc;
    HIPCHECK(hipMemsetAsync(ctx->sm[q], 0, nc * sizeof(double), ctx->stream));
  }
  if (!ctx->sm_planes && (rc = dev_alloc(ctx, &ctx->sm_planes, (size_t)SM_NPLANE * ctx->NZ))) return rc;
  ctx->sm_n = 0;
  HIPCHECK(hipStreamSynchronize(ctx->stream));
  return 0;
}

// glob_sample (stats.F90:91-197) on resident cell fields; the eddy viscosity
// term comes from NSX_F_EDDY (the reference's KeyLES build)
extern "C" int nsx_glob_sample(nsx_ctx *ctx, int u, int v, int w, int p) {
  NEED_COMMIT(ctx);
  if (!ctx->sm[0]) return nsx_fail(ctx, -1, "nsx_glob_sample: nsx_sm_init has not been called");
  if (kind_ok(ctx, u, FK_CELL) || kind_ok(ctx, v, FK_CELL) || kind_ok(ctx, w, FK_CELL) || kind_ok(ctx, p, FK_CELL))
    return -1;
  StatsBufs sb;
  for (int q = 0; q < SM_N; ++q) sb.a[q] = ctx->sm[q];
  LAUNCH(launch_glob_sample(ctx->g, ctx->stream, ctx->slot[u], ctx->slot[v], ctx->slot[w], ctx->slot[p],
                           ctx->slot[NSX_F_EDDY], sb));
  ctx->sm_n += 1;
  return 0;
}

extern "C" int nsx_sm_count(nsx_ctx *ctx, int *n) {
  NEED_COMMIT(ctx);
  if (!n) return nsx_fail(ctx, -1, "nsx_sm_count: n is NULL");
  *n = ctx->sm_n;
  return 0;
}

// one accumulator (NSX_SM_*) into host(N_CELL,NZ), rows 1..N_CELL0
extern "C" int nsx_sm_get(nsx_ctx *ctx, int q, double *host) {
  NEED_COMMIT(ctx);
  if (q < 0 || q >= SM_N) return nsx_fail(ctx, -1, "nsx_sm_get: no accumulator %d", q);
  if (!ctx->sm[q]) return nsx_fail(ctx, -1, "nsx_sm_get: nsx_sm_init has not been called");
  if (!host) return nsx_fail(ctx, -1, "nsx_sm_get: host is NULL");
  const int N0 = ctx->N0, NZ = ctx->NZ, NS = ctx->NS;
  const size_t NC = (size_t)ctx->NC;
  std::vector<double> buf((size_t)N0 * NS);
  HIPCHECK(hipMemcpyAsync(buf.data(), ctx->sm[q], buf.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHECK(hipStreamSynchronize(ctx->stream));
  for (int k = 0; k < NZ; ++k)
    for (int i = 0; i < N0; ++i) host[(size_t)i + (size_t)k * NC] = buf[(size_t)i * NS + k];
  return 0;
}

// the plane sums of glob_put_plane_mm (stats.F90:266-279), computed on the
// device in the reference's order: planes[q*NZ + k-1] = sum over i of
// sm_q(i,k)/ns for the nine velocity moments and k = 2..NZ-1, 0 at k = 1, NZ
extern "C" int nsx_sm_planes(nsx_ctx *ctx, double *planes, int *ns) {
  NEED_COMMIT(ctx);
  if (!ctx->sm[0]) return nsx_fail(ctx, -1, "nsx_sm_planes: nsx_sm_init has not been called");
  if (!planes) return nsx_fail(ctx, -1, "nsx_sm_planes: planes is NULL");
  StatsBufs sb;
  for (int q = 0; q < SM_N; ++q) sb.a[q] = ctx->sm[q];
  const size_t n = (size_t)SM_NPLANE * ctx->NZ;
  HIPCHECK(hipMemsetAsync(ctx->sm_planes, 0, n * sizeof(double), ctx->stream));
  LAUNCH(launch_plane_sums(ctx->g, ctx->stream, sb, ctx->sm_n, ctx->sm_planes));
  HIPCHECK(hipMemcpyAsync(planes, ctx->sm_planes, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHECK(hipStreamSynchronize(ctx->stream));
  if (ns) *ns = ctx->sm_n;
  return 0;
}

// glob_put_plane_mm (stats.F90:204-391): plane sums on the device, then the
// three files sm-, srms-, stau-NNNNN.dat in dir (the reference's
// ../output/Stats)
extern "C" int nsx_glob_put_plane_mm(nsx_ctx *ctx, const char *dir) {
  NEED_COMMIT(ctx);
  if (!dir) return nsx_fail(ctx, -1, "nsx_glob_put_plane_mm: dir is NULL");
  std::vector<double> planes((size_t)SM_NPLANE * ctx->NZ);
  int ns = 0, rc;
  if ((rc = nsx_sm_planes(ctx, planes.data(), &ns))) return rc;
  if (nsx_write_plane_stats(planes.data(), ctx->NZ, ctx->N0, ns, ctx->Re, dir))
    return nsx_fail(ctx, -5, "nsx_glob_put_plane_mm: cannot write the statistics files in %s", dir);
  return 0;
}

// -------------------------------------------------------------- snapshots --
extern "C" int nsx_savetec_vertex(nsx_ctx *ctx, const char *dir, int counter, double time, const double *xvt,
                                  const double *yvt, const double *zvt) {
  NEED_COMMIT(ctx);
  if (!dir || !xvt || !yvt || !zvt) return nsx_fail(ctx, -1, "nsx_savetec_vertex: NULL argument");
  const size_t nv = (size_t)ctx->NV * (ctx->NZ - 1);
  std::vector<double> f(4 * nv);
  for (int q = 0; q < 4; ++q) {
    int rc = nsx_get_field(ctx, NSX_F_UFV + q, f.data() + q * nv, 0);
    if (rc) return rc;
  }
  if (nsx_write_tec_vertex(dir, counter, time, ctx->N0, ctx->NV, ctx->NZ, ctx->No_vp.data(), xvt, yvt, zvt,
                           f.data(), f.data() + nv, f.data() + 2 * nv, f.data() + 3 * nv))
    return nsx_fail(ctx, -5, "nsx_savetec_vertex: cannot write V-%04d.tec in %s", counter, dir);
  return 0;
}

extern "C" int nsx_savetec_center(nsx_ctx *ctx, const char *dir, int counter, double time, const double *xvt,
                                  const double *yvt, const double *zvt, int u, int v, int w, int p) {
  NEED_COMMIT(ctx);
  if (!dir || !xvt || !yvt || !zvt) return nsx_fail(ctx, -1, "nsx_savetec_center: NULL argument");
  if (kind_ok(ctx, u, FK_CELL) || kind_ok(ctx, v, FK_CELL) || kind_ok(ctx, w, FK_CELL) || kind_ok(ctx, p, FK_CELL))
    return -1;
  const size_t nc = (size_t)ctx->NC * ctx->NZ;
  std::vector<double> f(4 * nc);
  const int ids[4] = {u, v, w, p};
  for (int q = 0; q < 4; ++q) {
    int rc = nsx_get_field(ctx, ids[q], f.data() + q * nc, 0);
    if (rc) return rc;
  }
  if (nsx_write_tec_center(dir, counter, time, ctx->N0, ctx->NC, ctx->NV, ctx->NZ, ctx->No_vp.data(), xvt, yvt, zvt,
                           f.data(), f.data() + nc, f.data() + 2 * nc, f.data() + 3 * nc))
    return nsx_fail(ctx, -5, "nsx_savetec_center: cannot write C-%05d.tec in %s", counter, dir);
  return 0;
}

extern "C" int nsx_save_paraview_vertex(nsx_ctx *ctx, const char *dir, int counter, const double *xvt,
                                        const double *yvt, const double *zvt) {
  NEED_COMMIT(ctx);
  if (!dir || !xvt || !yvt || !zvt) return nsx_fail(ctx, -1, "nsx_save_paraview_vertex: NULL argument");
  const size_t nv = (size_t)ctx->NV * (ctx->NZ - 1);
  std::vector<double> f(3 * nv);
  for (int q = 0; q < 3; ++q) {
    int rc = nsx_get_field(ctx, NSX_F_UFV + q, f.data() + q * nv, 0);
    if (rc) return rc;
  }
  if (nsx_write_vtk_vertex(dir, counter, ctx->N0, ctx->NV, ctx->NZ, ctx->No_vp.data(), xvt, yvt, zvt, f.data(),
                           f.data() + nv, f.data() + 2 * nv, nullptr, nullptr, nullptr))
    return nsx_fail(ctx, -5, "nsx_save_paraview_vertex: cannot write V-%04d.vtk in %s", counter, dir);
  return 0;
}

extern "C" int nsx_savetec_vc(nsx_ctx *ctx, const char *dir, int counter, double time, const double *xvt,
                              const double *yvt, const double *zvt, const double *xct, const double *yct,
                              const double *zct, int u, int v, int w, int p) {
  NEED_COMMIT(ctx);
  if (!dir || !xvt || !yvt || !zvt || !xct || !yct || !zct)
    return nsx_fail(ctx, -1, "nsx_savetec_vc: NULL argument");
  if (kind_ok(ctx, u, FK_CELL) || kind_ok(ctx, v, FK_CELL) || kind_ok(ctx, w, FK_CELL) || kind_ok(ctx, p, FK_CELL))
    return -1;
  const size_t nv = (size_t)ctx->NV * (ctx->NZ - 1), nc = (size_t)ctx->NC * ctx->NZ;
  std::vector<double> fv(4 * nv), fc(4 * nc);
  const int ids[4] = {u, v, w, p};
  for (int q = 0; q < 4; ++q) {
    int rc = nsx_get_field(ctx, NSX_F_UFV + q, fv.data() + q * nv, 0);
    if (!rc) rc = nsx_get_field(ctx, ids[q], fc.data() + q * nc, 0);
    if (rc) return rc;
  }
  if (nsx_write_tec_vc(dir, counter, time, ctx->N0, ctx->NC, ctx->NV, ctx->NZ, ctx->No_vp.data(), xvt, yvt, zvt, xct,
                       yct, zct, fv.data(), fv.data() + nv, fv.data() + 2 * nv, fv.data() + 3 * nv, fc.data(),
                       fc.data() + nc, fc.data() + 2 * nc, fc.data() + 3 * nc))
    return nsx_fail(ctx, -5, "nsx_savetec_vc: cannot write VC-%05d.tec in %s", counter, dir);
  return 0;
}

// ------------------------------------------------- differentiation form --
// The derivative of the time-averaged bottom-wall shear of the channel with
// respect to the per-prism streamwise body force (NSX_F_FORCE), by the
// discrete adjoint of the RK2 steps (nsmp3D.F90:448-548) and projection
// stages (hydro.F90:8-501), with checkpoints every few steps and the steps
// between recomputed; and the same derivative along one direction by the
// tangent-linear model.  Solves run in FP64 (BiCGSTAB) to eps.

static int adj_field(nsx_ctx *ctx, double **p, int kind) {
  const size_t n = field_elems(ctx, kind == FK_VERT ? FK_VERT : FK_CELL);
  int rc = dev_alloc(ctx, p, n);
  if (rc) return rc;
  HIPCHECK(hipMemsetAsync(*p, 0, n * sizeof(double), ctx->stream));
  return 0;
}

static int adj_alloc(nsx_ctx *ctx, size_t nckpt, size_t nseg) {
  AdjWork &W = ctx->adj;
  int rc;
  if (!W.ready) {
    for (int q = 0; q < 9; ++q) {
      if ((rc = adj_field(ctx, &W.fin[q], FK_CELL)) || (rc = adj_field(ctx, &W.Sb[q], FK_CELL))) return rc;
    }
    for (int q = 0; q < 4; ++q) {
      if ((rc = adj_field(ctx, &W.Yb[q], FK_CELL)) || (rc = adj_field(ctx, &W.Xb[0][q], FK_CELL)) ||
          (rc = adj_field(ctx, &W.Xb[1][q], FK_CELL)))
        return rc;
      for (int t = 0; t < 3; ++t)
        if ((rc = adj_field(ctx, &W.T[t][q], FK_CELL))) return rc;
    }
    for (int q = 0; q < 5; ++q)
      if ((rc = adj_field(ctx, &W.Fb[q], FK_CELL)) || (rc = adj_field(ctx, &W.Fib[0][q], FK_CELL)) ||
          (rc = adj_field(ctx, &W.Fib[1][q], FK_CELL)) || (rc = adj_field(ctx, &W.TF[q], FK_CELL)))
        return rc;
    for (int q = 0; q < 3; ++q)
      if ((rc = adj_field(ctx, &W.usb[q], FK_CELL)) || (rc = adj_field(ctx, &W.Tus[q], FK_CELL)) ||
          (rc = adj_field(ctx, &W.vb[q], FK_VERT)) || (rc = adj_field(ctx, &W.V[q], FK_VERT)) ||
          (rc = adj_field(ctx, &W.TV[q], FK_VERT)))
        return rc;
    if ((rc = adj_field(ctx, &W.eb, FK_CELL)) || (rc = adj_field(ctx, &W.lam, FK_CELL)) ||
        (rc = adj_field(ctx, &W.E, FK_CELL)) || (rc = adj_field(ctx, &W.TE, FK_CELL)) ||
        (rc = adj_field(ctx, &W.Trhs, FK_CELL)) || (rc = adj_field(ctx, &W.zero, FK_CELL)) ||
        (rc = adj_field(ctx, &W.pvb, FK_VERT)) || (rc = adj_field(ctx, &W.pvb2, FK_VERT)) ||
        (rc = adj_field(ctx, &W.valb, FK_VERT)) || (rc = adj_field(ctx, &W.Tpv, FK_VERT)))
      return rc;
    if ((rc = dev_alloc(ctx, &W.J, 2)) || (rc = dev_alloc(ctx, &W.Jpart, 64))) return rc;
    W.ready = true;
  }
  while (W.ckpt.size() < 9 * nckpt) {
    double *p;
    if ((rc = adj_field(ctx, &p, FK_CELL))) return rc;
    W.ckpt.push_back(p);
  }
  while (W.seg.size() < 18 * nseg) {
    double *p;
    if ((rc = adj_field(ctx, &p, FK_CELL))) return rc;
    W.seg.push_back(p);
  }
  return 0;
}


// the state of pool set `set` (u, v, w, p) and the faces into dst[9], or back
static int save_state(nsx_ctx *ctx, int set, double *const dst[9]) {
  for (int q = 0; q < 9; ++q) {
    const double *src = q < 4 ? ctx->pool_cell[set][q] : ctx->slot[NSX_F_U1FACE + q - 4];
    HIPCHECK(hipMemcpyAsync(dst[q], src, cell_bytes(ctx), hipMemcpyDeviceToDevice, ctx->stream));
  }
  return 0;
}

static int load_state(nsx_ctx *ctx, int set, const double *const src[9]) {
  for (int q = 0; q < 9; ++q) {
    double *dst = q < 4 ? ctx->pool_cell[set][q] : ctx->slot[NSX_F_U1FACE + q - 4];
    HIPCHECK(hipMemcpyAsync(dst, src[q], cell_bytes(ctx), hipMemcpyDeviceToDevice, ctx->stream));
  }
  return 0;
}

// 1 / (sum of the column areas): the mean over the bottom wall
static double inv_area(const nsx_ctx *ctx) {
  double a = 0.0;
  for (int i = 0; i < ctx->N0; ++i) a += ctx->areaCell[i];
  return 1.0 / a;
}

// One RK2 step of the resident state (as nsx_step), storing its input state
// S^n and the second stage's input S^a when sn / sa are given; J += scale x
// the bottom-wall shear of the new level
static int fwd_step(nsx_ctx *ctx, double *const *sn, double *const *sa, double scale) {
  const int N = ctx->set_n, A = (N + 1) % 3, B = (N + 2) % 3;
  int rc;
  if (sn && (rc = save_state(ctx, N, sn))) return rc;
  if ((rc = stage(ctx, N, A, 0))) return rc;
  if (ctx->info.exit_code == NSX_SOLVE_MAX_ITERS) ctx->adj.unconverged = true;
  if (sa && (rc = save_state(ctx, A, sa))) return rc;
  if ((rc = stage(ctx, A, B, 0))) return rc;
  if (ctx->info.exit_code == NSX_SOLVE_MAX_ITERS) ctx->adj.unconverged = true;
  const double *un[3] = {ctx->pool_cell[N][0], ctx->pool_cell[N][1], ctx->pool_cell[N][2]};
  double *unp[3] = {ctx->pool_cell[B][0], ctx->pool_cell[B][1], ctx->pool_cell[B][2]};
  LAUNCH(launch_rk2avg(ctx->g, ctx->stream, un, unp));
  ctx->set_n = B;
  point_slots(ctx, B, A, B);
  if (scale != 0.0) LAUNCH(launch_wall_shear(ctx->g, ctx->stream, ctx->pool_cell[B][0], scale, ctx->adj.Jpart,
                                             ctx->adj.J));
  return 0;
}

static void vert_acts(const Geo &g, const signed char *act[4]) {
  act[0] = g.vact_u;
  act[1] = g.vact_v;
  act[2] = g.vact_w;
  act[3] = g.vact_p;
}

// the stage's vertex values of u, v, w (tagdir 2) and eddy viscosity from
// its input X (ghost layers filled first, as the stage does)
static int stage_base(nsx_ctx *ctx, double *const X[4]) {
  const Geo &g = ctx->g;
  hipStream_t s = ctx->stream;
  AdjWork &W = ctx->adj;
  for (int q = 0; q < 3; ++q) BC_CENTER(X[q], q + 1);
  BC_CENTER(X[3], 0);
  const double *uvw[3] = {X[0], X[1], X[2]};
  const int vk[3] = {1, 2, 3};
  int rc = vertex_of(ctx, 3, uvw, W.V, vk, 2);
  if (rc) return rc;
  if (ctx->Iturbu == 1)
    LAUNCH(launch_les(g, s, X[0], X[1], X[2], W.V[0], W.V[1], W.V[2], W.E));
  else
    HIPCHECK(hipMemsetAsync(W.E, 0, cell_bytes(ctx), s));
  return 0;
}

// the adjoint of one stage (X, Fin) -> (Y, Fout): from the bars Yb of Y
// (ghost layers included) and Fb of Fout, the bars Xb of X and Fib of Fin;
// fb += the bar of the body force.  Yb and Fb are used up.
static int adj_stage(nsx_ctx *ctx, double *const X[4], const double *const Fin[5], double *const Yb[4],
                     double *const Fb[5], double *const Xb[4], double *const Fib[5], double *fb) {
  const Geo &g = ctx->g;
  hipStream_t s = ctx->stream;
  AdjWork &W = ctx->adj;
  const size_t cb = cell_bytes(ctx), vbytes = (size_t)ctx->NV * (ctx->NZ - 1) * sizeof(double);
  for (int q = 0; q < 4; ++q) HIPCHECK(hipMemsetAsync(Xb[q], 0, cb, s));
  for (int q = 0; q < 5; ++q) HIPCHECK(hipMemsetAsync(Fib[q], 0, cb, s));
  for (int q = 0; q < 3; ++q) {
    HIPCHECK(hipMemsetAsync(W.usb[q], 0, cb, s));
    HIPCHECK(hipMemsetAsync(W.vb[q], 0, vbytes, s));
  }
  HIPCHECK(hipMemsetAsync(W.eb, 0, cb, s));
  HIPCHECK(hipMemsetAsync(W.pvb, 0, vbytes, s));
  HIPCHECK(hipMemsetAsync(W.pvb2, 0, vbytes, s));
  int rc = stage_base(ctx, X);
  if (rc) return rc;
  const signed char *act[4];
  vert_acts(g, act);
  // F'' = F' - dt grad_n(Yp, Pv'): the face bars pass through
  LAUNCH(launch_adj_newu(g, s, Fb, Yb[3], W.pvb2));
  // Pv' = vertex values of Yp
  LAUNCH(launch_adj_vertex(g, s, W.pvb2, act[3], W.valb, Yb[3]));
  // Y = U* - dt grad_LSM(Yp), then the ghost layers of Y
  for (int q = 0; q < 3; ++q) LAUNCH(launch_adj_ghost(g, s, Yb[q], q + 1));
  const double *yb[3] = {Yb[0], Yb[1], Yb[2]};
  LAUNCH(launch_adj_corrector(g, s, yb, W.usb, Yb[3]));
  // Yp = N0 A^+ N1 rhsp with its Neumann ghost layers: rhsp_bar = N0 (A^T)^+ N1 Yp_bar
  LAUNCH(launch_adj_ghost(g, s, Yb[3], 0));
  // The adjoint solves are driven a hundred times tighter than eps: their
  // right-hand sides, the transposed corrector and face update of bars seeded
  // at the wall, are concentrated where the forward ones are spread, and the
  // same relative residual leaves a larger error in the smooth part of the
  // solution the gradient integrates.  The FP64 single-reduction CG (its
  // true residual tested in FP64) reaches such tolerances where BiCGSTAB's
  // recursive residual may stall; A^T is symmetric up to the last bit.
  HIPCHECK(hipMemsetAsync(W.lam, 0, cb, s));
  LAUNCH(launch_gnorm_sum(g, s, Yb[3], ctx->part));
  {
    const double eps0 = ctx->eps;
    ctx->eps = std::max(eps0 * 1e-2, 1e-16);
    ctx->g.trans = 1;
    rc = solve_cgs(ctx, Yb[3], W.lam, W.lam, true);
    ctx->g.trans = 0;
    ctx->eps = eps0;
  }
  if (rc) return rc;
  // F' = faces of U*, rhsp = div F' / dt - lagged flux of Pv
  const double *fbc[5] = {Fb[0], Fb[1], Fb[2], Fb[3], Fb[4]};
  LAUNCH(launch_adj_source(g, s, fbc, W.lam, W.usb, W.pvb));
  // U* = X + dt r, then the ghost layers of U*
  for (int q = 0; q < 3; ++q) LAUNCH(launch_adj_ghost(g, s, W.usb[q], q + 1));
  const double *uvw[3] = {X[0], X[1], X[2]}, *vtx[3] = {W.V[0], W.V[1], W.V[2]};
  const double *usb[3] = {W.usb[0], W.usb[1], W.usb[2]};
  double *xb[3] = {Xb[0], Xb[1], Xb[2]};
  LAUNCH(launch_adj_momentum(g, s, uvw, vtx, Fin, W.E, usb, xb, W.vb, Fib, W.eb, fb));
  // Pv = vertex values of p
  LAUNCH(launch_adj_vertex(g, s, W.pvb, act[3], W.valb, Xb[3]));
  // the eddy viscosity
  if (ctx->Iturbu == 1) LAUNCH(launch_adj_les(g, s, uvw, vtx, W.eb, xb, W.vb));
  // V = vertex values of u, v, w
  for (int q = 0; q < 3; ++q) LAUNCH(launch_adj_vertex(g, s, W.vb[q], act[q], W.valb, Xb[q]));
  // the ghost layers of X
  for (int q = 0; q < 3; ++q) LAUNCH(launch_adj_ghost(g, s, Xb[q], q + 1));
  LAUNCH(launch_adj_ghost(g, s, Xb[3], 0));
  return 0;
}

// the adjoint of one RK2 step from its stored input state Sn and second
// stage input Sa: Sb holds the bars of the new level on entry and those of
// level n on return; fb += the bar of the body force
static int adj_step(nsx_ctx *ctx, double *const Sn[9], double *const Sa[9], double *const Sb[9], double *fb) {
  hipStream_t s = ctx->stream;
  AdjWork &W = ctx->adj;
  const size_t nc = field_elems(ctx, FK_CELL);
  // u, v, w (n+1) = (u(n) + u_b)/2; p and the faces are the second stage's
  for (int q = 0; q < 3; ++q) LAUNCH(launch_axpby(s, nc, 0.5, Sb[q], 0.0, nullptr));
  for (int q = 0; q < 4; ++q) HIPCHECK(hipMemcpyAsync(W.Yb[q], Sb[q], cell_bytes(ctx), hipMemcpyDeviceToDevice, s));
  for (int q = 0; q < 5; ++q)
    HIPCHECK(hipMemcpyAsync(W.Fb[q], Sb[4 + q], cell_bytes(ctx), hipMemcpyDeviceToDevice, s));
  double *Xa[4] = {Sa[0], Sa[1], Sa[2], Sa[3]}, *Xn[4] = {Sn[0], Sn[1], Sn[2], Sn[3]};
  const double *Fa[5] = {Sa[4], Sa[5], Sa[6], Sa[7], Sa[8]}, *Fn[5] = {Sn[4], Sn[5], Sn[6], Sn[7], Sn[8]};
  int rc = adj_stage(ctx, Xa, Fa, W.Yb, W.Fb, W.Xb[0], W.Fib[0], fb);
  if (rc) return rc;
  if ((rc = adj_stage(ctx, Xn, Fn, W.Xb[0], W.Fib[0], W.Xb[1], W.Fib[1], fb))) return rc;
  for (int q = 0; q < 3; ++q) LAUNCH(launch_axpby(s, nc, 1.0, Sb[q], 1.0, W.Xb[1][q]));
  HIPCHECK(hipMemcpyAsync(Sb[3], W.Xb[1][3], cell_bytes(ctx), hipMemcpyDeviceToDevice, s));
  for (int q = 0; q < 5; ++q)
    HIPCHECK(hipMemcpyAsync(Sb[4 + q], W.Fib[1][q], cell_bytes(ctx), hipMemcpyDeviceToDevice, s));
  return 0;
}

// the tangent-linear stage (X, Fin) -> (Y, Fout) along (dX, dF, dforce),
// before the base stage runs (the faces are updated in place by both): dX
// gets its ghost layers, dF goes from Fin' to Fout', dY receives Y'
static int tlm_stage(nsx_ctx *ctx, double *const X[4], double *const dX[4], const double *dforce,
                     double *const dY[4]) {
  const Geo &g = ctx->g;
  hipStream_t s = ctx->stream;
  AdjWork &W = ctx->adj;
  int rc = stage_base(ctx, X);
  if (rc) return rc;
  for (int q = 0; q < 3; ++q) BC_CENTER(dX[q], q + 1);
  BC_CENTER(dX[3], 0);
  const double *uvw[3] = {X[0], X[1], X[2]}, *vtx[3] = {W.V[0], W.V[1], W.V[2]};
  const double *duvw[3] = {dX[0], dX[1], dX[2]};
  const int vk[3] = {1, 2, 3}, pk[3] = {0, 0, 0};
  if ((rc = vertex_of(ctx, 3, duvw, W.TV, vk, 2))) return rc;
  const double *dvtx[3] = {W.TV[0], W.TV[1], W.TV[2]};
  if (ctx->Iturbu == 1)
    LAUNCH(launch_tlm_les(g, s, uvw, vtx, duvw, dvtx, W.TE));
  else
    HIPCHECK(hipMemsetAsync(W.TE, 0, cell_bytes(ctx), s));
  const double *face[5] = {ctx->slot[NSX_F_U1FACE], ctx->slot[NSX_F_U2FACE], ctx->slot[NSX_F_U3FACE],
                           ctx->slot[NSX_F_UTFACE], ctx->slot[NSX_F_UBFACE]};
  const double *dface[5] = {W.TF[0], W.TF[1], W.TF[2], W.TF[3], W.TF[4]};
  LAUNCH(launch_tlm_momentum(g, s, uvw, vtx, face, W.E, duvw, dvtx, dface, W.TE, dforce, W.Tus));
  for (int q = 0; q < 3; ++q) BC_WALLS(W.Tus[q], q + 1);
  const double *dp[3] = {dX[3], nullptr, nullptr};
  double *dpv[3] = {W.Tpv, nullptr, nullptr};
  if ((rc = vertex_of(ctx, 1, dp, dpv, pk, 2))) return rc;
  const double *dus[3] = {W.Tus[0], W.Tus[1], W.Tus[2]};
  LAUNCH(launch_source(g, s, dus, W.TF, nullptr, W.Trhs, W.Tpv, ctx->part));
  // the tangent pressure by the FP64 single-reduction CG, as the adjoint
  // stage solves with A^T: at eps 1e-13 the reference's BiCGSTAB can stall on
  // its recursive residual and break down, depending on the last bits of the
  // right-hand side, where the CG tests its true residual in FP64
  if ((rc = solve_cgs(ctx, W.Trhs, W.zero, dY[3], true))) return rc;
  BC_WALLS(dY[3], 0);
  double *dun[3] = {dY[0], dY[1], dY[2]};
  LAUNCH(launch_corrector(g, s, dY[3], dus, dun));
  for (int q = 0; q < 3; ++q) BC_WALLS(dY[q], q + 1);
  const double *dyp[3] = {dY[3], nullptr, nullptr};
  if ((rc = vertex_of(ctx, 1, dyp, dpv, pk, 2))) return rc;
  LAUNCH(launch_newu(g, s, dY[3], W.Tpv, W.TF));
  return 0;
}

// one RK2 step of the base state with its tangent (pool-aligned tangent
// sets W.T, faces' tangent W.TF); J, dJ += scale x the wall shear of the new
// level and of its tangent
static int tlm_step(nsx_ctx *ctx, const double *dforce, double scale) {
  AdjWork &W = ctx->adj;
  hipStream_t s = ctx->stream;
  const int N = ctx->set_n, A = (N + 1) % 3, B = (N + 2) % 3;
  int rc;
  double *(*P)[4] = ctx->pool_cell;
  if ((rc = tlm_stage(ctx, P[N], W.T[N], dforce, W.T[A]))) return rc;
  if ((rc = stage(ctx, N, A, 0))) return rc;
  if (ctx->info.exit_code == NSX_SOLVE_MAX_ITERS) W.unconverged = true;
  if ((rc = tlm_stage(ctx, P[A], W.T[A], dforce, W.T[B]))) return rc;
  if ((rc = stage(ctx, A, B, 0))) return rc;
  if (ctx->info.exit_code == NSX_SOLVE_MAX_ITERS) W.unconverged = true;
  const double *un[3] = {P[N][0], P[N][1], P[N][2]}, *dn[3] = {W.T[N][0], W.T[N][1], W.T[N][2]};
  double *unp[3] = {P[B][0], P[B][1], P[B][2]}, *dnp[3] = {W.T[B][0], W.T[B][1], W.T[B][2]};
  LAUNCH(launch_rk2avg(ctx->g, s, un, unp));
  LAUNCH(launch_rk2avg(ctx->g, s, dn, dnp));
  ctx->set_n = B;
  point_slots(ctx, B, A, B);
  if (scale != 0.0) {
    LAUNCH(launch_wall_shear(ctx->g, s, P[B][0], scale, W.Jpart, W.J));
    LAUNCH(launch_wall_shear(ctx->g, s, W.T[B][0], scale, W.Jpart, W.J + 1));
  }
  return 0;
}

namespace {
// runs the differentiation form in FP64 and puts back the caller's precision
struct Fp64Scope {
  nsx_ctx *c;
  int p;
  explicit Fp64Scope(nsx_ctx *ctx) : c(ctx), p(ctx->precision) { ctx->precision = NSX_PREC_FP64; }
  ~Fp64Scope() { c->precision = p; }
};
}  // namespace

static int diff_checks(nsx_ctx *ctx) {
  if (ctx->FinRK != 2) return nsx_fail(ctx, -1, "the differentiation form needs RK2 steps (FinRK = 2)");
  if (ctx->TopBC == 2)
    return nsx_fail(ctx, -1, "the differentiation form does not transpose the lid (TopBC = 2): its vertex values "
                             "are constants the tangent and adjoint vertex kernels do not hold");
  if (ctx->has_inflow)
    return nsx_fail(ctx, -1, "the differentiation form does not take inflow boundaries (flag 2): their ghost and "
                             "vertex values hold constants the tangent kernels would add");
  return 0;
}

extern "C" int nsx_set_body_force(nsx_ctx *ctx, int on) {
  NEED_COMMIT(ctx);
  ctx->body_force = on != 0;
  return 0;
}

extern "C" int nsx_wall_shear_run(nsx_ctx *ctx, int nsteps, int every, double *J, int grad) {
  NEED_COMMIT(ctx);
  if (diff_checks(ctx)) return -1;
  if (nsteps < 1 || every < 1) return nsx_fail(ctx, -1, "nsx_wall_shear_run: nsteps and every must be >= 1");
  if (grad >= 0 && kind_ok(ctx, grad, FK_CELL)) return -1;
  if (grad == NSX_F_FORCE) return nsx_fail(ctx, -1, "nsx_wall_shear_run: grad must not be the body force");
  const int nck = (nsteps + every - 1) / every;
  int rc = adj_alloc(ctx, grad >= 0 ? (size_t)nck : 0, grad >= 0 ? (size_t)every : 0);
  if (rc) return rc;
  Fp64Scope fp64(ctx);
  AdjWork &W = ctx->adj;
  hipStream_t s = ctx->stream;
  const double scale = inv_area(ctx) / nsteps;
  W.unconverged = false;
  HIPCHECK(hipMemsetAsync(W.J, 0, 2 * sizeof(double), s));
  for (int n = 0; n < nsteps; ++n) {
    if (grad >= 0 && n % every == 0 && (rc = save_state(ctx, ctx->set_n, &W.ckpt[9 * (size_t)(n / every)]))) return rc;
    if ((rc = fwd_step(ctx, nullptr, nullptr, scale))) return rc;
  }
  double jh = 0.0;
  HIPCHECK(hipMemcpyAsync(&jh, W.J, sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHECK(hipStreamSynchronize(s));
  if (J) *J = jh;
  if (grad < 0) return W.unconverged ? NSX_SOLVE_MAX_ITERS : 0;
  // the reverse sweep, segment by segment from the last checkpoint: recompute
  // the segment's steps storing their stage inputs, then step back through it
  const int nfinal = ctx->set_n;
  if ((rc = save_state(ctx, nfinal, W.fin))) return rc;
  for (int q = 0; q < 9; ++q) HIPCHECK(hipMemsetAsync(W.Sb[q], 0, cell_bytes(ctx), s));
  double *fb = ctx->slot[grad];
  HIPCHECK(hipMemsetAsync(fb, 0, cell_bytes(ctx), s));
  for (int c = nck - 1; c >= 0; --c) {
    const int n0 = c * every, n1 = std::min(nsteps, n0 + every);
    if ((rc = load_state(ctx, ctx->set_n, &W.ckpt[9 * (size_t)c]))) return rc;
    for (int n = n0; n < n1; ++n) {
      double *const *sn = &W.seg[18 * (size_t)(n - n0)];
      if ((rc = fwd_step(ctx, sn, sn + 9, 0.0))) return rc;
    }
    for (int n = n1 - 1; n >= n0; --n) {
      double *const *sn = &W.seg[18 * (size_t)(n - n0)];
      LAUNCH(launch_wall_shear_seed(ctx->g, s, W.Sb[0], scale));
      if ((rc = adj_step(ctx, sn, sn + 9, W.Sb, fb))) return rc;
    }
  }
  if ((rc = load_state(ctx, ctx->set_n, W.fin))) return rc;
  HIPCHECK(hipStreamSynchronize(s));
  return W.unconverged ? NSX_SOLVE_MAX_ITERS : 0;
}

extern "C" int nsx_wall_shear_tangent(nsx_ctx *ctx, int nsteps, int dir, double *J, double *dJ) {
  NEED_COMMIT(ctx);
  if (diff_checks(ctx)) return -1;
  if (nsteps < 1) return nsx_fail(ctx, -1, "nsx_wall_shear_tangent: nsteps must be >= 1");
  if (kind_ok(ctx, dir, FK_CELL)) return -1;
  int rc = adj_alloc(ctx, 0, 0);
  if (rc) return rc;
  Fp64Scope fp64(ctx);
  AdjWork &W = ctx->adj;
  hipStream_t s = ctx->stream;
  for (int t = 0; t < 3; ++t)
    for (int q = 0; q < 4; ++q) HIPCHECK(hipMemsetAsync(W.T[t][q], 0, cell_bytes(ctx), s));
  for (int q = 0; q < 5; ++q) HIPCHECK(hipMemsetAsync(W.TF[q], 0, cell_bytes(ctx), s));
  HIPCHECK(hipMemsetAsync(W.J, 0, 2 * sizeof(double), s));
  const double scale = inv_area(ctx) / nsteps;
  W.unconverged = false;
  for (int n = 0; n < nsteps; ++n)
    if ((rc = tlm_step(ctx, ctx->slot[dir], scale))) return rc;
  double jh[2] = {0.0, 0.0};
  HIPCHECK(hipMemcpyAsync(jh, W.J, sizeof jh, hipMemcpyDeviceToHost, s));
  HIPCHECK(hipStreamSynchronize(s));
  if (J) *J = jh[0];
  if (dJ) *dJ = jh[1];
  return W.unconverged ? NSX_SOLVE_MAX_ITERS : 0;
}

// the nine fields of a state argument: u, v, w, p cell fields and five face
// fields
static int state_ids(nsx_ctx *ctx, const int *ids, double *out[9]) {
  if (!ids) return nsx_fail(ctx, -1, "state field ids are NULL");
  for (int q = 0; q < 9; ++q) {
    if (kind_ok(ctx, ids[q], q < 4 ? FK_CELL : FK_FACE)) return -1;
    out[q] = ctx->slot[ids[q]];
  }
  return 0;
}

extern "C" int nsx_step_tangent(nsx_ctx *ctx, const int *tin, const int *tout, int dforce) {
  NEED_COMMIT(ctx);
  if (diff_checks(ctx)) return -1;
  double *in[9], *out[9];
  if (state_ids(ctx, tin, in) || state_ids(ctx, tout, out)) return -1;
  if (dforce >= 0 && kind_ok(ctx, dforce, FK_CELL)) return -1;
  int rc = adj_alloc(ctx, 0, 0);
  if (rc) return rc;
  Fp64Scope fp64(ctx);
  AdjWork &W = ctx->adj;
  hipStream_t s = ctx->stream;
  const int n0 = ctx->set_n;
  double *slots[12];
  for (int q = 0; q < 12; ++q) slots[q] = ctx->slot[NSX_F_UFN + q];
  if ((rc = save_state(ctx, n0, W.fin))) return rc;
  for (int q = 0; q < 4; ++q)
    HIPCHECK(hipMemcpyAsync(W.T[n0][q], in[q], cell_bytes(ctx), hipMemcpyDeviceToDevice, s));
  for (int q = 0; q < 5; ++q)
    HIPCHECK(hipMemcpyAsync(W.TF[q], in[4 + q], cell_bytes(ctx), hipMemcpyDeviceToDevice, s));
  if (dforce < 0) HIPCHECK(hipMemsetAsync(W.zero, 0, cell_bytes(ctx), s));
  if ((rc = tlm_step(ctx, dforce >= 0 ? ctx->slot[dforce] : W.zero, 0.0))) return rc;
  const int n1 = ctx->set_n;
  for (int q = 0; q < 4; ++q)
    HIPCHECK(hipMemcpyAsync(out[q], W.T[n1][q], cell_bytes(ctx), hipMemcpyDeviceToDevice, s));
  for (int q = 0; q < 5; ++q)
    HIPCHECK(hipMemcpyAsync(out[4 + q], W.TF[q], cell_bytes(ctx), hipMemcpyDeviceToDevice, s));
  // the resident state is left as it was
  ctx->set_n = n0;
  for (int q = 0; q < 12; ++q) ctx->slot[NSX_F_UFN + q] = slots[q];
  if ((rc = load_state(ctx, n0, W.fin))) return rc;
  HIPCHECK(hipStreamSynchronize(s));
  return 0;
}

extern "C" int nsx_step_adjoint(nsx_ctx *ctx, const int *bin, const int *bout, int gforce) {
  NEED_COMMIT(ctx);
  if (diff_checks(ctx)) return -1;
  double *in[9], *out[9];
  if (state_ids(ctx, bin, in) || state_ids(ctx, bout, out)) return -1;
  if (kind_ok(ctx, gforce, FK_CELL)) return -1;
  int rc = adj_alloc(ctx, 0, 1);
  if (rc) return rc;
  Fp64Scope fp64(ctx);
  AdjWork &W = ctx->adj;
  hipStream_t s = ctx->stream;
  const int n0 = ctx->set_n;
  double *slots[12];
  for (int q = 0; q < 12; ++q) slots[q] = ctx->slot[NSX_F_UFN + q];
  if ((rc = save_state(ctx, n0, W.fin))) return rc;
  double *const *sn = &W.seg[0];
  if ((rc = fwd_step(ctx, sn, sn + 9, 0.0))) return rc;
  for (int q = 0; q < 9; ++q)
    HIPCHECK(hipMemcpyAsync(W.Sb[q], in[q], cell_bytes(ctx), hipMemcpyDeviceToDevice, s));
  HIPCHECK(hipMemsetAsync(ctx->slot[gforce], 0, cell_bytes(ctx), s));
  if ((rc = adj_step(ctx, sn, sn + 9, W.Sb, ctx->slot[gforce]))) return rc;
  for (int q = 0; q < 9; ++q)
    HIPCHECK(hipMemcpyAsync(out[q], W.Sb[q], cell_bytes(ctx), hipMemcpyDeviceToDevice, s));
  ctx->set_n = n0;
  for (int q = 0; q < 12; ++q) ctx->slot[NSX_F_UFN + q] = slots[q];
  if ((rc = load_state(ctx, n0, W.fin))) return rc;
  HIPCHECK(hipStreamSynchronize(s));
  return 0;
}

// ---------------------------------------------------------------- timing --
extern "C" int nsx_set_precision(nsx_ctx *ctx, int precision, int refine, double tol, int max_passes) {
  if (!ctx) return -1;
  if (precision < NSX_PREC_FP64 || precision > NSX_PREC_MIXED_FLUX32)
    return nsx_fail(ctx, -1, "unknown precision %d", precision);
  if (refine >= 0 && refine <= 2)
    return nsx_fail(ctx, -1,
                    "refinement %d (an FP32 BiCGSTAB form) is not built: FP32 BiCGSTAB breaks down on this "
                    "operator; use NSX_REFINE_CGSR_BLOCK",
                    refine);
  if (refine != NSX_REFINE_CG && refine != NSX_REFINE_CGSR && refine != NSX_REFINE_CGSR_BLOCK &&
      refine != NSX_REFINE_CGSR_LINE && refine != NSX_REFINE_MODAL)
    return nsx_fail(ctx, -1, "unknown refinement %d", refine);
  const bool sr = refine != NSX_REFINE_CG && refine != NSX_REFINE_MODAL;
  // a window of the line preconditioner holds whole columns of four-layer quads
  if (refine == NSX_REFINE_CGSR_LINE && ctx->g.NQ > NSX_BLOCK)
    return nsx_fail(ctx, -1, "line preconditioner: %d layers per column exceed %d", ctx->g.NS, 4 * NSX_BLOCK);
  if (!(tol > 0.0 && tol < 1.0)) return nsx_fail(ctx, -1, "tolerance %g outside (0,1)", tol);
  // the single-reduction recurrences drift faster in FP32: with updates
  // less often than about every 1.3 decades the solve stalls at its cap
  if (sr && tol < 0.05)
    return nsx_fail(ctx, -1, "single-reduction CG needs delta >= 0.05 (got %g); NSX_REFINE_CG takes smaller", tol);
  if (max_passes < 1) return nsx_fail(ctx, -1, "max_passes %d < 1", max_passes);
  ctx->precision = precision;
  ctx->refine = refine;
  ctx->delta_cg = tol;
  return 0;
}

extern "C" int nsx_set_initial_guess(nsx_ctx *ctx, int mode) {
  if (!ctx) return -1;
  if (mode != NSX_GUESS_PREVIOUS && mode != NSX_GUESS_EXTRAPOLATE)
    return nsx_fail(ctx, -1, "unknown initial-guess mode %d", mode);
  if (mode == NSX_GUESS_EXTRAPOLATE && !ctx->pguess) {
    if (!ctx->committed) return nsx_fail(ctx, -1, "nsx_set_initial_guess: the mesh is not committed");
    const size_t nc = field_elems(ctx, FK_CELL);
    int rc;
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b)
        if ((rc = dev_alloc(ctx, &ctx->phist[a][b], nc))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->pguess, nc))) return rc;
  }
  ctx->guess_mode = mode;
  reset_guess(ctx);
  return 0;
}

extern "C" int nsx_timing_enable(nsx_ctx *ctx, int on) {
  if (!ctx) return -1;
  ctx->timing = on != 0;
  return 0;
}

extern "C" int nsx_timing_get(nsx_ctx *ctx, const char *name, double *total_ms, long long *launches) {
  if (!ctx || !name) return -1;
  const Timer *t = nullptr;
  static const char *names[3] = {"bicg_a", "bicg_b", "bicg_c"};
  for (int q = 0; q < 3; ++q)
    if (!strcmp(name, names[q])) t = &ctx->t_k[q];
  if (!strcmp(name, "stage")) t = &ctx->t_stage;
  static const char *names_cg[2] = {"cg32_w", "cg32_x"};
  for (int q = 0; q < 2; ++q)
    if (!strcmp(name, names_cg[q])) t = &ctx->t_cg[q];
  if (!strcmp(name, "cg32_sr")) t = &ctx->t_cgs;
  if (!strcmp(name, "cg64_sr")) t = &ctx->t_cgs64;
  static const char *names_mod[5] = {"modal_fwd", "modal_mg", "modal_bwd", "modal_spmv", "modal_update"};
  for (int q = 0; q < 5; ++q)
    if (!strcmp(name, names_mod[q])) t = &ctx->t_mod[q];
  if (!t) return nsx_fail(ctx, -1, "unknown timer %s", name);
  if (total_ms) *total_ms = t->ms;
  if (launches) *launches = t->n;
  return 0;
}

extern "C" int nsx_timing_reset(nsx_ctx *ctx) {
  if (!ctx) return -1;
  for (int q = 0; q < 3; ++q) ctx->t_k[q] = Timer{};
  ctx->t_stage = ctx->t_cg[0] = ctx->t_cg[1] = ctx->t_cgs = ctx->t_cgs64 = Timer{};
  for (int q = 0; q < 5; ++q) ctx->t_mod[q] = Timer{};
  return 0;
}
