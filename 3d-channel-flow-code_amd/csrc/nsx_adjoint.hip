// nsx_adjoint.hip -- the differentiation form: hand-written tangent-linear
// and adjoint kernels of the projection stage (hydro.F90:8-501) and of the
// RK2 step (nsmp3D.F90:448-548), for the derivative of an objective with
// respect to a per-prism streamwise body force.
//
// The stage is, with X = (u, v, w, p) its input, F the face velocities
// U1FACE..UBFACE and f the body force:
//   1  ghost layers of X (BCvelcenter3D, BCpcenter)            linear
//   2  V  = vertex values of u, v, w (tagdir 2)                linear
//   3  E  = Smagorinsky eddy viscosity of (X, V)               nonlinear
//   4  U* = X + dt (f + (E + nu) diff(X, V) - adv(X, V; F))    nonlinear
//   5  Pv = vertex values of p                                 linear
//   6  F' = faces of U*, rhsp = div F' / dt - lagged flux(Pv)  linear
//   7  Yp = N0 A^+ N1 rhsp (the pressure solve)                linear
//   8  Y  = U* - dt grad_LSM(Yp)                               linear
//   9  Pv'= vertex values of Yp                                linear
//   10 F'' = F' - dt grad_n(Yp, Pv')                           linear
// The tangent-linear stage applies the forward kernels of the linear steps
// to the tangents and the kernels below to steps 3 and 4.  The adjoint stage
// runs the steps backwards with the transposed kernels below; the transposed
// pressure solve is the solver on A^T (N0 and N1 are each other's
// transposes).  A transposed gather is a scatter: every adjoint kernel adds
// its contributions to neighbouring prisms and vertices with atomic adds, so
// its sums are exact transposes up to the order of rounding.  Ghost layers
// that a forward step fills from the interior are folded back into the
// interior by k_adj_ghost.
//
// All arithmetic is FP64.  Bytes per prism (adjoint kernels, fused): momentum
// reads u, v, w, 3 vertex fields, 5 faces, eddy and the three bars and
// scatters into 3 + 3 vertex + 5 + 1 fields; these kernels run once per stage
// of the backward sweep and are not on the throughput path.
#include "nsx_kernels.h"

#define FOR_EACH_CELL(g, c)                                                         \
  uint32_t beg_, end_;                                                              \
  block_range((g).ncell, beg_, end_);                                               \
  for (uint32_t base_ = beg_; base_ < end_; base_ += NSX_BLOCK)                     \
    for (uint32_t c = base_ + threadIdx.x, once_ = 1; once_ && c < end_; once_ = 0)

namespace {

__device__ inline void aadd(double *p, double v) { unsafeAtomicAdd(p, v); }

// calcul_face's five face values of f at (i,k) (interpolationFace.F90:85-111),
// as k_momentum and k_les form them
__device__ inline void faces5(const Geo &g, int i, int k, size_t s, const double *f, const double *fv,
                              double (&out)[5]) {
  const double fi = f[s];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int e = i * 3 + j;
    const int nb = g.nbr[e];
    const int nv1 = g.vtx[e], nv2 = g.vtx[i * 3 + (j + 1) % 3];
    const double extra = g.doe[e] * (fv[(size_t)nv2 * g.NL + k] - fv[(size_t)nv1 * g.NL + k]) / g.dlVV[e];
    out[j] = f[(size_t)nb * g.NS + k] * 0.5 + fi * 0.5 + extra;
  }
  out[3] = 0.5 * (fi + f[s + 1]);
  out[4] = 0.5 * (fi + f[s - 1]);
}

// transpose of faces5: the bars fb of the five face values into the cell
// bars (own prism returned, neighbours and k +- 1 added atomically) and the
// vertex bars
__device__ inline double faces5_adj(const Geo &g, int i, int k, size_t s, const double (&fb)[5], double *bar,
                                    double *vbar) {
  double own = 0.0;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int e = i * 3 + j;
    const int nb = g.nbr[e];
    const int nv1 = g.vtx[e], nv2 = g.vtx[i * 3 + (j + 1) % 3];
    aadd(bar + (size_t)nb * g.NS + k, 0.5 * fb[j]);
    own += 0.5 * fb[j];
    const double c = fb[j] * g.doe[e] / g.dlVV[e];
    aadd(vbar + (size_t)nv2 * g.NL + k, c);
    aadd(vbar + (size_t)nv1 * g.NL + k, -c);
  }
  own += 0.5 * fb[3] + 0.5 * fb[4];
  aadd(bar + s + 1, 0.5 * fb[3]);
  aadd(bar + s - 1, 0.5 * fb[4]);
  return own;
}

// ------------------------------------------------------- Smagorinsky ----
// sgles's strain rate (sgm_les.F90:193-247): the nine gradients of the
// Green-Gauss fvm_gradient (dvdx, dvdy with the edge 2/3 swap of
// sgm_les.F90:444-446,453-455), from the faces of u, v, w
struct Grad9 {
  double ux, uy, uz, vx, vy, vz, wx, wy, wz;
};

__device__ inline Grad9 grad9(const Geo &g, int i, int k, const double (&U)[5], const double (&V)[5],
                              const double (&W)[5]) {
  const int e = i * 3;
  const double x1 = g.xnn[e], x2 = g.xnn[e + 1], x3 = g.xnn[e + 2];
  const double y1 = g.ynn[e], y2 = g.ynn[e + 1], y3 = g.ynn[e + 2];
  const double l1 = g.dlVV[e], l2 = g.dlVV[e + 1], l3 = g.dlVV[e + 2];
  const double a = g.area[i], dz = g.dsig[k];
  Grad9 d;
  d.ux = (U[0] * x1 * l1 + U[1] * x2 * l2 + U[2] * x3 * l3) / a;
  d.vx = (V[0] * x1 * l1 + V[1] * x2 * l3 + V[2] * x3 * l2) / a;
  d.wx = (W[0] * x1 * l1 + W[1] * x2 * l2 + W[2] * x3 * l3) / a;
  d.uy = (U[0] * y1 * l1 + U[1] * y2 * l2 + U[2] * y3 * l3) / a;
  d.vy = (V[0] * y1 * l1 + V[1] * y2 * l3 + V[2] * y3 * l2) / a;
  d.wy = (W[0] * y1 * l1 + W[1] * y2 * l2 + W[2] * y3 * l3) / a;
  d.uz = (U[3] - U[4]) / dz;
  d.vz = (V[3] - V[4]) / dz;
  d.wz = (W[3] - W[4]) / dz;
  return d;
}

// the damped length scale squared of sgles/sgdamp (sgm_les.F90:212-236,
// 1260-1278): a function of the geometry only, as k_les forms it
__device__ inline double les_nut(const Geo &g, int i, int k) {
  const double Cs = 0.1, karman = (double)0.415f;
  double DD = Cs * pow(g.area[i] * g.dsig[k] * 1.0, 1.0 / 3.0);
  const double lwall = karman * g.walld[k];
  const double dn1 = (1.0 / DD) * (1.0 / DD) + (1.0 / lwall) * (1.0 / lwall);
  const double del = sqrt(1.0 / dn1);
  return del * del;
}

// Q = 2 S_ij S_ij and its derivative along the tangent gradients dt
__device__ inline double strain_q(const Grad9 &d) {
  const double sxx = d.ux, syy = d.vy, szz = d.wz;
  const double sxy = (d.uy + d.vx) / 2.0, sxz = (d.uz + d.wx) / 2.0, syz = (d.vz + d.wy) / 2.0;
  return 2.0 * (sxx * sxx + syy * syy + szz * szz) + 4.0 * (sxy * sxy + sxz * sxz + syz * syz);
}

__device__ inline double strain_q_dot(const Grad9 &d, const Grad9 &t) {
  const double sxy = (d.uy + d.vx) / 2.0, sxz = (d.uz + d.wx) / 2.0, syz = (d.vz + d.wy) / 2.0;
  const double txy = (t.uy + t.vx) / 2.0, txz = (t.uz + t.wx) / 2.0, tyz = (t.vz + t.wy) / 2.0;
  return 4.0 * (d.ux * t.ux + d.vy * t.vy + d.wz * t.wz) + 8.0 * (sxy * txy + sxz * txz + syz * tyz);
}

// tangent of k_les: eddy' = nut Q' / (2 sqrt Q) (0 where Q = 0)
__global__ __launch_bounds__(NSX_BLOCK) void k_tlm_les(Geo g, const double *u, const double *v, const double *w,
                                                        const double *uv, const double *vv, const double *wv,
                                                        const double *du, const double *dv, const double *dw,
                                                        const double *duv, const double *dvv, const double *dwv,
                                                        double *deddy) {
  FOR_EACH_CELL(g, c) {
    int i, k;
    cell_of(g, c, i, k);
    const size_t s = (size_t)i * g.NS + k;
    double U[5], V[5], W[5], Ut[5], Vt[5], Wt[5];
    faces5(g, i, k, s, u, uv, U);
    faces5(g, i, k, s, v, vv, V);
    faces5(g, i, k, s, w, wv, W);
    faces5(g, i, k, s, du, duv, Ut);
    faces5(g, i, k, s, dv, dvv, Vt);
    faces5(g, i, k, s, dw, dwv, Wt);
    const Grad9 d = grad9(g, i, k, U, V, W), t = grad9(g, i, k, Ut, Vt, Wt);
    const double q = strain_q(d);
    deddy[s] = q > 0.0 ? les_nut(g, i, k) * strain_q_dot(d, t) / (2.0 * sqrt(q)) : 0.0;
  }
}

// transpose of k_les: the eddy-viscosity bar eb into the bars of u, v, w and
// of their vertex values
__global__ __launch_bounds__(NSX_BLOCK) void k_adj_les(Geo g, const double *u, const double *v, const double *w,
                                                        const double *uv, const double *vv, const double *wv,
                                                        const double *eb, double *ub, double *vb, double *wb,
                                                        double *uvb, double *vvb, double *wvb) {
  FOR_EACH_CELL(g, c) {
    int i, k;
    cell_of(g, c, i, k);
    const size_t s = (size_t)i * g.NS + k;
    const double ebar = eb[s];
    if (ebar == 0.0) continue;
    double U[5], V[5], W[5];
    faces5(g, i, k, s, u, uv, U);
    faces5(g, i, k, s, v, vv, V);
    faces5(g, i, k, s, w, wv, W);
    const Grad9 d = grad9(g, i, k, U, V, W);
    const double q = strain_q(d);
    if (!(q > 0.0)) continue;
    const double qb = ebar * les_nut(g, i, k) / (2.0 * sqrt(q));
    // Q = 2 (ux^2 + vy^2 + wz^2) + 4 (sxy^2 + sxz^2 + syz^2), s.. = (a + b)/2
    const double sxy = (d.uy + d.vx) / 2.0, sxz = (d.uz + d.wx) / 2.0, syz = (d.vz + d.wy) / 2.0;
    Grad9 b;
    b.ux = qb * 4.0 * d.ux;
    b.vy = qb * 4.0 * d.vy;
    b.wz = qb * 4.0 * d.wz;
    b.uy = b.vx = qb * 4.0 * sxy;
    b.uz = b.wx = qb * 4.0 * sxz;
    b.vz = b.wy = qb * 4.0 * syz;
    const int e = i * 3;
    const double x1 = g.xnn[e], x2 = g.xnn[e + 1], x3 = g.xnn[e + 2];
    const double y1 = g.ynn[e], y2 = g.ynn[e + 1], y3 = g.ynn[e + 2];
    const double l1 = g.dlVV[e], l2 = g.dlVV[e + 1], l3 = g.dlVV[e + 2];
    const double a = g.area[i], dz = g.dsig[k];
    double Ub[5], Vb[5], Wb[5];
    Ub[0] = (b.ux * x1 * l1 + b.uy * y1 * l1) / a;
    Ub[1] = (b.ux * x2 * l2 + b.uy * y2 * l2) / a;
    Ub[2] = (b.ux * x3 * l3 + b.uy * y3 * l3) / a;
    Vb[0] = (b.vx * x1 * l1 + b.vy * y1 * l1) / a;
    Vb[1] = (b.vx * x2 * l3 + b.vy * y2 * l3) / a;
    Vb[2] = (b.vx * x3 * l2 + b.vy * y3 * l2) / a;
    Wb[0] = (b.wx * x1 * l1 + b.wy * y1 * l1) / a;
    Wb[1] = (b.wx * x2 * l2 + b.wy * y2 * l2) / a;
    Wb[2] = (b.wx * x3 * l3 + b.wy * y3 * l3) / a;
    Ub[3] = b.uz / dz;
    Ub[4] = -b.uz / dz;
    Vb[3] = b.vz / dz;
    Vb[4] = -b.vz / dz;
    Wb[3] = b.wz / dz;
    Wb[4] = -b.wz / dz;
    aadd(ub + s, faces5_adj(g, i, k, s, Ub, ub, uvb));
    aadd(vb + s, faces5_adj(g, i, k, s, Vb, vb, vvb));
    aadd(wb + s, faces5_adj(g, i, k, s, Wb, wb, wvb));
  }
}

// ------------------------------------------- momentum flux + predictor ----
struct MomTan {
  // base state
  const double *uvw[3], *vtx[3], *face[5], *eddy;
  // tangents (dface, deddy, dforce may be null: zero)
  const double *duvw[3], *dvtx[3], *dface[5], *deddy, *dforce;
  double *dustar[3];
};

// Kim-Choi diffusion (diffusion3D.F90:110-191) and the five face values of
// phi at (i,k): src = sum dn_j dlVV_j sv + (dT + dB) area
__device__ inline double diff_src(const Geo &g, int i, int k, size_t s, const double *phi, const double *phiv) {
  const int e = i * 3;
  const double sv = g.dsigv[k - 1], pi = phi[s];
  double dn[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int ej = e + j;
    const int nb = g.nbr[ej];
    const int jv1 = g.vtx[ej], jv2 = g.vtx[e + (j + 1) % 3];
    const double diff1 = (phi[(size_t)nb * g.NS + k] - pi) / g.dlCC[ej];
    const double diff2 = (phiv[(size_t)jv2 * g.NL + k] - phiv[(size_t)jv1 * g.NL + k]) / g.dlVV[ej];
    dn[j] = diff1 * g.dle1[ej] + diff2 * g.dle2[ej];
  }
  const double dT = (phi[s + 1] - pi) / g.dsig[k];
  const double dB = (phi[s - 1] - pi) / g.dsig[k - 1];
  return dn[0] * g.dlVV[e] * sv + dn[1] * g.dlVV[e + 1] * sv + dn[2] * g.dlVV[e + 2] * sv + (dT + dB) * g.area[i];
}

// transpose of diff_src: the bar sb of src into the cell bars (own returned)
// and the vertex bars
__device__ inline double diff_src_adj(const Geo &g, int i, int k, size_t s, double sb, double *bar, double *vbar) {
  const int e = i * 3;
  const double sv = g.dsigv[k - 1];
  double own = 0.0;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int ej = e + j;
    const int nb = g.nbr[ej];
    const int jv1 = g.vtx[ej], jv2 = g.vtx[e + (j + 1) % 3];
    const double dnb = sb * g.dlVV[ej] * sv;
    const double c1 = dnb * g.dle1[ej] / g.dlCC[ej];
    aadd(bar + (size_t)nb * g.NS + k, c1);
    own -= c1;
    const double c2 = dnb * g.dle2[ej] / g.dlVV[ej];
    aadd(vbar + (size_t)jv2 * g.NL + k, c2);
    aadd(vbar + (size_t)jv1 * g.NL + k, -c2);
  }
  const double ab = sb * g.area[i];
  const double cT = ab / g.dsig[k], cB = ab / g.dsig[k - 1];
  aadd(bar + s + 1, cT);
  aadd(bar + s - 1, cB);
  own -= cT + cB;
  return own;
}

// tangent of k_momentum (RK2 stages: c1 = 1, no AB2 history): u*' = u' +
// dt (f' + (E' src + (E + nu) src') / Vol - flux' / Vol), flux' = sum_j
// (fc'_j UF_j + fc_j UF'_j); the ghost layers of u*' as BCvelcenter3D fills u*
__global__ __launch_bounds__(NSX_BLOCK) void k_tlm_momentum(Geo g, MomTan a) {
  FOR_EACH_CELL(g, c) {
    int i, k;
    cell_of(g, c, i, k);
    const size_t s = (size_t)i * g.NS + k;
    const int e = i * 3;
    const double area = g.area[i], sv = g.dsigv[k - 1];
    const double Vol = area * sv;
    const double Gam = a.eddy[s] + g.nu, dGam = a.deddy ? a.deddy[s] : 0.0;
    double UF[5], dUF[5];
    UF[0] = a.face[0][s] * g.dlVV[e] * sv;
    UF[1] = a.face[1][s] * g.dlVV[e + 1] * sv;
    UF[2] = a.face[2][s] * g.dlVV[e + 2] * sv;
    UF[3] = a.face[3][s] * area;
    UF[4] = a.face[4][s] * area;
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      const double fj = a.dface[j] ? a.dface[j][s] : 0.0;
      dUF[j] = j < 3 ? fj * g.dlVV[e + j] * sv : fj * area;
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double src = diff_src(g, i, k, s, a.uvw[q], a.vtx[q]);
      const double dsrc = diff_src(g, i, k, s, a.duvw[q], a.dvtx[q]);
      double fc[5], dfc[5];
      faces5(g, i, k, s, a.uvw[q], a.vtx[q], fc);
      faces5(g, i, k, s, a.duvw[q], a.dvtx[q], dfc);
      double dflux = 0.0;
#pragma unroll
      for (int j = 0; j < 5; ++j) dflux += dfc[j] * UF[j] + fc[j] * dUF[j];
      double dr = (q == 0 && a.dforce) ? a.dforce[s] : 0.0;
      dr += (dGam * src + Gam * dsrc) / Vol;
      dr -= dflux / Vol;
      const double dus = a.duvw[q][s] + g.dt * dr;
      a.dustar[q][s] = dus;
      if (k == 1) {
        const int sg = vel_ghost_sign(g.bcbot, g.BotBC, q);
        if (sg) a.dustar[q][s - 1] = sg > 0 ? dus : -dus;
      }
      if (k == g.NZ - 2) {
        const int sg = vel_ghost_sign(g.bctop, g.TopBC, q);
        if (sg) a.dustar[q][s + 1] = sg > 0 ? dus : -dus;
      }
    }
  }
}

struct MomAdj {
  const double *uvw[3], *vtx[3], *face[5], *eddy;
  const double *usb[3];  // bars of u* (interior; ghost bars already folded)
  double *ub[3], *vb[3], *faceb[5], *eddyb, *forceb;
};

// transpose of k_momentum's RK2 stage form
__global__ __launch_bounds__(NSX_BLOCK) void k_adj_momentum(Geo g, MomAdj a) {
  FOR_EACH_CELL(g, c) {
    int i, k;
    cell_of(g, c, i, k);
    const size_t s = (size_t)i * g.NS + k;
    const int e = i * 3;
    const double area = g.area[i], sv = g.dsigv[k - 1];
    const double Vol = area * sv;
    const double Gam = a.eddy[s] + g.nu;
    double UF[5];
    UF[0] = a.face[0][s] * g.dlVV[e] * sv;
    UF[1] = a.face[1][s] * g.dlVV[e + 1] * sv;
    UF[2] = a.face[2][s] * g.dlVV[e + 2] * sv;
    UF[3] = a.face[3][s] * area;
    UF[4] = a.face[4][s] * area;
    double UFb[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, Gb = 0.0;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double usb = a.usb[q][s];
      if (usb == 0.0) continue;
      const double rb = g.dt * usb;
      if (q == 0 && a.forceb) a.forceb[s] += rb;
      const double src = diff_src(g, i, k, s, a.uvw[q], a.vtx[q]);
      Gb += rb * src / Vol;
      double own = usb;
      own += diff_src_adj(g, i, k, s, rb * Gam / Vol, a.ub[q], a.vb[q]);
      double fc[5], fcb[5];
      faces5(g, i, k, s, a.uvw[q], a.vtx[q], fc);
      const double fluxb = -rb / Vol;
#pragma unroll
      for (int j = 0; j < 5; ++j) {
        fcb[j] = fluxb * UF[j];
        UFb[j] += fluxb * fc[j];
      }
      own += faces5_adj(g, i, k, s, fcb, a.ub[q], a.vb[q]);
      aadd(a.ub[q] + s, own);
    }
    a.eddyb[s] += Gb;
#pragma unroll
    for (int j = 0; j < 5; ++j) a.faceb[j][s] += j < 3 ? UFb[j] * g.dlVV[e + j] * sv : UFb[j] * area;
  }
}

// ----------------------------------------------------------- PPE source ----
// transpose of k_source (faces of u*, rhsp = div / dt - lagged Kim-Choi flux
// of pv): the face bars fb (interior layers) and the bar of rhsp into the
// bars of u*, v*, w* (ghost layers included) and of pv
__global__ __launch_bounds__(NSX_BLOCK) void k_adj_source(Geo g, const double *F1b, const double *F2b,
                                                           const double *F3b, const double *FTb, const double *FBb,
                                                           const double *rb, double *ub, double *vb, double *wb,
                                                           double *pvb) {
  FOR_EACH_CELL(g, c) {
    int i, k;
    cell_of(g, c, i, k);
    const size_t s = (size_t)i * g.NS + k;
    const int e = i * 3;
    const double sv = g.dsigv[k - 1], area = g.area[i];
    const double rhb = rb[s];
    const double resb = rhb / g.dt;
    const double Fb[3] = {F1b[s], F2b[s], F3b[s]};
    double uo = 0.0, vo = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double Ufb = Fb[j] + resb * g.dlVV[e + j] * sv;
      const double cu = 0.5 * Ufb * g.xnn[e + j], cv = 0.5 * Ufb * g.ynn[e + j];
      const size_t sn = (size_t)g.nbr[e + j] * g.NS + k;
      uo += cu;
      vo += cv;
      aadd(ub + sn, cu);
      aadd(vb + sn, cv);
      // lagged flux: rp -= dn_j dlVV_j sv, dn_j = (pv(jv2) - pv(jv1)) / dlVV_j dle2_j
      const int jv1 = g.vtx[e + j], jv2 = g.vtx[e + (j + 1) % 3];
      const double cp = -rhb * g.dlVV[e + j] * sv * g.dle2[e + j] / g.dlVV[e + j];
      aadd(pvb + (size_t)jv2 * g.NL + k, cp);
      aadd(pvb + (size_t)jv1 * g.NL + k, -cp);
    }
    aadd(ub + s, uo);
    aadd(vb + s, vo);
    const double UTb = FTb[s] + resb * area, UBb = FBb[s] + resb * area;
    aadd(wb + s, 0.5 * UTb - 0.5 * UBb);
    aadd(wb + s + 1, 0.5 * UTb);
    aadd(wb + s - 1, -0.5 * UBb);
  }
}

// ----------------------------------------------------------- corrector ----
// transpose of k_corrector (u = u* - dt grad_LSM p): the velocity bars (ghost
// bars folded) into the bars of u* and of p
__global__ __launch_bounds__(NSX_BLOCK) void k_adj_corrector(Geo g, const double *ub, const double *vb,
                                                              const double *wb, double *usb, double *vsb,
                                                              double *wsb, double *pb) {
  FOR_EACH_CELL(g, c) {
    int i, k;
    cell_of(g, c, i, k);
    const size_t s = (size_t)i * g.NS + k;
    const double bu = ub[s], bv = vb[s], bw = wb[s];
    usb[s] += bu;
    vsb[s] += bv;
    wsb[s] += bw;
    const double dxb = -g.dt * bu, dyb = -g.dt * bv, dsb = -g.dt * bw;
    const double sx2 = g.sum_xc2[i], sy2 = g.sum_yc2[i], sxy = g.sum_xcyc[i];
    const double deter = sx2 * sy2 - sxy * sxy;
    const double sfxb = (dxb * sy2 - dyb * sxy) / deter, sfyb = (dyb * sx2 - dxb * sxy) / deter;
    const int e = i * 3;
    double own = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double db = sfxb * g.dxCC[e + j] + sfyb * g.dyCC[e + j];
      aadd(pb + (size_t)g.nbr[e + j] * g.NS + k, db);
      own -= db;
    }
    const double dzp = g.sig[k + 1] - g.sig[k];
    const double dzm = g.sig[k - 1] - g.sig[k];
    const double cs = dsb / (dzp * dzp + dzm * dzm);
    aadd(pb + s + 1, cs * dzp);
    aadd(pb + s - 1, cs * dzm);
    own -= cs * (dzp + dzm);
    aadd(pb + s, own);
  }
}

// --------------------------------------------------------- face update ----
// transpose of k_newu (F -= dt dp/dn): the face bars pass through unchanged;
// their contribution to the bars of p and of its vertex values
__global__ __launch_bounds__(NSX_BLOCK) void k_adj_newu(Geo g, const double *F1b, const double *F2b,
                                                         const double *F3b, const double *FTb, const double *FBb,
                                                         double *pb, double *pvb) {
  FOR_EACH_CELL(g, c) {
    int i, k;
    cell_of(g, c, i, k);
    const size_t s = (size_t)i * g.NS + k;
    const int e = i * 3;
    const double Fb[3] = {F1b[s], F2b[s], F3b[s]};
    double own = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double dnb = -g.dt * Fb[j];
      const int jv1 = g.vtx[e + j], jv2 = g.vtx[e + (j + 1) % 3];
      const double c1 = dnb * g.dle1[e + j] / g.dlCC[e + j];
      aadd(pb + (size_t)g.nbr[e + j] * g.NS + k, c1);
      own -= c1;
      const double c2 = dnb * g.dle2[e + j] / g.dlVV[e + j];
      aadd(pvb + (size_t)jv2 * g.NL + k, c2);
      aadd(pvb + (size_t)jv1 * g.NL + k, -c2);
    }
    // the bottom face over dsig(k), as newu3D.F90:164 divides
    const double cT = -g.dt * FTb[s] / g.dsig[k], cB = -g.dt * FBb[s] / g.dsig[k];
    aadd(pb + s + 1, cT);
    aadd(pb + s - 1, cB);
    own -= cT + cB;
    aadd(pb + s, own);
  }
}

// ------------------------------------------------------- vertex values ----
// transpose of k_vertex (tagdir 2) in two passes: first the bars of the
// per-vertex inverse-distance values val(v) from the bars of the vertex
// outputs (a periodic vertex is the dlVsum-weighted mean of its pair or of
// the four corners; a Dirichlet vertex is 0), then those into the cells of
// each vertex's stencil
__global__ __launch_bounds__(NSX_BLOCK) void k_adj_vertex_val(Geo g, const double *fvb, const signed char *act,
                                                               double *valb) {
  const uint32_t t = blockIdx.x * NSX_BLOCK + threadIdx.x;
  if (t >= (uint32_t)g.NV * g.NL) return;
  const int v = (int)fdiv(t, g.divNL);
  const int l = (int)(t - (uint32_t)v * g.NL);
  const double b = fvb[(size_t)v * g.NL + l];
  if (b == 0.0) return;
  if (act[v] == VA_PERIODIC) {
    const signed char kind = g.vkind[v];
    if (kind == VK_CORNER) {
      double sd = 0.0;
      for (int q4 = 0; q4 < g.ncorner; ++q4) sd = sd + g.dlVsum[g.corner[q4]];
      for (int q4 = 0; q4 < g.ncorner; ++q4) {
        const int cv = g.corner[q4];
        aadd(valb + (size_t)cv * g.NL + l, b * g.dlVsum[cv] / sd);
      }
    } else if (kind == VK_PAIR) {
      const int p = g.vpair[v];
      const double sd = g.dlVsum[v] + g.dlVsum[p];
      aadd(valb + (size_t)v * g.NL + l, b * g.dlVsum[v] / sd);
      aadd(valb + (size_t)p * g.NL + l, b * g.dlVsum[p] / sd);
    }
  } else {
    aadd(valb + (size_t)v * g.NL + l, b);
  }
}

__global__ __launch_bounds__(NSX_BLOCK) void k_adj_vertex_idw(Geo g, const double *valb, const signed char *act,
                                                               double *fb) {
  const uint32_t t = blockIdx.x * NSX_BLOCK + threadIdx.x;
  if (t >= (uint32_t)g.NV * g.NL) return;
  const int v = (int)fdiv(t, g.divNL);
  const int l = (int)(t - (uint32_t)v * g.NL);
  if (act[v] == VA_ZERO || act[v] == VA_INFLOW) return;
  const double b = valb[(size_t)v * g.NL + l];
  if (b == 0.0) return;
  const double r = b / g.dlVsum[v];
  for (int j = g.vptr[v]; j < g.vptr[v + 1]; ++j) aadd(fb + (size_t)g.vcell[j] * g.NS + l, g.vw[j] * r);
}

// ------------------------------------------------------- ghost layers ----
// transpose of k_ghost_layers: a ghost layer filled from the first or last
// interior layer (sign +-1) passes its bar to that layer and is cleared; a
// ghost layer the forward step leaves alone keeps no bar either (its value
// does not depend on the interior)
__global__ __launch_bounds__(NSX_BLOCK) void k_adj_ghost(Geo g, double *fb, int kind) {
  const int i = blockIdx.x * NSX_BLOCK + threadIdx.x;
  if (i >= g.N0) return;
  const size_t b = (size_t)i * g.NS;
  int sb, st;
  if (kind == 0) {
    sb = g.bcbot == 1 ? 1 : 0;
    st = g.bctop == 1 ? 1 : 0;
  } else {
    sb = vel_ghost_sign(g.bcbot, g.BotBC, kind - 1);
    st = vel_ghost_sign(g.bctop, g.TopBC, kind - 1);
  }
  if (sb) fb[b + 1] += sb > 0 ? fb[b] : -fb[b];
  fb[b] = 0.0;
  if (st) fb[b + g.NZ - 2] += st > 0 ? fb[b + g.NZ - 1] : -fb[b + g.NZ - 1];
  fb[b + g.NZ - 1] = 0.0;
}

// transpose of k_wall_ghosts: the bar of a wall ghost column passes to its
// cell's column, layer by layer (negated for flag 0), and is cleared; two
// ghosts of one cell (a corner of a walled box) add into it atomically
__global__ __launch_bounds__(NSX_BLOCK) void k_adj_wall_ghosts(Geo g, double *fb, int kind) {
  const uint32_t t = blockIdx.x * NSX_BLOCK + threadIdx.x;
  if (t >= (uint32_t)g.NGW * (uint32_t)g.NZ) return;
  const int w = (int)(t / (uint32_t)g.NZ), k = (int)(t - (uint32_t)w * (uint32_t)g.NZ);
  const size_t d = (size_t)(g.N0 + w) * g.NS + k;
  const double b = fb[d];
  fb[d] = 0.0;
  if (b != 0.0) aadd(fb + (size_t)g.gw_owner[w] * g.NS + k, ((g.gw_neu[w] >> kind) & 1) ? b : -b);
}

// --------------------------------------------------------- objective ----
// the bottom-wall shear of the channel, sum_i w_i u(i, 1) with w_i = nu (1 -
// s) area_i / (dsig(1) sum area) (s the sign of the bottom ghost layer of u:
// -1 no slip), times `scale`: partial sums per block
__global__ __launch_bounds__(NSX_BLOCK) void k_wall_shear(Geo g, const double *u, double scale, double *P) {
  double acc[1] = {0.0};
  const int sgn = vel_ghost_sign(g.bcbot, g.BotBC, 0);
  for (uint32_t i = blockIdx.x * NSX_BLOCK + threadIdx.x; i < (uint32_t)g.N0; i += gridDim.x * NSX_BLOCK)
    acc[0] += g.area[i] * u[(size_t)i * g.NS + 1];
  const double wfac = scale * g.nu * (1.0 - (double)sgn) / g.dsig[0];
  acc[0] *= wfac;
  write_partials<1>(acc, P);
}

// J += sum of the partials (one block, fixed order)
__global__ __launch_bounds__(NSX_BLOCK) void k_accumulate(const double *P, uint32_t G, double *J) {
  double t[1];
  reduce_partials<1>(P, G, t);
  if (threadIdx.x == 0) J[0] += t[0];
}

// seed of the adjoint: u_bar(i, 1) += scale w_i (area normalised by atot)
__global__ __launch_bounds__(NSX_BLOCK) void k_wall_shear_seed(Geo g, double *ub, double scale) {
  const int i = blockIdx.x * NSX_BLOCK + threadIdx.x;
  if (i >= g.N0) return;
  const int sgn = vel_ghost_sign(g.bcbot, g.BotBC, 0);
  ub[(size_t)i * g.NS + 1] += scale * g.nu * (1.0 - (double)sgn) / g.dsig[0] * g.area[i];
}

// y = a y + b x over n doubles
__global__ __launch_bounds__(NSX_BLOCK) void k_axpby(size_t n, double a, double *y, double b, const double *x) {
  for (size_t t = (size_t)blockIdx.x * NSX_BLOCK + threadIdx.x; t < n; t += (size_t)gridDim.x * NSX_BLOCK)
    y[t] = a * y[t] + (x ? b * x[t] : 0.0);
}

}  // namespace

// ------------------------------------------------------------ launchers ---
static dim3 cells_grid(const Geo &g) { return dim3(nsx_grid_for(g.ncell)); }

int launch_tlm_les(const Geo &g, hipStream_t s, const double *const uvw[3], const double *const vtx[3],
                   const double *const duvw[3], const double *const dvtx[3], double *deddy) {
  hipLaunchKernelGGL(k_tlm_les, cells_grid(g), dim3(NSX_BLOCK), 0, s, g, uvw[0], uvw[1], uvw[2], vtx[0], vtx[1],
                     vtx[2], duvw[0], duvw[1], duvw[2], dvtx[0], dvtx[1], dvtx[2], deddy);
  return (int)hipGetLastError();
}

int launch_adj_les(const Geo &g, hipStream_t s, const double *const uvw[3], const double *const vtx[3],
                   const double *eddyb, double *const uvwb[3], double *const vtxb[3]) {
  hipLaunchKernelGGL(k_adj_les, cells_grid(g), dim3(NSX_BLOCK), 0, s, g, uvw[0], uvw[1], uvw[2], vtx[0], vtx[1],
                     vtx[2], eddyb, uvwb[0], uvwb[1], uvwb[2], vtxb[0], vtxb[1], vtxb[2]);
  return (int)hipGetLastError();
}

int launch_tlm_momentum(const Geo &g, hipStream_t s, const double *const uvw[3], const double *const vtx[3],
                        const double *const face[5], const double *eddy, const double *const duvw[3],
                        const double *const dvtx[3], const double *const dface[5], const double *deddy,
                        const double *dforce, double *const dustar[3]) {
  MomTan a;
  for (int q = 0; q < 3; ++q) {
    a.uvw[q] = uvw[q];
    a.vtx[q] = vtx[q];
    a.duvw[q] = duvw[q];
    a.dvtx[q] = dvtx[q];
    a.dustar[q] = dustar[q];
  }
  for (int j = 0; j < 5; ++j) {
    a.face[j] = face[j];
    a.dface[j] = dface ? dface[j] : nullptr;
  }
  a.eddy = eddy;
  a.deddy = deddy;
  a.dforce = dforce;
  hipLaunchKernelGGL(k_tlm_momentum, cells_grid(g), dim3(NSX_BLOCK), 0, s, g, a);
  return (int)hipGetLastError();
}

int launch_adj_momentum(const Geo &g, hipStream_t s, const double *const uvw[3], const double *const vtx[3],
                        const double *const face[5], const double *eddy, const double *const usb[3],
                        double *const uvwb[3], double *const vtxb[3], double *const faceb[5], double *eddyb,
                        double *forceb) {
  MomAdj a;
  for (int q = 0; q < 3; ++q) {
    a.uvw[q] = uvw[q];
    a.vtx[q] = vtx[q];
    a.usb[q] = usb[q];
    a.ub[q] = uvwb[q];
    a.vb[q] = vtxb[q];
  }
  for (int j = 0; j < 5; ++j) {
    a.face[j] = face[j];
    a.faceb[j] = faceb[j];
  }
  a.eddy = eddy;
  a.eddyb = eddyb;
  a.forceb = forceb;
  hipLaunchKernelGGL(k_adj_momentum, cells_grid(g), dim3(NSX_BLOCK), 0, s, g, a);
  return (int)hipGetLastError();
}

int launch_adj_source(const Geo &g, hipStream_t s, const double *const faceb[5], const double *rhspb,
                      double *const uvwb[3], double *pvb) {
  hipLaunchKernelGGL(k_adj_source, cells_grid(g), dim3(NSX_BLOCK), 0, s, g, faceb[0], faceb[1], faceb[2], faceb[3],
                     faceb[4], rhspb, uvwb[0], uvwb[1], uvwb[2], pvb);
  return (int)hipGetLastError();
}

int launch_adj_corrector(const Geo &g, hipStream_t s, const double *const ub[3], double *const usb[3], double *pb) {
  hipLaunchKernelGGL(k_adj_corrector, cells_grid(g), dim3(NSX_BLOCK), 0, s, g, ub[0], ub[1], ub[2], usb[0], usb[1],
                     usb[2], pb);
  return (int)hipGetLastError();
}

int launch_adj_newu(const Geo &g, hipStream_t s, const double *const faceb[5], double *pb, double *pvb) {
  hipLaunchKernelGGL(k_adj_newu, cells_grid(g), dim3(NSX_BLOCK), 0, s, g, faceb[0], faceb[1], faceb[2], faceb[3],
                     faceb[4], pb, pvb);
  return (int)hipGetLastError();
}

int launch_adj_vertex(const Geo &g, hipStream_t s, const double *fvb, const signed char *act, double *valb,
                      double *fb) {
  const uint32_t n = (uint32_t)g.NV * g.NL;
  const dim3 grid((n + NSX_BLOCK - 1) / NSX_BLOCK);
  hipError_t e = hipMemsetAsync(valb, 0, (size_t)n * sizeof(double), s);
  if (e != hipSuccess) return (int)e;
  hipLaunchKernelGGL(k_adj_vertex_val, grid, dim3(NSX_BLOCK), 0, s, g, fvb, act, valb);
  hipLaunchKernelGGL(k_adj_vertex_idw, grid, dim3(NSX_BLOCK), 0, s, g, (const double *)valb, act, fb);
  return (int)hipGetLastError();
}

int launch_adj_ghost(const Geo &g, hipStream_t s, double *fb, int kind) {
  // the forward fills the sigma ghost layers, then the wall ghost columns
  // (which copy them too): fold in the reverse order
  if (g.NGW) {
    const uint32_t n = (uint32_t)g.NGW * (uint32_t)g.NZ;
    hipLaunchKernelGGL(k_adj_wall_ghosts, dim3((n + NSX_BLOCK - 1) / NSX_BLOCK), dim3(NSX_BLOCK), 0, s, g, fb, kind);
  }
  hipLaunchKernelGGL(k_adj_ghost, dim3((g.N0 + NSX_BLOCK - 1) / NSX_BLOCK), dim3(NSX_BLOCK), 0, s, g, fb, kind);
  return (int)hipGetLastError();
}

int launch_wall_shear(const Geo &g, hipStream_t s, const double *u, double scale, double *part, double *J) {
  const int G = 64;
  hipLaunchKernelGGL(k_wall_shear, dim3(G), dim3(NSX_BLOCK), 0, s, g, u, scale, part);
  hipLaunchKernelGGL(k_accumulate, dim3(1), dim3(NSX_BLOCK), 0, s, (const double *)part, (uint32_t)G, J);
  return (int)hipGetLastError();
}

int launch_wall_shear_seed(const Geo &g, hipStream_t s, double *ub, double scale) {
  hipLaunchKernelGGL(k_wall_shear_seed, dim3((g.N0 + NSX_BLOCK - 1) / NSX_BLOCK), dim3(NSX_BLOCK), 0, s, g, ub,
                     scale);
  return (int)hipGetLastError();
}

int launch_axpby(hipStream_t s, size_t n, double a, double *y, double b, const double *x) {
  size_t blocks = (n + NSX_BLOCK - 1) / NSX_BLOCK;
  if (blocks > 8192) blocks = 8192;
  if (blocks == 0) return 0;
  hipLaunchKernelGGL(k_axpby, dim3((unsigned)blocks), dim3(NSX_BLOCK), 0, s, n, a, y, b, x);
  return (int)hipGetLastError();
}
