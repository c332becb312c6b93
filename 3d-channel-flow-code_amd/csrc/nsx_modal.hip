// nsx_modal.hip -- host set-up of the vertical-mode preconditioner of the
// precision form's CG (NSX_REFINE_MODAL).
//
// The PPE operator of the prismatic sigma grid is a Kronecker sum: a
// horizontal coupling between columns i and j at layer k is ppec(i,j)
// dsigv(k-1), a vertical one between layers k and k+1 of column i is area(i)
// t(k) (t = 1/dsig, zero at the Neumann top and bottom), and the diagonal
// closes every row to zero (PDE_center, PDE.F90:8-135).  So -A = L (x) S +
// D (x) T with L the weighted graph Laplacian of the columns, D = diag(area),
// S = diag(dsigv) and T the layers' 1D Laplacian.  The generalized
// eigenvectors Phi of (T, S) (Phi^T S Phi = I, Phi^T T Phi = diag(mu)) split
// it into one 2D problem per vertical mode m, L + mu_m D.  The preconditioner
// transforms a residual to modal coefficients, solves the lowest m0 2D
// problems with a V-cycle of aggregation multigrid on the column graph
// (Jacobi smoothing, piecewise-constant prolongation, Galerkin coarse
// operators, a dense inverse on the coarsest level), scales the other modes
// by their diagonal, and transforms back.  Every piece is symmetric, so the
// preconditioner is a fixed symmetric positive semi-definite operator and CG
// stays CG.  This file builds its tables on the host; the kernels are in
// nsx_bicgstab.hip.
#include <algorithm>
#include <cmath>
#include <string>
#include <unordered_map>
#include <vector>

#include "nsx_modal.h"

namespace {

// eigenvalues and orthonormal eigenvectors of the symmetric n x n matrix a
// (row-major) by cyclic Jacobi rotations; v[k*n+m] is component k of vector m
void jacobi_eigen(int n, std::vector<double> a, std::vector<double> &w, std::vector<double> &v) {
  v.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) v[(size_t)i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int p = 0; p < n; ++p)
      for (int q = 0; q < n; ++q) {
        const double x = a[(size_t)p * n + q] * a[(size_t)p * n + q];
        tot += x;
        if (p != q) off += x;
      }
    if (off <= 1e-30 * tot) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = a[(size_t)p * n + q];
        if (std::fabs(apq) < 1e-300) continue;
        const double app = a[(size_t)p * n + p], aqq = a[(size_t)q * n + q];
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        // A <- J^T A J on rows/columns p, q
        for (int k = 0; k < n; ++k) {
          const double akp = a[(size_t)k * n + p], akq = a[(size_t)k * n + q];
          a[(size_t)k * n + p] = c * akp - s * akq;
          a[(size_t)k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          const double apk = a[(size_t)p * n + k], aqk = a[(size_t)q * n + k];
          a[(size_t)p * n + k] = c * apk - s * aqk;
          a[(size_t)q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double vkp = v[(size_t)k * n + p], vkq = v[(size_t)k * n + q];
          v[(size_t)k * n + p] = c * vkp - s * vkq;
          v[(size_t)k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  w.resize(n);
  for (int i = 0; i < n; ++i) w[i] = a[(size_t)i * n + i];
}

// inverse of the n x n matrix a (row-major) by Gauss-Jordan elimination with
// partial pivoting; false when it is singular to working precision
bool invert(int n, std::vector<double> a, std::vector<double> &inv) {
  inv.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) inv[(size_t)i * n + i] = 1.0;
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(a[(size_t)r * n + c]) > std::fabs(a[(size_t)piv * n + c])) piv = r;
    if (!(std::fabs(a[(size_t)piv * n + c]) > 0.0)) return false;
    if (piv != c)
      for (int k = 0; k < n; ++k) {
        std::swap(a[(size_t)c * n + k], a[(size_t)piv * n + k]);
        std::swap(inv[(size_t)c * n + k], inv[(size_t)piv * n + k]);
      }
    const double d = 1.0 / a[(size_t)c * n + c];
    for (int k = 0; k < n; ++k) {
      a[(size_t)c * n + k] *= d;
      inv[(size_t)c * n + k] *= d;
    }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const double f = a[(size_t)r * n + c];
      if (f == 0.0) continue;
      for (int k = 0; k < n; ++k) {
        a[(size_t)r * n + k] -= f * a[(size_t)c * n + k];
        inv[(size_t)r * n + k] -= f * inv[(size_t)c * n + k];
      }
    }
  }
  return true;
}

// a weighted graph in CSR form (off-diagonal weights; the Laplacian's
// diagonal is the row sum)
struct Graph {
  int n = 0;
  std::vector<int> ptr, col;
  std::vector<double> w, area;
};

Graph from_pairs(int n, std::vector<std::unordered_map<int, double>> &rows, const std::vector<double> &area) {
  Graph G;
  G.n = n;
  G.area = area;
  G.ptr.assign(n + 1, 0);
  for (int i = 0; i < n; ++i) {
    std::vector<std::pair<int, double>> e(rows[i].begin(), rows[i].end());
    std::sort(e.begin(), e.end());
    for (auto &p : e) {
      G.col.push_back(p.first);
      G.w.push_back(p.second);
    }
    G.ptr[i + 1] = (int)G.col.size();
  }
  return G;
}

// greedy aggregation: each node not yet taken seeds an aggregate with its
// neighbours not yet taken
std::vector<int> aggregate(const Graph &G, int &na) {
  std::vector<int> agg(G.n, -1);
  na = 0;
  for (int i = 0; i < G.n; ++i) {
    if (agg[i] >= 0) continue;
    agg[i] = na;
    for (int q = G.ptr[i]; q < G.ptr[i + 1]; ++q)
      if (agg[G.col[q]] < 0) agg[G.col[q]] = na;
    ++na;
  }
  return agg;
}

Graph galerkin(const Graph &G, const std::vector<int> &agg, int na) {
  std::vector<std::unordered_map<int, double>> rows(na);
  std::vector<double> area(na, 0.0);
  for (int i = 0; i < G.n; ++i) {
    area[agg[i]] += G.area[i];
    for (int q = G.ptr[i]; q < G.ptr[i + 1]; ++q) {
      const int I = agg[i], J = agg[G.col[q]];
      if (I != J) rows[I][J] += G.w[q];
    }
  }
  return from_pairs(na, rows, area);
}

}  // namespace

int modal_build_host(int N0, int nzi, const int *nbr, const double *ppec, const double *area, const double *tlay,
                     const double *svl, int m0, ModalHost &H, std::string &err) {
  if (nzi < 1 || N0 < 2) {
    err = "modal preconditioner: empty mesh";
    return -1;
  }
  if (m0 > nzi) m0 = nzi;
  H.N0 = N0;
  H.nzi = nzi;
  H.m0 = m0;
  // vertical modes: C = S^-1/2 T S^-1/2, Phi = S^-1/2 Y
  std::vector<double> C((size_t)nzi * nzi, 0.0);
  for (int q = 0; q + 1 < nzi; ++q) {
    const double t = tlay[q + 1];
    C[(size_t)q * nzi + q] += t;
    C[(size_t)(q + 1) * nzi + q + 1] += t;
    C[(size_t)q * nzi + q + 1] -= t;
    C[(size_t)(q + 1) * nzi + q] -= t;
  }
  C[0] += tlay[0];
  C[(size_t)nzi * nzi - 1] += tlay[nzi];
  std::vector<double> rs(nzi);
  for (int q = 0; q < nzi; ++q) {
    if (!(svl[q] > 0.0)) {
      err = "modal preconditioner: a layer with dsigv <= 0";
      return -1;
    }
    rs[q] = 1.0 / std::sqrt(svl[q]);
  }
  for (int p = 0; p < nzi; ++p)
    for (int q = 0; q < nzi; ++q) C[(size_t)p * nzi + q] *= rs[p] * rs[q];
  std::vector<double> mu, Y;
  jacobi_eigen(nzi, C, mu, Y);
  std::vector<int> order(nzi);
  for (int m = 0; m < nzi; ++m) order[m] = m;
  std::sort(order.begin(), order.end(), [&](int a, int b) { return mu[a] < mu[b]; });
  H.mu.resize(nzi);
  H.phi.assign((size_t)nzi * nzi, 0.0f);
  for (int m = 0; m < nzi; ++m) {
    const int src = order[m];
    H.mu[m] = (float)std::max(mu[src], 0.0);
    for (int k = 0; k < nzi; ++k) H.phi[(size_t)k * nzi + m] = (float)(rs[k] * Y[(size_t)k * nzi + src]);
  }
  // level 0: the columns, edges to interior neighbours, weights made
  // symmetric (the average of the two directions)
  std::vector<std::unordered_map<int, double>> rows(N0);
  for (int i = 0; i < N0; ++i)
    for (int j = 0; j < 3; ++j) {
      const int n = nbr[i * 3 + j];
      if (n < 0 || n >= N0 || n == i) continue;
      const double w = 0.5 * ppec[i * 3 + j];
      rows[i][n] += w;
      rows[n][i] += w;
    }
  std::vector<Graph> lev;
  lev.push_back(from_pairs(N0, rows, std::vector<double>(area, area + N0)));
  std::vector<std::vector<int>> aggs;
  while (lev.back().n > 400 && (int)lev.size() < NSX_MODAL_MAXLEV) {
    int na = 0;
    std::vector<int> agg = aggregate(lev.back(), na);
    if (na > 0.9 * lev.back().n) break;
    Graph c = galerkin(lev.back(), agg, na);
    aggs.push_back(agg);
    lev.push_back(std::move(c));
  }
  const int L = (int)lev.size();
  H.levels.resize(L);
  for (int l = 0; l < L; ++l) {
    ModalHost::Level &o = H.levels[l];
    const Graph &G = lev[l];
    o.n = G.n;
    o.ptr = G.ptr;
    o.col = G.col;
    o.w.assign(G.w.begin(), G.w.end());
    o.diag.resize(G.n);
    o.area.resize(G.n);
    for (int i = 0; i < G.n; ++i) {
      double d = 0.0;
      for (int q = G.ptr[i]; q < G.ptr[i + 1]; ++q) d += G.w[q];
      o.diag[i] = (float)d;
      o.area[i] = (float)G.area[i];
    }
    if (l + 1 < L) {
      o.agg = aggs[l];
      o.eagg.resize(G.col.size());
      for (size_t e = 0; e < G.col.size(); ++e) o.eagg[e] = aggs[l][G.col[e]];
      // children of each coarse node, for the restriction
      ModalHost::Level &c = H.levels[l + 1];
      const int nc = lev[l + 1].n;
      std::vector<std::vector<int>> kids(nc);
      for (int i = 0; i < G.n; ++i) kids[aggs[l][i]].push_back(i);
      c.dptr.assign(nc + 1, 0);
      std::vector<std::pair<int, std::pair<double, double>>> row;
      for (int J = 0; J < nc; ++J) {
        row.clear();
        for (int i : kids[J]) {
          row.push_back({i, {1.0 - (double)NSX_MODAL_OMEGA, 0.0}});
          for (int e = G.ptr[i]; e < G.ptr[i + 1]; ++e)
            row.push_back({G.col[e], {0.0, (double)NSX_MODAL_OMEGA * G.w[e]}});
        }
        std::stable_sort(row.begin(), row.end(),
                         [](const std::pair<int, std::pair<double, double>> &a,
                            const std::pair<int, std::pair<double, double>> &b) { return a.first < b.first; });
        for (size_t q = 0; q < row.size();) {
          double vb = 0.0, vz = 0.0;
          size_t r = q;
          for (; r < row.size() && row[r].first == row[q].first; ++r) {
            vb += row[r].second.first;
            vz += row[r].second.second;
          }
          c.dcol.push_back(row[q].first);
          c.db.push_back((float)vb);
          c.dz.push_back((float)vz);
          q = r;
        }
        c.dptr[J + 1] = (int)c.dcol.size();
      }
    }
  }
  // the coarsest level's inverses, one per low mode; the Neumann mode
  // (mu = 0, L singular with the constant null vector) by the inverse with
  // node 0 pinned, between projections onto zero mean
  const Graph &G = lev.back();
  const int nc = G.n;
  H.nc = nc;
  H.cinv.assign((size_t)m0 * nc * nc, 0.0f);
  for (int m = 0; m < m0; ++m) {
    std::vector<double> A((size_t)nc * nc, 0.0);
    for (int i = 0; i < nc; ++i) {
      double d = 0.0;
      for (int q = G.ptr[i]; q < G.ptr[i + 1]; ++q) {
        A[(size_t)i * nc + G.col[q]] -= G.w[q];
        d += G.w[q];
      }
      A[(size_t)i * nc + i] += d + (double)H.mu[m] * G.area[i];
    }
    std::vector<double> inv;
    const bool neumann = m == 0 && H.mu[0] <= 1e-9f * H.mu[nzi - 1];
    if (!neumann) {
      if (!invert(nc, A, inv)) {
        err = "modal preconditioner: singular coarse operator";
        return -1;
      }
    } else {
      std::vector<double> B((size_t)(nc - 1) * (nc - 1));
      for (int i = 1; i < nc; ++i)
        for (int j = 1; j < nc; ++j) B[(size_t)(i - 1) * (nc - 1) + (j - 1)] = A[(size_t)i * nc + j];
      std::vector<double> Bi;
      if (!invert(nc - 1, B, Bi)) {
        err = "modal preconditioner: singular coarse operator";
        return -1;
      }
      std::vector<double> Gm((size_t)nc * nc, 0.0);
      for (int i = 1; i < nc; ++i)
        for (int j = 1; j < nc; ++j) Gm[(size_t)i * nc + j] = Bi[(size_t)(i - 1) * (nc - 1) + (j - 1)];
      // inv = P Gm P, P = I - 1 1^T / nc
      std::vector<double> rowm(nc, 0.0), colm(nc, 0.0);
      double all = 0.0;
      for (int i = 0; i < nc; ++i)
        for (int j = 0; j < nc; ++j) {
          rowm[i] += Gm[(size_t)i * nc + j] / nc;
          colm[j] += Gm[(size_t)i * nc + j] / nc;
          all += Gm[(size_t)i * nc + j] / ((double)nc * nc);
        }
      inv.assign((size_t)nc * nc, 0.0);
      for (int i = 0; i < nc; ++i)
        for (int j = 0; j < nc; ++j) inv[(size_t)i * nc + j] = Gm[(size_t)i * nc + j] - rowm[i] - colm[j] + all;
    }
    for (size_t q = 0; q < (size_t)nc * nc; ++q) H.cinv[(size_t)m * nc * nc + q] = (float)inv[q];
  }
  return 0;
}
