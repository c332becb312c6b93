"""Differentiation of the pressure solve: the adjoint of bicgstab with respect
to its right-hand side (nsx_bicgstab_adjoint, A^T through the neighbours'
lateral couplings) and a torch.autograd.Function over the C ABI.

CPU: the transposed couplings built by the reverse-edge rule are the exact
transpose of the PDE_center matrix.  GPU: the dot-product test <w, F d> =
<F^T w, d> (F linear), the adjoint against a dense solve of the transposed
system on the 6x4x4 mesh, and the autograd gradient against central finite
differences along seeded directions.  Tolerances: both solves stop at the
reference's 1e-10 residual of an operator with condition number O(1e4), so
the identities hold to 1e-6 relative (1e-5 for finite differences).
"""
import numpy as np
import pytest
import scipy.sparse as sp

from common import interior, tiny_case
from nsx.case import config_case


def ppe_matrix(case, trans=False):
    """The PPE operator on interior unknowns (i-major, k fastest), from the
    stored AmC* coefficients; with trans, the lateral couplings of row i are
    taken from the neighbour's row towards i (the rule of nsx_commit)."""
    A = case.arrays
    n0, nz = case.N_CELL0, case.NZ
    nzi = nz - 2
    ncp = A["No_cp"].astype(np.int64) - 1
    nb = np.zeros((n0, 3), np.int64)
    for j in range(3):
        jc = ncp[:n0, j].copy()
        gh = jc >= n0
        jc[gh] = ncp[jc[gh], 2]
        nb[:, j] = jc
    coef = [A["AmC1"], A["AmC2"], A["AmC3"]]
    rows, cols, vals = [], [], []
    idx = lambda i, k: i * nzi + (k - 1)
    for k in range(1, nz - 1):
        for i in range(n0):
            r = idx(i, k)
            rows.append(r), cols.append(r), vals.append(A["AmC0"][i, k])
            for j in range(3):
                n = nb[i, j]
                if trans:
                    cands = [jj for jj in range(3) if nb[n, jj] == i]
                    jj = min(cands, key=lambda q: abs(coef[q][n, k] - coef[j][i, k]))
                    c = coef[jj][n, k]
                else:
                    c = coef[j][i, k]
                rows.append(r), cols.append(idx(n, k)), vals.append(c)
            if k < nz - 2:
                rows.append(r), cols.append(idx(i, k + 1)), vals.append(A["AmCT"][i, k])
            if k > 1:
                rows.append(r), cols.append(idx(i, k - 1)), vals.append(A["AmCB"][i, k])
    N = n0 * nzi
    return sp.csr_matrix((vals, (rows, cols)), shape=(N, N))


@pytest.mark.parametrize("which", ["tiny", "cfg0"])
def test_transposed_couplings_are_the_exact_transpose(which):
    c = tiny_case()[0] if which == "tiny" else config_case(0, jitter=0.1)
    A, AT = ppe_matrix(c), ppe_matrix(c, trans=True)
    d = (AT - A.T.tocsr())
    d.eliminate_zeros()
    assert d.nnz == 0
    # and the operator is not symmetric bit for bit, so the rule matters
    s = (A - A.T.tocsr())
    s.eliminate_zeros()
    assert which == "tiny" or s.nnz > 0


def _fields(c, seed):
    rng = np.random.default_rng(seed)
    d, w = c.cell_field(), c.cell_field()
    d[:c.N_CELL0, 1:-1] = rng.standard_normal((c.N_CELL0, c.NZ - 2))
    w[:c.N_CELL0, 1:-1] = rng.standard_normal((c.N_CELL0, c.NZ - 2))
    return d, w


def _ctx(case):
    from nsx.device import Context
    return Context(case)


@pytest.mark.gpu
@pytest.mark.parametrize("which", ["tiny", "cfg0"])
def test_adjoint_dot_product(which):
    c = tiny_case()[0] if which == "tiny" else config_case(0, jitter=0.1)
    d, w = _fields(c, 3)
    G = _ctx(c)
    G.init_state()
    phi, _, info = G.bicgstab(c.cell_field(), d)
    lam, ainfo = G.bicgstab_adjoint(w)
    G.close()
    assert info["exit"] == "converged" and ainfo["exit"] == "converged"
    lhs = float((interior(c, w) * interior(c, phi)).sum())
    rhs = float((interior(c, lam) * interior(c, d)).sum())
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs), (lhs, rhs)


@pytest.mark.gpu
def test_adjoint_against_dense_transposed_solve():
    """6x4 quads x 4 layers: lam = N0 (A^T)^+ N1 w from a dense least-squares
    solve of the transposed system (N1 subtracts the volume-weighted mean,
    N0 the volume mean, as Glob_Norm(.,1) and Glob_Norm(.,0) do)."""
    c = tiny_case()[0]
    _, w = _fields(c, 4)
    G = _ctx(c)
    G.init_state()
    lam, info = G.bicgstab_adjoint(w)
    G.close()
    AT = ppe_matrix(c, trans=True).toarray()
    vol = (c.areaCell[:c.N_CELL0, None] * c.dsig[None, 1:-1]).ravel()
    Vt = c.TotalVolume * (c.params.sigFin - c.params.sigIni)
    g = interior(c, w).ravel()
    g = g - vol * g.sum() / Vt
    x = np.linalg.lstsq(AT, g, rcond=None)[0]
    x = x - (x * vol).sum() / Vt
    got = interior(c, lam).ravel()
    assert np.linalg.norm(got - x) <= 1e-8 * np.linalg.norm(x)


@pytest.mark.gpu
def test_autograd_function_against_finite_differences():
    torch = pytest.importorskip("torch")
    c = config_case(0, jitter=0.1)
    G = _ctx(c)
    G.init_state()
    shape = (c.N_CELL0, c.NZ - 2)

    class PressureSolve(torch.autograd.Function):
        """phi = bicgstab(rhs) on interior cells; backward through the adjoint solve."""

        @staticmethod
        def forward(ctx, rhs):
            full = c.cell_field()
            full[:c.N_CELL0, 1:-1] = rhs.detach().cpu().numpy()
            phi, _, info = G.bicgstab(c.cell_field(), full)
            assert info["exit"] == "converged"
            return torch.from_numpy(np.ascontiguousarray(interior(c, phi)))

        @staticmethod
        def backward(ctx, grad):
            full = c.cell_field()
            full[:c.N_CELL0, 1:-1] = grad.detach().cpu().numpy()
            lam, info = G.bicgstab_adjoint(full)
            assert info["exit"] == "converged"
            return torch.from_numpy(np.ascontiguousarray(interior(c, lam)))

    rng = np.random.default_rng(8)
    wt = torch.from_numpy(rng.standard_normal(shape))
    rhs = torch.from_numpy(rng.standard_normal(shape)).requires_grad_(True)
    loss = (wt * PressureSolve.apply(rhs)).sum()
    loss.backward()
    grad = rhs.grad.numpy()
    for q in range(4):
        dirn = rng.standard_normal(shape)
        # the map is linear, so the central difference is exact for any step;
        # a unit step keeps the solves' ~1e-6 relative error from being
        # amplified by 1/h
        h = 1.0
        with torch.no_grad():
            lp = float((wt * PressureSolve.apply(rhs + h * torch.from_numpy(dirn))).sum())
            lm = float((wt * PressureSolve.apply(rhs - h * torch.from_numpy(dirn))).sum())
        fd = (lp - lm) / (2 * h)
        ad = float((grad * dirn).sum())
        assert abs(fd - ad) <= 1e-5 * abs(ad), (q, fd, ad)
    G.close()
