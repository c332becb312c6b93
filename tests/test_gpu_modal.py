"""The precision form's CG preconditioned by the vertical modes of the
operator (NSX_REFINE_MODAL): per column the generalized eigenvectors of the
vertical couplings split A into one 2D problem per mode, the 16 lowest go
through a V-cycle of aggregation multigrid on the column graph, the others
are scaled by their diagonal.

  * random right-hand sides on the 32x16 quads x 32 layers channel and the
    walled duct: the solve meets the reference's stopping rule on its true
    residual (checked on the host with the stored PDE_center coefficients, 2
    eps for the summation order) and agrees with the FP64 BiCGSTAB and the
    oracle to 1e-6 (both stop at 1e-10 of an operator with condition number
    O(1e4));
  * a 24x16 quads x 140 layers channel, whose columns hold more slots than
    the matrix-core transforms take (128): the vector-unit transforms, same
    checks against the FP64 BiCGSTAB;
  * an 8x6 quads x 6 layers channel with a bed: fewer layers than the
    multigrid's 16 modes (all 6 go through it), a single level (the dense
    coarse solve on the 96 columns), one mode tile of the transforms; a
    stage of nsx_hydro in the precision form against the FP64 path;
  * the 64x32 quads x 48 layers channel: a random right-hand side solved to
    eps 1e-13 in a fraction of the block-Jacobi instance's iterations, and
    four RK2 steps of the precision form at eps 1e-13 within 1e-8 of the
    FP64 path in every field."""
import numpy as np
import pytest

import oracle
from common import bc_case, interior, load_bc, ppe_residual, rel_err
from nsx.abi import PREC_FP64, PREC_MIXED, REFINE_CGSR_BLOCK, REFINE_MODAL
from nsx.case import config_case, make_case

pytestmark = pytest.mark.gpu


def _solve_pair(c, seed):
    from nsx.device import Context
    rng = np.random.default_rng(seed)
    rhs = c.cell_field()
    rhs[:c.N_CELL0, 1:-1] = rng.standard_normal((c.N_CELL0, c.NZ - 2))
    phi0 = c.cell_field()
    G = Context(c)
    G.init_state()
    phi64, pv64, i64 = G.bicgstab(phi0, rhs)
    G.set_precision(PREC_MIXED, REFINE_MODAL, 0.1, 30)
    phim, pvm, im = G.bicgstab(phi0, rhs)
    G.close()
    return rhs, phi0, (phi64, pv64, i64), (phim, pvm, im)


@pytest.mark.parametrize("seed", [5, 6])
def test_modal_solve_random_rhs(seed):
    c = config_case(0, jitter=0.1)
    rhs, phi0, (phi64, pv64, _), (phim, pvm, im) = _solve_pair(c, seed)
    print(f"modal CG: {im['iters']} iterations, {im['outer_passes']} true residuals", flush=True)
    assert im["exit"] == "converged" and im["iters_fp32"] > 0, im
    assert min(im["e1"], im["e2"]) < c.params.eps
    e1, e2 = ppe_residual(c, phim, rhs)
    assert min(e1, e2) < 2 * c.params.eps, (e1, e2)
    phi = phi0.copy(order="F")
    O = oracle.Oracle(c)
    O.call("bicgstab", phi, O.vert(), rhs.copy(order="F"))
    assert rel_err(interior(c, phim), interior(c, phi)) <= 1e-6
    assert rel_err(interior(c, phim), interior(c, phi64)) <= 1e-6
    assert rel_err(pvm, pv64) <= 1e-6


def test_modal_solve_walled_duct():
    c = bc_case(load_bc(), "duct", "steps")
    rhs, _, (phi64, _, _), (phim, _, im) = _solve_pair(c, 7)
    assert im["exit"] == "converged", im
    e1, e2 = ppe_residual(c, phim, rhs)
    assert min(e1, e2) < 2 * c.params.eps, (e1, e2)
    assert rel_err(interior(c, phim), interior(c, phi64)) <= 1e-6


def test_modal_solve_tall_columns():
    c = make_case(24, 16, 140, jitter=0.1)
    assert c.NZ > 130
    rhs, _, (phi64, pv64, _), (phim, pvm, im) = _solve_pair(c, 8)
    print(f"modal CG, {c.NZ} layers: {im['iters']} iterations", flush=True)
    assert im["exit"] == "converged" and im["iters_fp32"] > 0, im
    e1, e2 = ppe_residual(c, phim, rhs)
    assert min(e1, e2) < 2 * c.params.eps, (e1, e2)
    assert rel_err(interior(c, phim), interior(c, phi64)) <= 1e-6
    assert rel_err(pvm, pv64) <= 1e-6


def test_modal_tiny_mesh_single_level():
    from nsx.device import Context
    c = make_case(8, 6, 6, seed=3, jitter=0.1, bed=True)
    rhs, _, (phi64, pv64, _), (phim, pvm, im) = _solve_pair(c, 9)
    assert im["exit"] == "converged" and im["iters_fp32"] > 0, im
    e1, e2 = ppe_residual(c, phim, rhs)
    assert min(e1, e2) < 2 * c.params.eps, (e1, e2)
    assert rel_err(interior(c, phim), interior(c, phi64)) <= 1e-6
    out = {}
    for prec in (PREC_FP64, PREC_MIXED):
        G = Context(c)
        G.set_params(eps=1e-13)
        G.set_precision(prec, REFINE_MODAL, 0.1, 30)
        G.init_state()
        for a, b in (("UF", "UFN"), ("VF", "VFN"), ("WF", "WFN"), ("PF", "PFN")):
            G.copy(a, b)
        info = G.hydro(0)
        assert info["exit"] == "converged", (prec, info)
        out[prec] = {f: interior(c, G.get(f)) for f in ("UFNP", "VFNP", "WFNP", "PFNP")}
        G.close()
    for f in ("UFNP", "VFNP", "WFNP", "PFNP"):
        e = rel_err(out[PREC_MIXED][f], out[PREC_FP64][f])
        assert e <= 1e-8, (f, e)


def test_modal_iterations_and_steps_64x32x48():
    from nsx.device import Context
    c = config_case(3)
    rng = np.random.default_rng(3)
    rhs = c.cell_field()
    rhs[:c.N_CELL0, 1:-1] = rng.standard_normal((c.N_CELL0, c.NZ - 2))
    phi0 = c.cell_field()
    its = {}
    for refine in (REFINE_CGSR_BLOCK, REFINE_MODAL):
        G = Context(c)
        G.set_params(eps=1e-13, max_iters=6000)
        G.init_state()
        G.set_precision(PREC_MIXED, refine, 0.1, 30)
        phi, _, info = G.bicgstab(phi0, rhs)
        G.close()
        assert info["exit"] == "converged", (refine, info)
        e1, e2 = ppe_residual(c, phi, rhs)
        assert min(e1, e2) < 2e-13, (refine, e1, e2)
        its[refine] = info["iters"]
    print(f"iterations at eps 1e-13: 4-layer blocks {its[REFINE_CGSR_BLOCK]}, vertical modes {its[REFINE_MODAL]}",
          flush=True)
    assert its[REFINE_MODAL] * 4 < its[REFINE_CGSR_BLOCK]
    out = {}
    for prec in (PREC_FP64, PREC_MIXED):
        G = Context(c)
        G.set_params(eps=1e-13, max_iters=6000)
        G.set_precision(PREC_MIXED, REFINE_MODAL, 0.1, 30)
        G.init_state()
        info = G.step(4, prec)
        assert info["exit"] == "converged", (prec, info)
        out[prec] = {f: interior(c, G.get(f)) for f in ("UFN", "VFN", "WFN", "PFN")}
        G.close()
    for f in ("UFN", "VFN", "WFN", "PFN"):
        e = rel_err(out[PREC_MIXED][f], out[PREC_FP64][f])
        assert e <= 1e-8, (f, e)
