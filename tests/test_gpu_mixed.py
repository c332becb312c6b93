"""Mixed-precision form of the pressure solve (FP32 conjugate gradients with
reliable updates against the FP64 solution: the two-kernel and the
single-reduction arrangements, with point-Jacobi, 4-layer block and line
preconditioners; and the FP64 instance of the single-reduction CG) against
the FP64 path of the same library, the oracle and a host-side residual.

Tolerances: the refined solution meets the reference's stopping rule on its
true FP64 residual (checked on the host with the stored PDE_center
coefficients, 2 eps to allow for summation order); the solution agrees with
the FP64 solves to 1e-6 relative (both stop at a residual of 1e-10 of an
operator whose condition number is O(1e4)); over five RK2 steps of the 32x16
quads x 32 layers channel the velocities agree with the FP64 path to 1e-8 and
the pressure to 1e-6.
"""
import numpy as np
import pytest

import oracle
from common import interior, ppe_residual, rel_err
from nsx.abi import PREC_FP64, PREC_FP64_CG, PREC_MIXED, REFINE_CG, REFINE_CGSR, REFINE_CGSR_BLOCK, REFINE_CGSR_LINE
from nsx.case import config_case

pytestmark = pytest.mark.gpu


def _ctx(case):
    from nsx.device import Context
    return Context(case)


MODES = {"cg": (REFINE_CG, 0.1), "cgsr": (REFINE_CGSR, 0.1), "cgsr_block": (REFINE_CGSR_BLOCK, 0.1),
         "cgsr_line": (REFINE_CGSR_LINE, 0.1)}


@pytest.mark.parametrize("mode", sorted(MODES))
@pytest.mark.parametrize("seed", [5, 6])
def test_mixed_solve_random_rhs(seed, mode):
    c = config_case(0, jitter=0.1)
    rng = np.random.default_rng(seed)
    rhs = c.cell_field()
    rhs[:c.N_CELL0, 1:-1] = rng.standard_normal((c.N_CELL0, c.NZ - 2))
    phi0 = c.cell_field()
    G = _ctx(c)
    G.init_state()
    phi64, pv64, i64 = G.bicgstab(phi0, rhs)
    G.set_precision(PREC_MIXED, *MODES[mode], 30)
    phim, pvm, im = G.bicgstab(phi0, rhs)
    G.close()
    assert im["exit"] == "converged" and im["iters_fp32"] > 0, im
    assert im["outer_passes"] >= 2, im
    assert min(im["e1"], im["e2"]) < c.params.eps
    e1, e2 = ppe_residual(c, phim, rhs)
    assert min(e1, e2) < 2 * c.params.eps, (e1, e2)
    phi = phi0.copy(order="F")
    O = oracle.Oracle(c)
    O.call("bicgstab", phi, O.vert(), rhs.copy(order="F"))
    assert rel_err(interior(c, phim), interior(c, phi)) <= 1e-6
    assert rel_err(interior(c, phim), interior(c, phi64)) <= 1e-6
    assert rel_err(pvm, pv64) <= 1e-6


@pytest.mark.parametrize("refine", [0, 1, 2])
def test_fp32_bicgstab_forms_are_refused(refine):
    """The FP32 BiCGSTAB refinements (values 0..2) are not built: FP32
    BiCGSTAB breaks down on this operator, and nsx_set_precision says so."""
    c = config_case(0, jitter=0.1)
    G = _ctx(c)
    with pytest.raises(RuntimeError, match="FP32 BiCGSTAB"):
        G.set_precision(PREC_MIXED, refine, 0.1, 30)
    G.close()


@pytest.mark.parametrize("seed", [5, 6])
def test_fp64_cg_solve_random_rhs(seed):
    """The single-reduction CG with FP64 vectors (PREC_FP64_CG): the solution
    meets the reference's rule on its host-checked residual, agrees with the
    FP64 BiCGSTAB and the oracle to the solve's tolerance, and takes about
    the FP32 form's iteration count (the same algorithm)."""
    c = config_case(0, jitter=0.1)
    rng = np.random.default_rng(seed)
    rhs = c.cell_field()
    rhs[:c.N_CELL0, 1:-1] = rng.standard_normal((c.N_CELL0, c.NZ - 2))
    phi0 = c.cell_field()
    G = _ctx(c)
    G.init_state()
    phi64, _, i64 = G.bicgstab(phi0, rhs)
    G.set_precision(PREC_FP64_CG, REFINE_CGSR_BLOCK, 0.1, 30)
    phic, _, ic = G.bicgstab(phi0, rhs)
    G.set_precision(PREC_MIXED, REFINE_CGSR_BLOCK, 0.1, 30)
    _, _, im = G.bicgstab(phi0, rhs)
    G.close()
    assert ic["exit"] == "converged" and ic["iters"] > 0 and ic["iters_fp32"] == 0, ic
    e1, e2 = ppe_residual(c, phic, rhs)
    assert min(e1, e2) < 2 * c.params.eps, (e1, e2)
    phi = phi0.copy(order="F")
    O = oracle.Oracle(c)
    O.call("bicgstab", phi, O.vert(), rhs.copy(order="F"))
    assert rel_err(interior(c, phic), interior(c, phi)) <= 1e-6
    assert rel_err(interior(c, phic), interior(c, phi64)) <= 1e-6
    assert abs(ic["iters"] - im["iters"]) <= max(5, 0.1 * im["iters"]), (ic, im)


@pytest.mark.parametrize("refine", [REFINE_CG, REFINE_CGSR, REFINE_CGSR_BLOCK, REFINE_CGSR_LINE])
def test_cg_no_solve(refine):
    """A zero right-hand side: the start meets the rule, no CG iteration runs."""
    c = config_case(0, jitter=0.1)
    G = _ctx(c)
    G.init_state()
    G.set_precision(PREC_MIXED, refine, 0.1, 30)
    phi, pv, info = G.bicgstab(c.cell_field(), c.cell_field())
    G.close()
    assert info["exit"] == "no solve" and info["iters"] == 0, info
    assert np.all(interior(c, phi) == 0.0)


@pytest.mark.parametrize("refine", [REFINE_CGSR, REFINE_CGSR_BLOCK, REFINE_CGSR_LINE])
@pytest.mark.parametrize("tol", [0.1, 0.3])
def test_single_reduction_cg_solution_meets_rule(tol, refine):
    """The x a solve returns is the one whose true residual met the rule:
    host-checked on several right-hand sides.  The launches queued after a
    stop pass it on instead of running again from a slot written before
    it (which folded alpha p into x once more)."""
    c = config_case(0, jitter=0.1)
    G = _ctx(c)
    G.init_state()
    G.set_precision(PREC_MIXED, refine, tol, 30)
    for seed in range(5, 10):
        rng = np.random.default_rng(seed)
        rhs = c.cell_field()
        rhs[:c.N_CELL0, 1:-1] = rng.standard_normal((c.N_CELL0, c.NZ - 2))
        phi, _, info = G.bicgstab(c.cell_field(), rhs)
        assert info["exit"] == "converged", (seed, info)
        e1, e2 = ppe_residual(c, phi, rhs)
        assert min(e1, e2) < 2 * c.params.eps, (seed, e1, e2, info)
        # one reliable update per decade or so, not a count inflated by
        # launches that ran again after the stop
        assert info["outer_passes"] <= 40, (seed, info)
    G.close()


@pytest.mark.parametrize("refine", [REFINE_CG, REFINE_CGSR, REFINE_CGSR_BLOCK, REFINE_CGSR_LINE])
def test_solve_after_no_solve(refine):
    """A solve right after one that stopped at its start (no launch of the
    iteration ran) starts from its own state, not the previous solve's."""
    c = config_case(0, jitter=0.1)
    rng = np.random.default_rng(13)
    rhs = c.cell_field()
    rhs[:c.N_CELL0, 1:-1] = rng.standard_normal((c.N_CELL0, c.NZ - 2))
    G = _ctx(c)
    G.init_state()
    G.set_precision(PREC_MIXED, refine, 0.1, 30)
    for _ in range(2):
        G.bicgstab(c.cell_field(), rhs)
        _, _, none = G.bicgstab(c.cell_field(), c.cell_field())
        assert none["exit"] == "no solve", none
    phi, _, info = G.bicgstab(c.cell_field(), rhs)
    G.close()
    assert info["exit"] == "converged", info
    e1, e2 = ppe_residual(c, phi, rhs)
    assert min(e1, e2) < 2 * c.params.eps, (e1, e2)


@pytest.mark.parametrize("cfg", [0, 3])
def test_preconditioners_cut_iterations(cfg):
    """Point Jacobi, 4-layer blocks and whole lines of the same matrix: each
    wider preconditioner holds more of A, and CG needs fewer iterations
    (a wrong block or line solve would slow it down or stall it)."""
    c = config_case(cfg, jitter=0.1)
    rng = np.random.default_rng(21)
    rhs = c.cell_field()
    rhs[:c.N_CELL0, 1:-1] = rng.standard_normal((c.N_CELL0, c.NZ - 2))
    its = {}
    for refine in (REFINE_CGSR, REFINE_CGSR_BLOCK, REFINE_CGSR_LINE):
        G = _ctx(c)
        G.init_state()
        G.set_precision(PREC_MIXED, refine, 0.1, 30)
        phi, _, info = G.bicgstab(c.cell_field(), rhs)
        G.close()
        assert info["exit"] == "converged", (refine, info)
        e1, e2 = ppe_residual(c, phi, rhs)
        assert min(e1, e2) < 2 * c.params.eps, (refine, e1, e2)
        its[refine] = info["iters"]
    assert its[REFINE_CGSR_LINE] < its[REFINE_CGSR_BLOCK] < its[REFINE_CGSR], its


def test_line_preconditioner_rejects_tall_columns():
    """A window of the line preconditioner holds whole columns of four-layer
    quads: columns of more than 1024 layers are refused, the other forms
    take them."""
    from nsx.case import make_case
    c = make_case(4, 2, 1100, seed=3, jitter=0.0)
    G = _ctx(c)
    with pytest.raises(RuntimeError, match="line preconditioner"):
        G.set_precision(PREC_MIXED, REFINE_CGSR_LINE, 0.1, 30)
    G.set_precision(PREC_MIXED, REFINE_CGSR_BLOCK, 0.1, 30)
    G.close()


def test_single_reduction_cg_rejects_rare_updates():
    """The single-reduction form needs a reliable update at least every 1.3
    decades (delta >= 0.05); the two-kernel form takes smaller deltas."""
    c = config_case(0, jitter=0.1)
    G = _ctx(c)
    with pytest.raises(RuntimeError, match="delta >= 0.05"):
        G.set_precision(PREC_MIXED, REFINE_CGSR, 0.01, 30)
    G.set_precision(PREC_MIXED, REFINE_CG, 0.01, 30)
    G.close()


def test_single_reduction_cg_tracks_cg():
    """The single-reduction form runs CG's iteration in another arrangement:
    in FP32 its iteration and update counts stay within a few percent of the
    two-kernel form's, and both solutions agree to the solve's tolerance."""
    c = config_case(0, jitter=0.1)
    rng = np.random.default_rng(11)
    rhs = c.cell_field()
    rhs[:c.N_CELL0, 1:-1] = rng.standard_normal((c.N_CELL0, c.NZ - 2))
    out = {}
    for refine in (REFINE_CG, REFINE_CGSR):
        G = _ctx(c)
        G.init_state()
        G.set_precision(PREC_MIXED, refine, 0.1, 30)
        out[refine] = G.bicgstab(c.cell_field(), rhs)
        G.close()
    (p1, _, i1), (p2, _, i2) = out[REFINE_CG], out[REFINE_CGSR]
    assert i1["exit"] == i2["exit"] == "converged", (i1, i2)
    assert abs(i2["iters"] - i1["iters"]) <= max(3, 0.05 * i1["iters"]), (i1, i2)
    assert abs(i2["outer_passes"] - i1["outer_passes"]) <= 3, (i1, i2)
    assert rel_err(interior(c, p2), interior(c, p1)) <= 1e-6


@pytest.mark.parametrize("refine", [REFINE_CG, REFINE_CGSR, REFINE_CGSR_BLOCK, REFINE_CGSR_LINE])
def test_cg_cap_falls_back_to_fp64(refine):
    """A CG that reaches its cap (2 MaxIters + 100 iterations) keeps x at its
    last update and the FP64 BiCGSTAB finishes from there: the solution
    still meets the rule, and the FP32 iterations are those of the cap.
    On the 64x32 x 48 mesh, where even the line-preconditioned CG needs
    enough iterations that the FP64 solve is left MaxIters of its own to
    finish with."""
    c = config_case(3)
    rng = np.random.default_rng(12)
    rhs = c.cell_field()
    rhs[:c.N_CELL0, 1:-1] = rng.standard_normal((c.N_CELL0, c.NZ - 2))
    G = _ctx(c)
    G.init_state()
    G.set_precision(PREC_MIXED, refine, 0.1, 30)
    _, _, free = G.bicgstab(c.cell_field(), rhs)
    n_cg = free["iters"]
    if n_cg < 160:
        pytest.skip(f"CG converges in {n_cg} iterations, too few to cap")
    m = (n_cg - 130) // 2  # cap = n_cg - 30 or - 31
    G.set_max_iters(m)
    phi, _, info = G.bicgstab(c.cell_field(), rhs)
    G.close()
    assert info["iters_fp32"] == 2 * m + 100, (free, info)
    assert info["exit"] == "converged" and info["iters"] > info["iters_fp32"], (free, info)
    e1, e2 = ppe_residual(c, phi, rhs)
    assert min(e1, e2) < 2 * c.params.eps, (e1, e2)


@pytest.mark.parametrize("mode", sorted(MODES))
def test_mixed_steps_vs_fp64(mode):
    c = config_case(0)
    A = _ctx(c)
    A.init_state()
    ia = A.step(5, PREC_FP64)
    B = _ctx(c)
    B.set_precision(PREC_MIXED, *MODES[mode], 30)
    B.init_state()
    ib = B.step(5, PREC_MIXED)
    assert ia["exit"] == "converged" and ib["exit"] == "converged"
    errs = {}
    for f in ("UFN", "VFN", "WFN", "PFN"):
        errs[f] = rel_err(interior(c, B.get(f)), interior(c, A.get(f)))
    A.close()
    B.close()
    assert errs["UFN"] <= 1e-8 and errs["VFN"] <= 1e-8 and errs["WFN"] <= 1e-8, errs
    assert errs["PFN"] <= 1e-6, errs


@pytest.mark.parametrize("refine", [REFINE_CG, REFINE_CGSR_BLOCK])
def test_fp32_overflow_falls_back_to_fp64(refine):
    """A right-hand side beyond FP32's range (1e40, fine in FP64): the FP32
    vectors overflow, the CG stops as a breakdown without folding a
    non-finite correction into x, and the FP64 BiCGSTAB finishes from x with
    a finite pressure that meets the reference's rule."""
    c = config_case(0, jitter=0.1)
    rng = np.random.default_rng(17)
    rhs = c.cell_field()
    rhs[:c.N_CELL0, 1:-1] = 1e40 * rng.standard_normal((c.N_CELL0, c.NZ - 2))
    G = _ctx(c)
    G.init_state()
    G.set_precision(PREC_MIXED, refine, 0.1, 30)
    phi, pv, info = G.bicgstab(c.cell_field(), rhs)
    G.close()
    assert info["exit"] == "converged", info
    assert np.all(np.isfinite(interior(c, phi))) and np.all(np.isfinite(pv))
    e1, e2 = ppe_residual(c, phi, rhs)
    assert min(e1, e2) < 2 * c.params.eps, (e1, e2)


def test_fp64_cg_steps_vs_fp64():
    """Five RK2 steps with the FP64 single-reduction CG against the FP64
    BiCGSTAB at eps 1e-13: the two FP64 solvers agree to 1e-8 in every field."""
    c = config_case(0)
    out = []
    for prec in (PREC_FP64, PREC_FP64_CG):
        G = _ctx(c)
        G.set_params(eps=1e-13)
        G.init_state()
        info = G.step(5, prec)
        assert info["exit"] == "converged", (prec, info)
        out.append({f: interior(c, G.get(f)) for f in ("UFN", "VFN", "WFN", "PFN")})
        G.close()
    errs = {f: rel_err(out[1][f], out[0][f]) for f in out[0]}
    assert max(errs.values()) <= 1e-8, errs


@pytest.mark.parametrize("cfg", [0, 3])
def test_fp32_flux_form_errors(cfg):
    """The precision form with FP32 flux and eddy-viscosity arithmetic
    (PREC_MIXED_FLUX32) against the FP64 path over three RK2 steps, both
    solves at eps 1e-13: the FP32 rate enters u* = u + dt r through dt, so
    the velocities stay within 1e-6 (measured 2e-8 to 2.3e-7); the eddy
    viscosity within 1e-6 (2e-7); the pressure, whose right-hand side is
    the divergence of the FP32 rate, within 1e-5 (2e-6).  These are above
    the 1e-8 gate of the default precision form, which keeps the fluxes in
    FP64 (printed: the errors the form is reported with)."""
    from nsx.abi import PREC_MIXED_FLUX32
    c = config_case(cfg)
    out = []
    for prec in (PREC_FP64, PREC_MIXED_FLUX32):
        G = _ctx(c)
        G.set_params(eps=1e-13)
        if prec == PREC_MIXED_FLUX32:
            G.set_precision(PREC_MIXED_FLUX32, REFINE_CGSR_BLOCK, 0.1, 30)
        G.init_state()
        info = G.step(3, prec)
        assert info["exit"] == "converged", (prec, info)
        out.append({f: interior(c, G.get(f)) for f in ("UFN", "VFN", "WFN", "PFN", "EDDY")})
        G.close()
    errs = {f: rel_err(out[1][f], out[0][f]) for f in out[0]}
    print(f"workload {cfg}: FP32 flux form vs FP64:", errs, flush=True)
    assert max(errs["UFN"], errs["VFN"], errs["WFN"]) <= 1e-6, errs
    assert errs["PFN"] <= 1e-5 and errs["EDDY"] <= 1e-6, errs
