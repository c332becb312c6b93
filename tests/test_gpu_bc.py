"""The HIP library on boundary sets other than the channel's, against the
oracle (which equals the compiled reference on them hash for hash,
tests/test_bc.py): walled sides (wall ghost columns filled by the rules of
BCvelcenter3D and BCpcenter: no slip, free slip, inflow) and the lid
(TopBC = 2).

  * three RK2 steps on the 12x8 quads x 8 layers mesh of each set, FP64:
    velocities within 1e-8, pressure within 1e-6 (the solver's summation
    order, as for the channel), the wall ghost rows of the new velocity
    within 1e-8;
  * the same steps in the precision form at eps 1e-13 against the FP64 path
    at 1e-13, every field within 1e-8;
  * the adjoint of one RK2 step is the transpose of its tangent on the
    walled sets (dot-product test), so the wall ghost folds transpose the
    wall ghost fills;
  * a 64x32 quads x 32 layers duct: the no-slip walls hold (the face
    velocity through a wall face is zero), the pressure keeps a zero volume
    mean and the velocities stay finite and bounded.
The per-routine kernels run on the walled tiny meshes in
tests/test_gpu_parity.py (its bc_* cases)."""
import numpy as np
import pytest

import oracle
from common import bc_case, interior, load_bc, rel_err
from nsx.case import gaussian_random_field, make_case

pytestmark = pytest.mark.gpu

REC = load_bc()
SETS = sorted(REC["sets"])


def _ctx(case, eps=None):
    from nsx.device import Context
    c = Context(case)
    if eps is not None:
        c.set_params(eps=eps, max_iters=6000)
    return c


@pytest.mark.parametrize("name", SETS)
def test_bc_trajectory_vs_oracle(name):
    c = bc_case(REC, name, "steps")
    O = oracle.Oracle(c)
    G = _ctx(c)
    G.init_state()
    for _ in range(REC["steps_case"]["nsteps"]):
        O.step()
        G.step(1)
    for dev, orc, tol in (("UFN", "ufn", 1e-8), ("VFN", "vfn", 1e-8), ("WFN", "wfn", 1e-8), ("PFN", "pfn", 1e-6)):
        e = rel_err(interior(c, G.get(dev)), interior(c, O.state(orc)))
        assert e <= tol, (name, dev, e)
    ghosts = slice(c.N_CELL0, c.N_CELLexact)
    for dev, orc in (("UFN", "ufn"), ("VFN", "vfn"), ("WFN", "wfn")):
        a, b = G.get(dev)[ghosts, 1:-1], O.state(orc)[ghosts, 1:-1]
        assert np.abs(a - b).max() <= 1e-8 * max(np.abs(b).max(), 1.0), (name, dev)
    G.close()


@pytest.mark.parametrize("name", SETS)
def test_bc_precision_form_vs_fp64(name):
    from nsx import abi
    c = bc_case(REC, name, "steps")
    out = {}
    for prec in (abi.PREC_FP64, abi.PREC_MIXED):
        G = _ctx(c, 1e-13)
        G.init_state()
        G.step(REC["steps_case"]["nsteps"], prec)
        out[prec] = {f: interior(c, G.get(f)) for f in ("UFN", "VFN", "WFN", "PFN")}
        G.close()
    for f in ("UFN", "VFN", "WFN", "PFN"):
        e = rel_err(out[abi.PREC_MIXED][f], out[abi.PREC_FP64][f])
        assert e <= 1e-8, (name, f, e)


def _rand_state(case, rng):
    out = []
    for q in range(9):
        a = case.cell_field() if q < 4 else np.zeros((case.N_CELL0, case.NZ), order="F")
        a[:case.N_CELL0, 1:-1] = rng.standard_normal((case.N_CELL0, case.NZ - 2))
        out.append(a)
    return out


def _dot(a, b, case):
    return float(sum(np.sum(x[:case.N_CELL0] * y[:case.N_CELL0]) for x, y in zip(a, b)))


@pytest.mark.parametrize("name", ["duct", "slip_x", "box"])
def test_bc_step_adjoint_is_transpose_of_tangent(name):
    c = bc_case(REC, name, "steps")
    G = _ctx(c, 1e-13)
    rng = np.random.default_rng(21)
    G.set_body_force(1.0 + gaussian_random_field(c, 3))
    G.init_state()
    G.step(1)
    t = _rand_state(c, rng)
    df = rng.standard_normal((c.N_CELL0, c.NZ - 2))
    b = _rand_state(c, rng)
    tout = G.step_tangent(t, df)
    bout, gf = G.step_adjoint(b)
    G.close()
    lhs = _dot(b, tout, c)
    rhs = _dot(bout, t, c) + float(np.sum(gf[:c.N_CELL0, 1:-1] * df))
    print(f"{name}: <b, T t> = {lhs:.15e}, <A b, t> = {rhs:.15e}", flush=True)
    assert abs(lhs - rhs) <= 1e-8 * max(abs(lhs), abs(rhs)), (lhs, rhs)


@pytest.mark.parametrize("name,why", [("lid", "lid"), ("inflow_x", "inflow")])
def test_refused_by_the_differentiation_form(name, why):
    c = bc_case(REC, name, "steps")
    G = _ctx(c, 1e-12)
    G.init_state()
    with pytest.raises(RuntimeError, match=why):
        G.wall_shear(2, 1)
    G.close()


def test_duct_64x32x32_walls_hold():
    """Two RK2 steps of a jittered 64x32 quads x 32 layers duct (no-slip
    walls in y): the solves converge, the face velocities through the wall
    faces are zero, the pressure has zero volume mean (Glob_Norm) and the
    velocities stay finite and bounded."""
    c = make_case(64, 32, 32, seed=3, jitter=0.1, walls=(1, 0, 0, 0))
    G = _ctx(c)
    G.init_state()
    info = G.step(2)
    assert info["exit"] == "converged" and min(info["e1"], info["e2"]) < c.params.eps
    wall = np.zeros((c.N_CELL0, 3), bool)
    for j in range(3):
        g = c.No_cp[:c.N_CELL0, j] - 1
        wall[:, j] = (g >= c.N_CELL0) & (c.bcuc[np.maximum(g, 0)] == 0)
    assert wall.any()
    for j, f in enumerate(("U1FACE", "U2FACE", "U3FACE")):
        F = G.get(f)[:, 1:-1]
        if wall[:, j].any():
            assert np.abs(F[wall[:, j]]).max() <= 1e-12 * np.abs(F).max(), f
    p = interior(c, G.get("PFN"))
    vol = c.areaCell[:c.N_CELL0, None] * c.dsig[None, 1:-1]
    assert abs((p * vol).sum()) <= 1e-10 * np.abs(p * vol).sum()
    u = interior(c, G.get("UFN"))
    assert np.isfinite(u).all() and np.abs(u).max() < 5 * c.arrays["u0"].max()
    G.close()
