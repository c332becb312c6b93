"""The differentiation form on the GPU: hand-written tangent-linear and
adjoint kernels of the projection stage and the RK2 step, the checkpointed
adjoint of the time-averaged bottom-wall shear with respect to a per-prism
streamwise body force, and its torch.autograd harness.

Checks, each fixed by the mathematics rather than by a second copy of the
kernels:
- one RK2 step: the adjoint map is the transpose of the tangent-linear map,
  <b, T t> = <A b, t> on random interior fields (dot-product test), to the
  pressure solves' tolerance (eps = 1e-13: 1e-8 relative);
- the tangent-linear model against central finite differences of the
  nonlinear forward steps (h = 1e-6);
- the 64x32 quads x 48 layers channel (196,608 prisms), 20 steps at dt 5e-4,
  body force 1 + a seeded Gaussian random field (correlation length 0.5, std
  0.1): the adjoint gradient (checkpoints every 5 steps) along 8 seeded
  random directions against central finite differences with h = 1e-6 (1e-5
  relative) and against the FP64 tangent-linear run (1e-6 relative); solves
  to eps = 1e-13, so that the finite differences are not swamped by the
  solver's stopping error (FP64 BiCGSTAB does not reach 1e-14 on every
  perturbed force of this mesh).  Both derivatives are exact up to their
  pressure solves' stopping errors; the adjoint solves run a hundred times
  tighter than eps (FP64 single-reduction CG), and the adjoint and tangent
  gradients agree to about 1e-15;
- torch.autograd: the gradient of J and of a function of J through
  WallShear.backward equals the adjoint gradient (to 1e-7 of its largest
  entry: the adjoint kernels scatter with atomic adds, so two runs differ by
  rounding, which the adjoint pressure solves carry to their stopping error).
"""
import numpy as np
import pytest

from nsx.case import config_case, gaussian_random_field, make_case

pytestmark = pytest.mark.gpu


def _ctx(case, eps):
    from nsx.device import Context
    c = Context(case)
    c.set_params(eps=eps, max_iters=6000)
    return c


def _rand_state(case, rng):
    """u, v, w, p (N_CELL,NZ) and five faces (N_CELL0,NZ), random on the
    interior prisms and zero elsewhere"""
    out = []
    for q in range(9):
        a = case.cell_field() if q < 4 else np.zeros((case.N_CELL0, case.NZ), order="F")
        a[:case.N_CELL0, 1:-1] = rng.standard_normal((case.N_CELL0, case.NZ - 2))
        out.append(a)
    return out


def _dot(a, b, case):
    return float(sum(np.sum(x[:case.N_CELL0] * y[:case.N_CELL0]) for x, y in zip(a, b)))


@pytest.mark.parametrize("steps_before", [0, 2])
def test_step_adjoint_is_transpose_of_tangent(steps_before):
    case = config_case(0, jitter=0.1)
    G = _ctx(case, 1e-13)
    rng = np.random.default_rng(11 + steps_before)
    G.set_body_force(1.0 + gaussian_random_field(case, 3))
    G.init_state()
    if steps_before:
        G.step(steps_before)
    t = _rand_state(case, rng)
    df = rng.standard_normal((case.N_CELL0, case.NZ - 2))
    b = _rand_state(case, rng)
    u_before = G.get("UFN")
    tout = G.step_tangent(t, df)
    bout, gf = G.step_adjoint(b)
    assert np.array_equal(G.get("UFN"), u_before), "the single-step maps changed the resident state"
    G.close()
    lhs = _dot(b, tout, case)
    rhs = _dot(bout, t, case) + float(np.sum(gf[:case.N_CELL0, 1:-1] * df))
    print(f"dot-product test: <b, T t> = {lhs:.15e}, <A b, t> = {rhs:.15e}", flush=True)
    assert abs(lhs - rhs) <= 1e-8 * max(abs(lhs), abs(rhs)), (lhs, rhs)


def test_tangent_linear_step_vs_finite_differences():
    """The tangent of the wall shear after 3 steps against central
    differences of the nonlinear steps, on a small jittered channel."""
    case = make_case(8, 6, 8, seed=5, jitter=0.1, bed=True)
    G = _ctx(case, 1e-14)
    f = 1.0 + gaussian_random_field(case, 4)
    d = gaussian_random_field(case, 40, std=1.0)
    h = 1e-6
    G.set_body_force(f)
    G.init_state()
    J0, dJ = G.wall_shear_tangent(3, d)
    Jp, Jm = [], []
    for sgn, out in ((1.0, Jp), (-1.0, Jm)):
        G.set_body_force(f + sgn * h * d)
        G.init_state()
        out.append(G.wall_shear(3))
    G.close()
    fd = (Jp[0] - Jm[0]) / (2 * h)
    print(f"J {J0:.15e}, tangent {dJ:.15e}, central difference {fd:.15e}", flush=True)
    assert abs(dJ - fd) <= 1e-6 * abs(fd), (dJ, fd)


def test_wall_shear_gradient_64x32x48():
    from nsx.adjoint import WallShearProblem

    case = config_case(3)
    K, every, h = 20, 5, 1e-6
    f = 1.0 + gaussian_random_field(case, 7, corr=0.5, std=0.1)
    P = WallShearProblem(case, K, every, eps=1e-13)
    J0, g = P.gradient(f)
    J1 = P.objective(f)
    assert J1 == J0, (J0, J1)  # the forward is bit-reproducible
    rows = []
    for m in range(8):
        d = gaussian_random_field(case, 100 + m, corr=0.5, std=1.0)
        adj = float(np.sum(g * d))
        _, tlm = P.tangent(f, d)
        fd = (P.objective(f + h * d) - P.objective(f - h * d)) / (2 * h)
        rows.append((adj, tlm, fd))
        print(f"direction {m}: adjoint {adj:.12e} tangent {tlm:.12e} central difference {fd:.12e} "
              f"(adj-fd)/fd {(adj - fd) / fd:.2e} (adj-tlm)/tlm {(adj - tlm) / tlm:.2e}", flush=True)
    P.close()
    for adj, tlm, fd in rows:
        assert abs(adj - fd) <= 1e-5 * abs(fd), (adj, fd)
        assert abs(adj - tlm) <= 1e-6 * abs(tlm), (adj, tlm)


def test_autograd_harness():
    import torch

    from nsx.adjoint import WallShearProblem, wall_shear

    case = make_case(8, 6, 8, seed=5, jitter=0.1, bed=True)
    P = WallShearProblem(case, 4, 2, eps=1e-13)
    f0 = 1.0 + gaussian_random_field(case, 4)
    f = torch.tensor(f0, dtype=torch.float64, requires_grad=True)
    J = wall_shear(f, P)
    loss = 3.0 * J ** 2
    loss.backward()
    Jn, g = P.gradient(f0)
    P.close()
    assert J.item() == Jn
    ref = 6.0 * Jn * g
    np.testing.assert_allclose(f.grad.numpy(), ref, rtol=0.0, atol=1e-7 * np.abs(ref).max())
