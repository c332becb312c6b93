"""The extrapolated initial guess of the precision forms' pressure solves
(nsx_set_initial_guess, NSX_GUESS_EXTRAPOLATE: 2 p_k - p_{k-1} from the same
RK stage's last two solutions).

  * eight RK2 steps of the 64x32 quads x 48 layers channel in the precision
    form at eps 1e-13, started from the stage's input pressure and from the
    extrapolation: every field agrees to 1e-9 (both solves meet the same
    stopping rule on their true FP64 residual, so they differ by the
    solver's tolerance, not by the guess), and the extrapolated run needs
    fewer iterations;
  * the FP64 path (the reference's BiCGSTAB and its initial guess) is left
    alone: its steps are bit for bit the same with the option set;
  * the stored solutions start again after nsx_init_state: two runs of the
    same steps from the initial field are bit for bit the same;
  * unknown modes are refused."""
import numpy as np
import pytest

from common import interior, rel_err
from nsx.abi import GUESS_EXTRAPOLATE, GUESS_PREVIOUS, PREC_FP64, PREC_MIXED, REFINE_CGSR_BLOCK
from nsx.case import config_case

pytestmark = pytest.mark.gpu

FIELDS = ("UFN", "VFN", "WFN", "PFN")


def _run(case, mode, precision, nsteps, eps=1e-13):
    from nsx.device import Context
    c = Context(case)
    c.set_params(eps=eps)
    c.set_precision(PREC_MIXED, REFINE_CGSR_BLOCK, 0.1, 30)
    c.set_initial_guess(mode)
    c.init_state()
    info = c.step(nsteps, precision)
    out = {f: interior(case, c.get(f)) for f in FIELDS}
    c.close()
    return out, info


def test_extrapolated_guess_same_steps_fewer_iterations():
    case = config_case(3)
    a, ia = _run(case, GUESS_PREVIOUS, PREC_MIXED, 8)
    b, ib = _run(case, GUESS_EXTRAPOLATE, PREC_MIXED, 8)
    assert ia["exit"] == ib["exit"] == "converged"
    for f in FIELDS:
        e = rel_err(b[f], a[f])
        assert e <= 1e-9, (f, e)
    print(f"iterations: input pressure {ia['iters_total']}, extrapolated {ib['iters_total']}", flush=True)
    assert ib["iters_total"] < ia["iters_total"]


def test_fp64_path_untouched():
    case = config_case(3)
    a, _ = _run(case, GUESS_PREVIOUS, PREC_FP64, 3, eps=1e-10)
    b, _ = _run(case, GUESS_EXTRAPOLATE, PREC_FP64, 3, eps=1e-10)
    for f in FIELDS:
        assert np.array_equal(a[f], b[f]), f


def test_guess_restarts_after_init_state():
    from nsx.device import Context
    case = config_case(3)
    c = Context(case)
    c.set_params(eps=1e-13)
    c.set_precision(PREC_MIXED, REFINE_CGSR_BLOCK, 0.1, 30)
    c.set_initial_guess(GUESS_EXTRAPOLATE)
    runs = []
    for _ in range(2):
        c.init_state()
        c.step(4, PREC_MIXED)
        runs.append({f: interior(case, c.get(f)) for f in FIELDS})
    with pytest.raises(RuntimeError, match="initial-guess"):
        c.set_initial_guess(2)
    c.close()
    for f in FIELDS:
        assert np.array_equal(runs[0][f], runs[1][f]), f
