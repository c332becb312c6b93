"""Boundary conditions other than the channel's: walled sides and the lid.

tests/golden/bc_cases.json holds what the compiled reference produced for
six boundary sets (tests/golden/make_golden.py, reference builds with XPB,
YPB, XBC, YBC and TopBC set in common.mpf):
  duct    periodic in x, no-slip walls in y (ghost flags u,v,w 0, p 1);
  slip_x  free-slip walls in x, periodic in y (u 0; v, w, p 1);
  box     no-slip walls on all four sides;
  lid     the channel with TopBC = 2 (u ghost 2 - u at the top, u = 1 at
          the top level of the interior vertices);
  inflow_x  inflow and outflow in x (XBC = 2): on the inflow side the ghosts
          are 2 inflow() - phi and the vertices inflow(), whose value is 1
          for v and 0 for u and w (Functions.F90:109-121); Neumann outflow;
  slip_y  free-slip walls in y (YBC = 1), where update_bc flags u inflow.
The host set-up, every hot-path routine of the oracle on the 6x4 x 4 mesh
and three RK2 steps on a 12x8 x 8 mesh are compared hash for hash.  Closed
forms pin the ghost rules themselves (a ghost is -u or u of its cell, an
inflow ghost of v is 2 - v, the lid's top ghost of u is 2 - u)."""
import numpy as np
import pytest

import oracle
from common import STATE_MAP, bc_case, digest, load_bc, setup_views, sha
from test_oracle import run_unit_sequence

REC = load_bc()
SETS = sorted(REC["sets"])


@pytest.mark.parametrize("name", SETS)
def test_bc_setup_bitwise(name):
    """calcul_geometry + adjust_bc + PDE_center of a walled box (mirrored
    ghost centres, vertex pairs only across periodic sides, the flags of
    update_bc.F90) == the reference's, array by array."""
    e = REC["sets"][name]
    c, _ = bc_case(REC, name, "tiny")
    assert [c.N_CELLexact, c.N_CELLghost, c.bctop, c.bcbot] == e["sizes"]
    assert c.TotalVolume == e["TotalVolume"]
    views = setup_views({k: c.arrays[k] for k in e["setup"]}, c.N_CELL, c.N_CELLexact, c.Dimsurrounding)
    bad = [k for k in e["setup"] if sha(views[k]) != e["setup"][k]]
    assert not bad, bad


@pytest.mark.parametrize("name", SETS)
def test_bc_oracle_routines_bitwise(name):
    """Every hot-path routine of the oracle on the walled (or lid) set-up ==
    the reference's, bit for bit."""
    e = REC["sets"][name]
    c, p0 = bc_case(REC, name, "tiny")
    O = oracle.Oracle(c, p0=p0)
    r = run_unit_sequence(O, c, p0)
    checked = 0
    for k, mine in r.items():
        if k in e["unit"]:
            assert digest(c, np.asarray(mine)) == e["unit"][k], k
            checked += 1
    assert checked >= 40, checked
    assert list(r["solves"]) == e["unit_solves"]


@pytest.mark.parametrize("name", SETS)
def test_bc_oracle_trajectory_bitwise(name):
    """Three RK2 steps of the oracle == the reference's, state by state,
    with the same BiCGSTAB iteration counts."""
    e = REC["sets"][name]
    c = bc_case(REC, name, "steps")
    O = oracle.Oracle(c)
    for s in range(1, REC["steps_case"]["nsteps"] + 1):
        O.step()
        for nm, want in e["steps"][str(s)].items():
            assert digest(c, O.state(STATE_MAP[nm])) == want, (s, nm)
    assert O.solver_record()["iter_total"] == sum(e["solves"])


def test_ghost_rules_closed_form():
    """The ghost rules on a field that is 1 + the cell index: a wall ghost of
    u is -u of its cell (no slip), a free-slip ghost of v is v, an inflow
    ghost of v is 2 - v and of u is -u, a periodic ghost is its partner's
    value, and the lid's top ghost of u is 2 - u."""
    for name, checks in (("duct", "wall"), ("slip_x", "slip"), ("lid", "lid"), ("inflow_x", "inflow")):
        c, _ = bc_case(REC, name, "tiny")
        O = oracle.Oracle(c)
        N0, NZ = c.N_CELL0, c.NZ
        f = np.zeros((c.N_CELL, NZ), order="F")
        f[:N0, :] = 1.0 + np.arange(N0)[:, None] + 0.01 * np.arange(NZ)[None, :]
        u, v = f.copy(order="F"), f.copy(order="F")
        w = f.copy(order="F")
        O.call("bcglobalvc", u, v, w, O.vert(), O.vert(), O.vert(), tagdir=2)
        ncp = c.No_cp
        for i in range(N0):
            for j in range(3):
                g = ncp[i, j] - 1
                if g < N0:
                    continue
                owner = u[i, 1:-1]
                if c.bcuc[g] == 3:
                    partner = ncp[g, 2] - 1
                    assert np.array_equal(u[g, 1:-1], u[partner, 1:-1])
                elif c.bcuc[g] == 0:
                    assert np.array_equal(u[g, 1:-1], -owner)
                if c.bcvc[g] == 1:
                    assert np.array_equal(v[g, 1:-1], v[i, 1:-1])
                if c.bcvc[g] == 2:
                    assert np.array_equal(v[g, 1:-1], 2.0 - v[i, 1:-1])
                    assert np.array_equal(u[g, 1:-1], -owner) and c.bcuc[g] == 2
        if checks == "lid":
            assert np.array_equal(u[:N0, NZ - 1], 2.0 - u[:N0, NZ - 2])
            assert np.array_equal(v[:N0, NZ - 1], -v[:N0, NZ - 2])
        if checks == "slip":
            assert (c.bcvc[N0:c.N_CELLexact] == 1).sum() > 0
        if checks == "inflow":
            assert (c.bcvc[N0:c.N_CELLexact] == 2).sum() > 0
