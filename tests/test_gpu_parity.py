"""HIP library (through the C ABI) against the oracle on the same seeded inputs.

Element-wise operators are compiled without FMA contraction and keep the
reference's operation order, so they must match the oracle bit for bit.  The
exceptions and their tolerances:
  * sgles: the cube root of the filter width uses the device pow
    (relative 1e-14);
  * everything downstream of a BiCGSTAB solve: the dot products are summed
    in a different order than the reference's sequential loops, so the
    Krylov iterates drift apart at round-off level and the solve stops within
    a few iterations of the reference's.  Tolerances (relative L2 over
    interior cells): pressure (volume-mean removed as the reference does)
    1e-6, velocities and face fluxes 1e-8; iteration counts within 10 % for
    the PPE of a time step, 25 % for a random right-hand side.
"""
import numpy as np
import pytest

import oracle
from common import bc_case, interior, load_bc, rel_err, tiny_case
from nsx.case import config_case

pytestmark = pytest.mark.gpu


def _ctx(case):
    from nsx.device import Context
    return Context(case)


def _rows(case, a):
    """interior + ghost rows (what the reference fills), all layers"""
    return np.asarray(a)[:case.N_CELLexact]


# the tiny mesh, workload 0, and the tiny mesh with the boundary sets of the
# reference's walled builds (a duct, free-slip x walls, a walled box, the lid,
# inflow and outflow in x, free-slip y walls)
@pytest.fixture(scope="module", params=["tiny", "cfg0", "bc_duct", "bc_slip_x", "bc_box", "bc_lid", "bc_inflow_x",
                                        "bc_slip_y"])
def pair(request):
    if request.param == "tiny":
        c, p0 = tiny_case()
    elif request.param.startswith("bc_"):
        c, p0 = bc_case(load_bc(), request.param[3:], "tiny")
    else:
        c = config_case(0, jitter=0.1)
        rng = np.random.default_rng(11)
        p0 = np.zeros((c.N_CELL, c.NZ), order="F")
        p0[:c.N_CELL0, 1:-1] = rng.uniform(-1, 1, (c.N_CELL0, c.NZ - 2))
    O = oracle.Oracle(c, p0=p0)
    G = _ctx(c)
    G.init_state(p0=p0)
    yield c, p0, O, G
    G.close()


def _st(O, n):
    return O.state(n).copy(order="F")


def test_init_state_bitwise(pair):
    c, p0, O, G = pair
    for dev, orc in (("UFN", "ufn"), ("VFN", "vfn"), ("WFN", "wfn"), ("PFN", "pfn")):
        assert np.array_equal(_rows(c, G.get(dev)), _rows(c, O.state(orc))), dev
    for dev, orc in (("UFV", "ufv"), ("VFV", "vfv"), ("WFV", "wfv"), ("PFV", "pfv")):
        assert np.array_equal(G.get(dev), O.state(orc)), dev
    for dev, orc in (("U1FACE", "U1"), ("U2FACE", "U2"), ("U3FACE", "U3"), ("UTFACE", "UT"), ("UBFACE", "UB")):
        assert np.array_equal(G.get(dev)[:, 1:-1], O.state(orc)[:, 1:-1]), dev


@pytest.mark.parametrize("tagdir", [2, 3])
def test_bcglobalvc_bcglobalp_bitwise(pair, tagdir):
    c, p0, O, G = pair
    u, v, w, p = _st(O, "ufn"), _st(O, "vfn"), _st(O, "wfn"), _st(O, "pfn")
    got = G.BCglobalVC(u, v, w, tagdir)
    uv, vv, wv = O.vert(), O.vert(), O.vert()
    O.call("bcglobalvc", u, v, w, uv, vv, wv, tagdir=tagdir)
    for a, b in zip(got[:3], (u, v, w)):
        assert np.array_equal(_rows(c, a), _rows(c, b))
    for a, b in zip(got[3:], (uv, vv, wv)):
        assert np.array_equal(a, b)
    gp, gpv = G.BCglobalP(p, tagdir)
    pv = O.vert()
    O.call("bcglobalp", p, pv, tagdir=tagdir)
    assert np.array_equal(_rows(c, gp), _rows(c, p))
    assert np.array_equal(gpv, pv)


def test_gradient_and_faces_bitwise(pair):
    c, p0, O, G = pair
    p, u = _st(O, "pfn"), _st(O, "ufn")
    gx, gy, gs = O.cell(), O.cell(), O.cell()
    O.call("grad_lsm", gx, gy, gs, p)
    for a, b in zip(G.grandientLSM(p), (gx, gy, gs)):
        assert np.array_equal(_rows(c, a)[:, 1:-1], _rows(c, b)[:, 1:-1])
        assert np.array_equal(a[:c.N_CELL0], b[:c.N_CELL0])
    uv = O.state("ufv").copy(order="F")
    f = [O.face() for _ in range(5)]
    O.call("calcul_face", *f, u, uv)
    for a, b in zip(G.calcul_face(u, uv), f):
        assert np.array_equal(a[:, 1:-1], b[:, 1:-1])


def test_sgles(pair):
    c, p0, O, G = pair
    u, v, w = _st(O, "ufn"), _st(O, "vfn"), _st(O, "wfn")
    ed = G.sgles(u, v, w)
    O.call("sgles", u.copy(order="F"), v.copy(order="F"), w.copy(order="F"))
    ref = O.state("eddy")
    assert rel_err(interior(c, ed), interior(c, ref)) <= 1e-14
    assert np.abs(interior(c, ed) - interior(c, ref)).max() <= 1e-14 * np.abs(interior(c, ref)).max()


def test_momentum_source_rhsp_newu_bitwise(pair):
    c, p0, O, G = pair
    prm = c.params
    u, v, w, p = _st(O, "ufn"), _st(O, "vfn"), _st(O, "wfn"), _st(O, "pfn")
    # Source3D: divergence and the face velocities that AdvDiffNS then advects with
    got = G.Source3D(u, v, w)
    src = O.cell()
    O.call("source3d", src, u.copy(order="F"), v.copy(order="F"), w.copy(order="F"))
    assert np.array_equal(got[0][:c.N_CELL0, 1:-1], src[:c.N_CELL0, 1:-1])
    for a, n in zip(got[1:], ("U1", "U2", "U3", "UT", "UB")):
        assert np.array_equal(a[:, 1:-1], O.state(n)[:, 1:-1]), n
    # AdvDiffNS with an eddy field computed by the oracle
    O.call("sgles", u.copy(order="F"), v.copy(order="F"), w.copy(order="F"))
    ed = _st(O, "eddy")
    rb = [O.cell() + prm.bdx, O.cell() + prm.bdy, O.cell() + prm.bdz]
    got = G.AdvDiffNS(rb[0], rb[1], rb[2], u, v, w, ed)
    O.call("advdiff", rb[0], rb[1], rb[2], u.copy(order="F"), v.copy(order="F"), w.copy(order="F"), ed)
    for a, b in zip(got, rb):
        assert np.array_equal(interior(c, a), interior(c, b))
    # update_rhsp
    rhs = src.copy(order="F")
    rhs[:c.N_CELL0, 1:-1] /= prm.dt
    grhs, gpv = G.update_rhsp(rhs, p)
    pv = O.vert()
    O.call("update_rhsp", rhs, p.copy(order="F"), pv)
    assert np.array_equal(interior(c, grhs), interior(c, rhs))
    assert np.array_equal(gpv, pv)
    # newu3D on the resident faces (those of Source3D above)
    faces, gpv2 = G.newu3D(p)
    O.call("source3d", O.cell(), u.copy(order="F"), v.copy(order="F"), w.copy(order="F"))
    pv2 = O.vert()
    O.call("newu3d", p.copy(order="F"), pv2)
    assert np.array_equal(gpv2, pv2)
    for a, n in zip(faces, ("U1", "U2", "U3", "UT", "UB")):
        assert np.array_equal(a[:, 1:-1], O.state(n)[:, 1:-1]), n


def test_bicgstab_vs_oracle(pair):
    c, p0, O, G = pair
    rng = np.random.default_rng(5)
    rhs = c.cell_field()
    rhs[:c.N_CELL0, 1:-1] = rng.standard_normal((c.N_CELL0, c.NZ - 2))
    phi0 = _st(O, "pfn")
    gphi, gphiv, info = G.bicgstab(phi0, rhs)
    phi, phiv = phi0.copy(order="F"), O.vert()
    O.call("bicgstab", phi, phiv, rhs.copy(order="F"))
    rec = O.solver_record()
    assert info["exit"] == "converged" and min(info["e1"], info["e2"]) < c.params.eps
    # a random right-hand side converges erratically: the iteration count at
    # which round-off-perturbed iterates cross eps varies more than for the
    # smooth PPE sources of a time step
    assert abs(info["iters"] - rec["iter"]) <= max(5, 0.25 * rec["iter"])
    assert rel_err(interior(c, gphi), interior(c, phi)) <= 1e-6
    assert rel_err(gphiv, phiv) <= 1e-6


@pytest.fixture(scope="module")
def cfg0_runs():
    c = config_case(0)
    O = oracle.Oracle(c)
    G = _ctx(c)
    G.init_state()
    return c, O, G


def test_one_stage_vs_oracle(cfg0_runs):
    c, O, G = cfg0_runs
    G.copy("UF", "UFN"), G.copy("VF", "VFN"), G.copy("WF", "WFN"), G.copy("PF", "PFN")
    info = G.hydro(0)
    O2 = oracle.Oracle(c)
    for n in ("uf", "vf", "wf", "pf"):
        O2.state(n)[:] = O2.state(n[0] + "fn")
    O2.hydro()
    rec = O2.solver_record()
    assert abs(info["iters"] - rec["iter"]) <= max(3, 0.1 * rec["iter"])
    for dev, orc, tol in (("UFNP", "ufnp", 1e-8), ("VFNP", "vfnp", 1e-8), ("WFNP", "wfnp", 1e-8),
                          ("PFNP", "pfnp", 1e-6), ("EDDY", "eddy", 1e-13), ("RHSUF", "rhsuf", 1e-13),
                          ("U1FACE", "U1", 1e-8), ("UTFACE", "UT", 1e-8), ("UFV", "ufv", 1e-8),
                          ("PFV", "pfv", 1e-6)):
        a, b = G.get(dev), O2.state(orc)
        if dev.startswith("U") and "FACE" in dev:
            a, b = a[:, 1:-1], b[:, 1:-1]
        elif dev.endswith("V"):
            pass
        else:
            a, b = interior(c, a), interior(c, b)
        assert rel_err(a, b) <= tol, (dev, rel_err(a, b))


def test_trajectory_cfg0_vs_oracle():
    """Ten RK2 steps of workload 0 (32x16 quads x 32 layers)."""
    c = config_case(0)
    O = oracle.Oracle(c)
    G = _ctx(c)
    G.init_state()
    its_o, its_g = [], []
    for s in range(10):
        O.step()
        its_o.append(O.solver_record()["iter_total"])
        its_g.append(G.step(1)["iters_total"])
    assert abs(its_g[-1] - its_o[-1]) <= 0.1 * its_o[-1]
    for dev, orc, tol in (("UFN", "ufn", 1e-7), ("VFN", "vfn", 1e-7), ("WFN", "wfn", 1e-7), ("PFN", "pfn", 1e-5)):
        e = rel_err(interior(c, G.get(dev)), interior(c, O.state(orc)))
        assert e <= tol, (dev, e)
    G.close()


def test_deterministic_reductions():
    """Two runs of the same case give the same bits (fixed-order grid reductions)."""
    c = config_case(0)
    out = []
    for _ in range(2):
        G = _ctx(c)
        G.init_state()
        info = G.step(2)
        out.append((G.get("UFN"), G.get("PFN"), info["iters_total"]))
        G.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    assert out[0][2] == out[1][2]


def test_full_size_step_properties():
    """Workload 1 (128x64 quads x 64 layers, 1,048,576 prisms), one RK2 step:
    both solves converge below eps, the pressure has zero volume mean
    (Glob_Norm) and the fields stay finite and bounded."""
    c = config_case(1)
    G = _ctx(c)
    G.init_state()
    info = G.step(1)
    assert info["n_solves"] == 2 and info["exit"] == "converged"
    assert min(info["e1"], info["e2"]) < c.params.eps
    p = interior(c, G.get("PFN"))
    vol = c.areaCell[:c.N_CELL0, None] * c.dsig[None, 1:-1]
    assert abs((p * vol).sum()) <= 1e-10 * np.abs(p * vol).sum()
    u = interior(c, G.get("UFN"))
    assert np.isfinite(u).all() and np.abs(u).max() < 5 * c.arrays["u0"].max()
    G.close()


def test_glob_cfl_vs_oracle():
    """glob_cfl on the device against the oracle on the same fields and faces
    (the oracle's state after three steps of workload 0): the selected
    extremes and the CFL number are per-cell values and must be equal; the
    sums differ in order only."""
    c = config_case(0)
    O = oracle.Oracle(c)
    for _ in range(3):
        O.step()
    ref = O.glob_cfl()
    G = _ctx(c)
    G.init_state()
    for dev, orc in (("UFN", "ufn"), ("VFN", "vfn"), ("WFN", "wfn"), ("U1FACE", "U1"), ("U2FACE", "U2"),
                     ("U3FACE", "U3"), ("UTFACE", "UT"), ("UBFACE", "UB")):
        G.put(dev, O.state(orc))
    got = G.glob_cfl()
    G.close()
    for k in ("umax", "vmax", "wmax", "cfl", "flux_max"):
        assert got[k] == ref[k], (k, got[k], ref[k])
    for k in ("aveu", "avev", "avew"):
        assert abs(got[k] - ref[k]) <= 1e-13 * max(abs(ref["aveu"]), 1.0), (k, got[k], ref[k])
    scale = abs(ref["flux_max"]) * c.n_interior
    for k in ("flux_total", "flux_bc", "flux_bot", "flux_top", "flux_side"):
        assert abs(got[k] - ref[k]) <= 1e-13 * scale, (k, got[k], ref[k])


def test_cached_setup_drives_the_device(tmp_path):
    """A device context set up from the cached set-up file (nsx_case_load)
    steps exactly as one set up from the freshly built case: the same bits
    after two RK2 steps."""
    from nsx.case import load_case, make_case
    path = str(tmp_path / "case.bin")
    c = make_case(32, 16, 32, seed=1234, jitter=0.1, save=path)
    d = load_case(path, params=c.params)
    out = []
    for case in (c, d):
        G = _ctx(case)
        G.init_state()
        G.step(2)
        out.append([G.get(n) for n in ("UFN", "VFN", "WFN", "PFN")])
        G.close()
    for a, b in zip(*out):
        assert np.array_equal(a, b)
