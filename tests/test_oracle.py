"""The oracle (CPU restatement) and the host set-up against the reference.

The fixtures in tests/golden were written by the reference's own Fortran
routines (tests/golden/make_golden.py); every comparison here is bit-exact.
Closed-form checks pin what a fixture could not tell apart by itself.
"""
import hashlib
import os

import numpy as np
import pytest

import oracle
from common import STATE_MAP, cell_coordinates, digest, load_steps, load_tiny, tiny_case, vertex_coordinates
from nsx.case import config_case, make_case

GEOM_CELL = ["xc", "yc", "areaCell", "No_cp", "bcuc", "bcvc", "bcwc", "bcpc"]
GEOM_FULL = ["dxVV", "dyVV", "dlVV", "xnn", "ynn", "dxCC", "dyCC", "dlCC", "xme", "yme", "xmo", "ymo", "doe",
             "dle1", "dle2", "dle3", "sum_xc2", "sum_yc2", "sum_xcyc", "dlVsum", "Dimsurrounding",
             "vertex_pair", "vertex_edge", "tagXV", "tagYV", "bcuv", "bcvv", "bcwv", "bcpv", "sig", "dsig", "sigv",
             "dsigv", "ic1tec", "ic2tec", "ic3tec", "Numictec"]


def test_setup_geometry_bitwise():
    """host/nsx_case.c == calcul_geometry + calcul_bc + PDE_center of the reference."""
    c, _ = tiny_case()
    g = load_tiny()
    ne = c.N_CELLexact
    for k in GEOM_CELL:
        assert np.array_equal(g[k][:ne], c.arrays[k][:ne]), k
    for k in GEOM_FULL:
        assert np.array_equal(g[k], c.arrays[k]), k
    for k in ["AmC0", "AmC1", "AmC2", "AmC3", "AmCT", "AmCB"]:
        assert np.array_equal(g[k][:, 1:-1], c.arrays[k][:, 1:-1]), k
    ds = c.Dimsurrounding
    for v in range(c.N_VERT):
        assert np.array_equal(g["surrounding"][v, :ds[v]], c.surrounding[v, :ds[v]])
        assert np.array_equal(g["weight"][v, :ds[v]], c.weight[v, :ds[v]])
    assert list(g["sizes"]) == [c.N_CELLexact, c.N_CELLghost, c.bctop, c.bcbot]
    assert g["TotalVolume"][0] == c.TotalVolume


def run_unit_sequence(O, c, p0):
    """The single-routine sequence of nsx_ref_driver's unit mode, on the oracle."""
    r = {}
    for nm, st in (("ufn_0", "ufn"), ("vfn_0", "vfn"), ("wfn_0", "wfn"), ("pfn_0", "pfn"), ("ufv_0", "ufv"),
                   ("vfv_0", "vfv"), ("wfv_0", "wfv"), ("pfv_0", "pfv"), ("U1FACE_0", "U1"), ("U2FACE_0", "U2"),
                   ("U3FACE_0", "U3"), ("UTFACE_0", "UT"), ("UBFACE_0", "UB")):
        r[nm] = O.state(st).copy()

    def st(n):
        return O.state(n).copy(order="F")

    u, v, w, p = st("ufn"), st("vfn"), st("wfn"), st("pfn")
    uv, vv, wv = O.vert(), O.vert(), O.vert()
    O.call("bcglobalvc", u, v, w, uv, vv, wv, tagdir=2)
    r["vc2_u"], r["vc2_v"], r["vc2_w"] = uv.copy(), vv.copy(), wv.copy()
    pv = O.vert()
    O.call("bcglobalp", p, pv, tagdir=2)
    r["pv2"] = pv.copy()
    pv3 = O.vert()
    O.call("bcglobalp", p, pv3, tagdir=3)
    r["pv3"] = pv3
    gx, gy, gs = O.cell(), O.cell(), O.cell()
    O.call("grad_lsm", gx, gy, gs, p)
    r["lsm_p_x"], r["lsm_p_y"], r["lsm_p_s"] = gx.copy(), gy.copy(), gs.copy()
    O.call("grad_lsm", gx, gy, gs, u)
    r["lsm_u_x"], r["lsm_u_y"], r["lsm_u_s"] = gx.copy(), gy.copy(), gs.copy()
    f = [O.face() for _ in range(5)]
    O.call("calcul_face", *f, u, uv)
    for q, s in enumerate("123TB"):
        r["face_u_" + s] = f[q]
    O.call("sgles", st("ufn"), st("vfn"), st("wfn"))
    r["eddy_0"] = O.state("eddy").copy()
    prm = c.params
    ru, rv, rw = O.cell() + prm.bdx, O.cell() + prm.bdy, O.cell() + prm.bdz
    O.call("advdiff", ru, rv, rw, st("ufn"), st("vfn"), st("wfn"), st("eddy"))
    r["rhsu"], r["rhsv"], r["rhsw"] = ru, rv, rw
    src = O.cell()
    O.call("source3d", src, st("ufn"), st("vfn"), st("wfn"))
    r["src"] = src.copy()
    for s, nm in zip("123TB", ("U1", "U2", "U3", "UT", "UB")):
        r["src_U" + s] = O.state(nm).copy()
    src[:c.N_CELL0, 1:-1] /= prm.dt
    O.call("update_rhsp", src, st("pfn"), O.vert())
    r["rhsp"] = src.copy()
    p6, pv6 = st("pfn"), O.vert()
    O.call("bicgstab", p6, pv6, src)
    r["solves"] = np.array([O.solver_record()["iter"]], dtype=np.int32)
    r["psol"], r["psol_v"] = p6.copy(), pv6.copy()
    O.call("grad_lsm", gx, gy, gs, p6)
    r["gp_x"], r["gp_y"], r["gp_s"] = gx.copy(), gy.copy(), gs.copy()
    O.call("newu3d", p6, pv6)
    for s, nm in zip("123TB", ("U1", "U2", "U3", "UT", "UB")):
        r["newu_U" + s] = O.state(nm).copy()
    r["newu_pv"] = pv6
    return r


def test_oracle_routines_bitwise():
    """Every hot-path routine of the oracle == the reference's, bit for bit."""
    c, p0 = tiny_case()
    g = load_tiny()
    O = oracle.Oracle(c, p0=p0)
    r = run_unit_sequence(O, c, p0)
    assert r.keys() <= g.keys()
    for k, mine in r.items():
        ref = g[k]
        n = ref.shape[0]
        assert np.array_equal(ref, np.asarray(mine)[:n]), k


@pytest.mark.parametrize("cfg", [0, 3])
def test_oracle_trajectory_bitwise(cfg):
    """Whole RK2 steps (hydro x2, average, update) == the reference's, hash
    for hash; with glob_sample after every step, the twelve accumulators of
    the channel statistics too, and the three files glob_put_plane_mm wrote
    come out byte for byte from the library's writer fed the oracle's plane
    sums."""
    import tempfile

    from nsx._lib import device
    rec = load_steps(cfg)
    c = config_case(cfg)
    O = oracle.Oracle(c)
    O.sm_init()
    for s in range(1, rec["nsteps"] + 1):
        before = O.solver_record()["n_solves"]
        O.step()
        assert O.solver_record()["n_solves"] == before + 2
        for nm, want in rec["steps"][str(s)].items():
            assert digest(c, O.state(STATE_MAP[nm])) == want["sha256"], (s, nm)
        O.glob_sample()
    assert O.solver_record()["iter_total"] == sum(rec["solves"])
    assert O.solver_record()["iter"] == rec["solves"][-1]
    st = rec["stats"]
    for nm, want in st["accumulators"].items():
        a = np.ascontiguousarray(O.state(nm)[:c.N_CELL0], dtype=np.float64) + 0.0
        assert hashlib.sha256(a.tobytes()).hexdigest() == want["sha256"], nm
    planes, ns = O.sm_planes()
    assert ns == st["sm_n"]
    lib = device()
    with tempfile.TemporaryDirectory() as td:
        assert lib.nsx_write_plane_stats(np.ascontiguousarray(planes).ctypes.data, c.NZ, c.N_CELL0, ns, c.params.Re,
                                         td.encode()) == 0
        for fn, text in st["files"].items():
            assert open(os.path.join(td, fn)).read() == text, fn
    check_snapshots(lib, c, rec["serial"], {n: O.state(n) for n in ("ufv", "vfv", "wfv", "pfv", "ufnp", "vfnp",
                                                                    "wfnp", "pfnp")})


def check_snapshots(lib, c, serial, f):
    """The snapshots SavetecVertex, SavetecCenter, SavetecVC (Tecplot) and
    SaveParaviewVertex (VTK) wrote of the final state == the library's
    writers fed the same fields, byte for byte"""
    import tempfile

    xvt, yvt, zvt = vertex_coordinates(c)
    xct, yct, zct = cell_coordinates(c)
    for nm, a in (("xvt", xvt), ("yvt", yvt), ("zvt", zvt), ("xct", xct), ("yct", yct), ("zct", zct)):
        assert hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest() == serial["coords"][nm], nm
    P = lambda a: np.asfortranarray(a, dtype=np.float64).ctypes.data  # noqa: E731
    no_vp = np.asfortranarray(c.arrays["No_vp"], dtype=np.int32)
    keep = [f[n] for n in f] + [xvt, yvt, zvt, xct, yct, zct, no_vp]  # noqa: F841  (alive across the calls)
    with tempfile.TemporaryDirectory() as td:
        assert lib.nsx_write_tec_vertex(td.encode(), serial["counter"], serial["time"], c.N_CELL0, c.N_VERT, c.NZ,
                                        no_vp.ctypes.data, P(xvt), P(yvt), P(zvt), P(f["ufv"]), P(f["vfv"]),
                                        P(f["wfv"]), P(f["pfv"])) == 0
        assert lib.nsx_write_tec_center(td.encode(), serial["counter"], serial["time"], c.N_CELL0, c.N_CELL,
                                        c.N_VERT, c.NZ, no_vp.ctypes.data, P(xvt), P(yvt), P(zvt), P(f["ufnp"]),
                                        P(f["vfnp"]), P(f["wfnp"]), P(f["pfnp"])) == 0
        assert lib.nsx_write_vtk_vertex(td.encode(), serial["counter"], c.N_CELL0, c.N_VERT, c.NZ, no_vp.ctypes.data,
                                        P(xvt), P(yvt), P(zvt), P(f["ufv"]), P(f["vfv"]), P(f["wfv"]), None, None,
                                        None) == 0
        assert lib.nsx_write_tec_vc(td.encode(), serial["counter"], serial["time"], c.N_CELL0, c.N_CELL, c.N_VERT,
                                    c.NZ, no_vp.ctypes.data, P(xvt), P(yvt), P(zvt), P(xct), P(yct), P(zct),
                                    P(f["ufv"]), P(f["vfv"]), P(f["wfv"]), P(f["pfv"]), P(f["ufnp"]), P(f["vfnp"]),
                                    P(f["wfnp"]), P(f["pfnp"])) == 0
        assert sorted(os.listdir(td)) == sorted(serial["files"])
        for fn, want in serial["files"].items():
            data = open(os.path.join(td, fn), "rb").read()
            lines = data.decode().split("\n")
            assert lines[:5] == want["head"], fn
            assert (len(data), len(lines) - 1) == (want["size"], want["lines"]), fn
            assert hashlib.sha256(data).hexdigest() == want["sha256"], fn


def test_list_directed_real_format():
    """The library's list-directed REAL(8) formatting (the "TKE =" line of
    glob_put_plane_mm) against what amdflang, the compiler of the reference
    build, printed for 394 values (tests/golden/make_list_directed.py)."""
    import ctypes
    import json

    from nsx._lib import device
    with open(os.path.join(os.path.dirname(__file__), "golden", "list_directed.json")) as f:
        pairs = json.load(f)["pairs"]
    buf = ctypes.create_string_buffer(64)
    lib = device()
    bad = []
    for v, want in pairs:
        assert lib.nsx_format_real(float(v), 1, 0, 0, buf, 64) == 0
        if buf.value.decode() != want:
            bad.append((v, want, buf.value.decode()))
    assert not bad, bad[:5]


def test_e_edit_format():
    """The library's Ew.d output (E12.5, the data lines of glob_put_plane_mm)
    against what amdflang printed for the same values."""
    import ctypes
    import json

    from nsx._lib import device
    with open(os.path.join(os.path.dirname(__file__), "golden", "list_directed.json")) as f:
        pairs = json.load(f)["e12.5"]
    buf = ctypes.create_string_buffer(64)
    lib = device()
    bad = []
    for v, want in pairs:
        assert lib.nsx_format_real(float(v), 0, 12, 5, buf, 64) == 0
        if buf.value.decode() != want:
            bad.append((v, want, buf.value.decode()))
    assert not bad, bad[:5]


# ---------------------------------------------------------- closed forms ----

def _linear(c, a, b, e):
    f = c.cell_field()
    ne = c.N_CELLexact
    f[:ne] = a * c.xc[:ne, None] + b * c.yc[:ne, None] + e * c.sig[None, :]
    return f


def test_lsm_gradient_exact_for_linear_fields():
    """grandientLSM is exact for f = a x + b y (+ any function of sigma) in the interior."""
    c = make_case(8, 6, 5, seed=3, jitter=0.1)
    O = oracle.Oracle(c)
    # avoid the periodic seam: cells whose stencil does not wrap
    f = _linear(c, 0.7, -1.3, 0.0)
    gx, gy, gs = O.cell(), O.cell(), O.cell()
    O.call("grad_lsm", gx, gy, gs, f)
    xc = c.xc
    inner = [i for i in range(c.N_CELL0) if c.nbe[i] == 0 and
             all(abs(xc[c.No_cp[i, j] - 1] - xc[i]) < 1.0 for j in range(3))]
    assert len(inner) > 10
    np.testing.assert_allclose(gx[inner, 1:-1], 0.7, rtol=0, atol=1e-12)
    np.testing.assert_allclose(gy[inner, 1:-1], -1.3, rtol=0, atol=1e-12)


def test_vertical_gradient_exact_for_quadratic_on_uniform_levels():
    c = make_case(4, 4, 6, seed=1, stretch=0.0)
    O = oracle.Oracle(c)
    f = c.cell_field()
    f[:c.N_CELLexact] = 3.0 * c.sig[None, :] - 2.0
    gx, gy, gs = O.cell(), O.cell(), O.cell()
    O.call("grad_lsm", gx, gy, gs, f)
    np.testing.assert_allclose(gs[:c.N_CELL0, 1:-1], 3.0, atol=1e-12)


def test_idw_reproduces_constants():
    c = make_case(6, 5, 4, seed=5, jitter=0.1)
    O = oracle.Oracle(c)
    p = c.cell_field() + 2.5
    pv = O.vert()
    O.call("bcglobalp", p, pv, tagdir=2)
    np.testing.assert_allclose(pv, 2.5, rtol=1e-15)


def test_ppe_rows_sum_to_zero_and_manufactured_solve():
    """PDE_center rows sum to zero; BICGSTAB_SOLVE recovers x from b = A x."""
    c = make_case(6, 4, 6, seed=2, jitter=0.1)
    A = [c.AmC0, c.AmC1, c.AmC2, c.AmC3, c.AmCT, c.AmCB]
    rows = sum(a[:, 1:-1] for a in A)
    assert np.abs(rows).max() < 1e-12 * np.abs(c.AmC0[:, 1:-1]).max()
    # manufactured zero-mean solution
    rng = np.random.default_rng(0)
    x = c.cell_field()
    x[:c.N_CELL0, 1:-1] = rng.standard_normal((c.N_CELL0, c.NZ - 2))
    O = oracle.Oracle(c)
    O.call("bcpcenter", x)
    b = c.cell_field()
    nz = c.NZ
    cp = c.No_cp[:c.N_CELL0] - 1
    part = [x[cp[:, j]] for j in range(3)]
    for k in range(1, nz - 1):
        b[:c.N_CELL0, k] = (A[0][:, k] * x[:c.N_CELL0, k] + A[1][:, k] * part[0][:, k] + A[2][:, k] * part[1][:, k] +
                            A[3][:, k] * part[2][:, k] + A[4][:, k] * x[:c.N_CELL0, k + 1] +
                            A[5][:, k] * x[:c.N_CELL0, k - 1])
    phi = c.cell_field()
    pv = O.vert()
    O.call("bicgstab", phi, pv, b.copy(order="F"))
    # the solution is fixed up to a constant; the solver removes the volume mean
    vol = c.areaCell[:c.N_CELL0, None] * c.dsig[None, 1:-1]
    xm = x[:c.N_CELL0, 1:-1] - (x[:c.N_CELL0, 1:-1] * vol).sum() / (c.TotalVolume * 2.0)
    np.testing.assert_allclose(phi[:c.N_CELL0, 1:-1], xm, atol=1e-7 * np.abs(xm).max())


@pytest.mark.parametrize("cfg", [0, 3])
def test_glob_cfl_vs_reference(cfg):
    """glob_cfl (cfl.F90:84-200) on the final state of the recorded steps: the
    values the compiled reference printed (17 significant digits, so equal as
    doubles) and the means it wrote to meanuv.dat (f12.5)."""
    rec = load_steps(cfg)
    c = config_case(cfg)
    O = oracle.Oracle(c)
    for _ in range(rec["nsteps"]):
        O.step()
    got = O.glob_cfl()
    ref = rec["glob_cfl"]
    for k in ("umax", "vmax", "wmax", "cfl", "flux_total", "flux_max", "flux_bc", "flux_bot", "flux_top",
              "flux_side"):
        assert got[k] == ref[k], (k, got[k], ref[k])
    for k in ("aveu", "avev", "avew"):
        assert abs(got[k] - ref[k]) <= 0.5e-5 + 1e-12, (k, got[k], ref[k])


def test_glob_cfl_closed_form():
    """Uniform velocity u = 1, v = w = 0 with face velocities of a uniform
    horizontal flow: the means are (1, 0, 0), every cell's net flux is the
    sum of its lateral face fluxes and the CFL number follows from the face
    sums, computed here independently of the oracle."""
    c = make_case(8, 6, 6, seed=3, jitter=0.1)
    O = oracle.Oracle(c)
    n0, nz = c.N_CELL0, c.NZ
    u = c.cell_field() + 1.0
    zero = c.cell_field()
    xnn = c.arrays["xnn"][:n0]
    dl = c.arrays["dlVV"][:n0]
    for j, name in enumerate(("U1", "U2", "U3")):
        O.state(name)[:, :] = xnn[:, j:j + 1]            # face-normal component of (1, 0)
    O.state("UT")[:, :] = 0.0
    O.state("UB")[:, :] = 0.0
    got = O.glob_cfl(u, zero, zero)
    assert got["umax"] == 1.0 and got["vmax"] == 0.0 and got["wmax"] == 0.0
    assert abs(got["aveu"] - 1.0) < 1e-13 and got["avev"] == 0.0
    # closed polygon: sum_j n_j l_j = 0, so every cell's flux vanishes to round-off
    assert abs(got["flux_max"]) < 1e-13 and abs(got["flux_total"]) < 1e-11
    area = c.areaCell[:n0]
    dsig = c.dsig
    cfl = max(0.5 * c.params.dt * np.abs(xnn * dl * dsig[k]).sum(axis=1).max() / (area * dsig[k]).min()
              for k in range(1, nz - 1))
    assert got["cfl"] <= cfl * (1 + 1e-12)
    per_cell = [0.5 * c.params.dt * (np.abs(xnn * dl).sum(axis=1) * dsig[k]) / (area * dsig[k])
                for k in range(1, nz - 1)]
    assert abs(got["cfl"] - max(p.max() for p in per_cell)) <= 1e-14 * got["cfl"]


def test_parallel_oracle_build_matches_serial():
    """The OpenMP build of the oracle (all-cores CPU baseline only) solves the
    same pressure system as the serial checker: its sums run in another
    order, so the iteration count may move a little and the solution agrees
    to the solver's tolerance."""
    import oracle as orc
    c = config_case(0, jitter=0.1)
    rng = np.random.default_rng(3)
    rhs = c.cell_field()
    rhs[:c.N_CELL0, 1:-1] = rng.standard_normal((c.N_CELL0, c.NZ - 2))
    sols, its = [], []
    for par in (False, True):
        O = orc.Oracle(c, parallel=par)
        phi = c.cell_field()
        its.append(O.call("bicgstab", phi, O.vert(), rhs.copy(order="F")))
        sols.append(phi[:c.N_CELL0, 1:-1].copy())
        if par:
            assert O.threads() >= 1
    assert abs(its[1] - its[0]) <= 0.1 * its[0], its
    a, b = sols
    a = a - a.mean()
    b = b - b.mean()
    assert np.linalg.norm(a - b) <= 1e-6 * np.linalg.norm(a)
