"""Channel statistics (stats.F90) on the device against the oracle and the
reference's own files.

* glob_sample on the oracle's states (u, v, w, p and the eddy viscosity put
  on the device): the twelve accumulators and the plane sums of
  glob_put_plane_mm are bit-identical to the oracle's (element-wise sums in
  the reference's order; plane sums in its sequential order), and the three
  files are byte-identical to the ones the library's writer makes from the
  oracle's plane sums.
* glob_sample after every one of ten device RK2 steps of the 32x16 quads x 32
  layers channel: the files hold the reference's values (its own files,
  tests/golden) to 1e-6 relative or 1e-8 absolute, the spread of two
  trajectories whose pressure solves stop at a 1e-10 residual.
* the snapshots (SavetecVertex, SavetecCenter, SavetecVC, SaveParaviewVertex)
  of resident fields: byte-identical to the reference's files on the same
  state.
"""
import os
import tempfile

import numpy as np
import pytest

import oracle
from common import cell_coordinates, load_steps, vertex_coordinates
from nsx.abi import SM_NAMES
from nsx.case import config_case

pytestmark = pytest.mark.gpu


def _ctx(case):
    from nsx.device import Context
    return Context(case)


def test_glob_sample_bitwise_vs_oracle():
    c = config_case(0)
    O = oracle.Oracle(c)
    G = _ctx(c)
    G.init_state()
    O.sm_init()
    G.sm_init()
    for _ in range(3):
        O.step()
        for f, s in (("UFN", "ufn"), ("VFN", "vfn"), ("WFN", "wfn"), ("PFN", "pfn"), ("EDDY", "eddy")):
            G.put(f, O.state(s))
        O.glob_sample()
        G.glob_sample()
    assert G.sm_count() == 3
    n0 = c.N_CELL0
    for nm in SM_NAMES:
        dev, ref = G.sm_get(nm)[:n0], O.state(nm)[:n0]
        assert np.array_equal(dev, ref), nm
    pd, nsd = G.sm_planes()
    po, nso = O.sm_planes()
    assert nsd == nso == 3
    assert np.array_equal(pd, po)
    from nsx._lib import device
    with tempfile.TemporaryDirectory() as a, tempfile.TemporaryDirectory() as b:
        G.glob_put_plane_mm(a)
        assert device().nsx_write_plane_stats(np.ascontiguousarray(po).ctypes.data, c.NZ, n0, nso, c.params.Re,
                                              b.encode()) == 0
        names = sorted(os.listdir(a))
        assert names == ["sm-00003.dat", "srms-00003.dat", "stau-00003.dat"]
        for fn in names:
            assert open(os.path.join(a, fn)).read() == open(os.path.join(b, fn)).read(), fn
    G.close()


def _numbers(text):
    out = []
    for ln in text.splitlines():
        ln = ln.replace("TKE =", "")
        for tok in ln.split():
            t = tok.replace("E", "e")
            if t.startswith(".") or t.startswith("-."):
                t = t.replace(".", "0.", 1)
            out.append(float(t))
    return np.array(out)


def test_stats_over_device_steps_vs_reference_files():
    rec = load_steps(0)
    c = config_case(0)
    G = _ctx(c)
    G.init_state()
    G.sm_init()
    for _ in range(rec["nsteps"]):
        G.step(1, 0)
        G.glob_sample()
    with tempfile.TemporaryDirectory() as td:
        G.glob_put_plane_mm(td)
        for fn, want in rec["stats"]["files"].items():
            got = open(os.path.join(td, fn)).read()
            a, b = _numbers(got), _numbers(want)
            assert a.shape == b.shape, fn
            assert np.all(np.abs(a - b) <= 1e-6 * np.abs(b) + 1e-8), (fn, np.max(np.abs(a - b)))
    G.close()


def _snapshot_numbers(path):
    """the numbers of a Tecplot snapshot after its header lines (E12.5 blocks
    then the integer brick connectivity)"""
    with open(path) as f:
        lines = f.read().splitlines()
    body = [ln for ln in lines if not any(k in ln for k in ("TITLE", "VARIABLES", "ZONE", "VARLOCATION", "StrandID"))]
    return np.array([float(t) for ln in body for t in ln.split()])


def test_snapshots_of_resident_fields_match_reference_files():
    """nsx_savetec_vertex / nsx_savetec_center download the resident fields:
    fed the oracle's final 10-step state (== the reference's, hash for hash),
    the files are the reference's byte for byte; after ten device steps they
    hold the oracle's values up to one unit in the last printed digit plus
    1e-6 of each field's largest value (the spread of two trajectories whose
    solves stop at a 1e-10 residual), with the same connectivity."""
    import hashlib

    rec = load_steps(0)
    serial = rec["serial"]
    c = config_case(0)
    O = oracle.Oracle(c)
    for _ in range(rec["nsteps"]):
        O.step()
    xvt, yvt, zvt = vertex_coordinates(c)
    xct, yct, zct = cell_coordinates(c)
    G = _ctx(c)
    G.init_state()
    for f in ("UFV", "VFV", "WFV", "PFV", "UFNP", "VFNP", "WFNP", "PFNP"):
        G.put(f, O.state(f.lower()))

    def write_all(ctx, td):
        ctx.savetec_vertex(td, serial["counter"], serial["time"], xvt, yvt, zvt)
        ctx.savetec_center(td, serial["counter"], serial["time"], xvt, yvt, zvt)
        ctx.save_paraview_vertex(td, serial["counter"], xvt, yvt, zvt)
        ctx.savetec_vc(td, serial["counter"], serial["time"], xvt, yvt, zvt, xct, yct, zct)

    with tempfile.TemporaryDirectory() as td:
        write_all(G, td)
        assert sorted(os.listdir(td)) == sorted(serial["files"])
        for fn, want in serial["files"].items():
            with open(os.path.join(td, fn), "rb") as f:
                assert hashlib.sha256(f.read()).hexdigest() == want["sha256"], fn
        # the numeric comparison after device steps covers the two Tecplot
        # files whose blocks are one field each
        ref = {fn: _snapshot_numbers(os.path.join(td, fn)) for fn in ("V-0003.tec", "C-00003.tec")}
    G2 = _ctx(c)
    G2.init_state()
    for _ in range(rec["nsteps"]):
        G2.step(1, 0)
    with tempfile.TemporaryDirectory() as td:
        write_all(G2, td)
        for fn, b in ref.items():
            a = _snapshot_numbers(os.path.join(td, fn))
            assert a.shape == b.shape, fn
            # field blocks (coordinates and fields), then the connectivity
            nv, ncl = c.N_VERT * (c.NZ - 1), c.N_CELL0 * (c.NZ - 2)
            sizes = [nv] * 7 if fn.startswith("V") else [nv] * 3 + [ncl] * 4
            o = 0
            for q, n in enumerate(sizes):
                x, y = a[o:o + n], b[o:o + n]
                o += n
                # one unit in the fifth (last printed) digit of E12.5, plus
                # the spread of the two trajectories, 1e-6 of the field's scale
                ulp = 10.0 ** (np.floor(np.log10(np.maximum(np.abs(y), 1e-300))) - 4)
                bad = np.abs(x - y) > 1.5 * ulp + 1e-6 * np.abs(y).max()
                assert not bad.any(), (fn, q, int(bad.sum()), x[bad][:4], y[bad][:4])
            assert np.array_equal(a[o:], b[o:]), fn
    G.close()
    G2.close()


def test_sample_needs_init():
    c = config_case(0)
    G = _ctx(c)
    with pytest.raises(RuntimeError):
        G.glob_sample()
    G.close()
