"""The set-up path on the host: the reference's mesh file (data.txt,
input.F90:147-194) and calcul_geometry + adjust_bc + PDE_center on a given
mesh (nsx_case_from_mesh).

* The reader takes every field as the reference build's compiler does
  (tests/golden/mesh_fields.json: a crafted file with blanks, short records,
  implied decimal points and exponents, read by amdflang).
* A jittered mesh written as data.txt and read back is set up exactly as the
  reference sets it up from the same file through its own input.F90
  (tests/golden/mesh_setup.json: SHA-256 of every set-up array).
* On the synthetic meshes the set-up from arrays equals the generator's,
  which the geometry tests pin to the reference.
* The cached binary set-up (nsx_case_save / nsx_case_load) gives back every
  array bit for bit and refuses a truncated, corrupt or foreign file.
"""
import hashlib
import json
import os
import struct

import numpy as np
import pytest

from conftest import GOLDEN
from nsx.case import case_from_mesh, config_case, load_case, make_case, read_mesh, write_mesh

MESH_KEYS = ("No_vp", "xv", "yv", "No_cp", "zbv", "nbe", "nbev")


def _bits(a):
    return [struct.pack(">d", float(x)).hex().upper() for x in a]


def test_mesh_fields_as_the_compiler_reads_them(tmp_path):
    d = json.load(open(os.path.join(GOLDEN, "mesh_fields.json")))
    p = tmp_path / "data.txt"
    p.write_text(d["text"])
    m = read_mesh(str(p), d["N_CELL0"], d["N_VERT"])
    assert list(m["No_vp"].ravel(order="F")) == d["No_vp"]
    assert list(m["No_cp"].ravel(order="F")) == d["No_cp"]
    assert list(m["nbe"]) == d["nbe"] and list(m["nbev"]) == d["nbev"]
    for k in ("xv", "yv", "zbv"):
        assert _bits(m[k]) == d[k + "_bits"], k


def test_mesh_file_round_trip(tmp_path):
    """Connectivity and flags come back as written; coordinates as the F15.5
    and F8.3 fields round them (the nearest double to the printed decimal)."""
    c = config_case(0, jitter=0.1, bed=True)
    p = str(tmp_path / "data.txt")
    write_mesh(p, c)
    m = read_mesh(p, c.N_CELL0, c.N_VERT)
    assert np.array_equal(m["No_vp"], c.No_vp)
    assert np.array_equal(m["No_cp"], c.No_cp[:c.N_CELL0])
    assert np.array_equal(m["nbe"], c.nbe) and np.array_equal(m["nbev"], c.nbev)
    assert np.array_equal(m["xv"], [float("%.5f" % x) for x in c.xv])
    assert np.array_equal(m["yv"], [float("%.5f" % y) for y in c.yv])
    assert np.array_equal(m["zbv"], [float("%.3f" % z) for z in c.zbv])


def _views(arrays, N_CELL, N_CELLexact, dims):
    """the canonical parts hashed by tests/golden/make_golden.py"""
    out = {}
    for k, a in arrays.items():
        a = np.asarray(a)
        if k.startswith("AmC"):
            v = a[:, 1:-1]
        elif k in ("surrounding", "weight"):
            v = a.copy()
            for j in range(v.shape[1]):
                v[dims <= j, j] = 0
        elif a.ndim >= 1 and a.shape[0] == N_CELL:
            v = a[:N_CELLexact]
        else:
            v = a
        v = np.ascontiguousarray(v)
        out[k] = hashlib.sha256((v + 0.0 if v.dtype.kind == "f" else v).tobytes()).hexdigest()
    return out


def test_mesh_file_setup_matches_reference(tmp_path):
    g = json.load(open(os.path.join(GOLDEN, "mesh_setup.json")))
    m = g["case"]
    c = make_case(m["nx"], m["ny"], m["nn"], seed=m["seed"], jitter=m["jitter"], Lx=m["Lx"], Ly=m["Ly"])
    p = str(tmp_path / "data.txt")
    write_mesh(p, c)
    assert hashlib.sha256(open(p, "rb").read()).hexdigest() == g["data_txt_sha256"]
    mesh = read_mesh(p, c.N_CELL0, c.N_VERT)
    for k, want in g["mesh_read"].items():
        a = np.ascontiguousarray(mesh[k])
        assert hashlib.sha256((a + 0.0 if a.dtype.kind == "f" else a).tobytes()).hexdigest() == want, k
    s = case_from_mesh(mesh, m["nn"], box=(0.0, m["Lx"], 0.0, m["Ly"]))
    assert [s.N_CELLexact, s.N_CELLghost, s.bctop, s.bcbot] == g["sizes"]
    assert s.TotalVolume == g["TotalVolume"]
    got = _views({k: s.arrays[k] for k in g["arrays"]}, s.N_CELL, s.N_CELLexact, s.Dimsurrounding)
    assert [k for k in g["arrays"] if got[k] != g["arrays"][k]] == []


@pytest.mark.parametrize("kw", [dict(cfg=0), dict(cfg=0, jitter=0.1, bed=True), dict(cfg=3)])
def test_setup_from_arrays_equals_generator(kw):
    cfg = kw.pop("cfg")
    c = config_case(cfg, **kw)
    mesh = {k: (c.No_cp[:c.N_CELL0] if k == "No_cp" else c.arrays[k]) for k in MESH_KEYS}
    s = case_from_mesh(mesh, c.nn, Ubulk=15.6, seed=1234)
    assert (s.N_CELLexact, s.N_CELLghost, s.TotalVolume) == (c.N_CELLexact, c.N_CELLghost, c.TotalVolume)
    for k, a in c.arrays.items():
        if k == "grid":  # the generator's nx, ny: a mesh file has none
            continue
        assert np.array_equal(s.arrays[k], a), k


def test_mesh_errors(tmp_path):
    c = config_case(0)
    with pytest.raises(OSError):
        read_mesh(str(tmp_path / "missing.txt"), c.N_CELL0, c.N_VERT)
    p = tmp_path / "data.txt"
    write_mesh(str(p), c)
    lines = p.read_text().splitlines(keepends=True)
    short = tmp_path / "short.txt"
    short.write_text("".join(lines[:100]))
    with pytest.raises(ValueError, match="record"):
        read_mesh(str(short), c.N_CELL0, c.N_VERT)
    mesh = read_mesh(str(p), c.N_CELL0, c.N_VERT)
    bad = dict(mesh, No_vp=mesh["No_vp"].copy())
    bad["No_vp"][3, 1] = c.N_VERT + 1
    with pytest.raises(RuntimeError, match="not a vertex"):
        case_from_mesh(bad, c.nn)
    # the box of the file's rounded coordinates differs from 2*pi x pi: the
    # periodic partners are not found, as in the reference
    with pytest.raises(RuntimeError, match="periodic matching failed"):
        case_from_mesh(mesh, c.nn)


@pytest.mark.parametrize("walls", [None, (1, 0, 0, 0)])
def test_cached_setup_round_trip(tmp_path, walls):
    """A set-up written to its cache file and read back: the same sizes,
    volume and every array, bit for bit (the periodic channel and a duct)."""
    path = str(tmp_path / "case.bin")
    c = make_case(24, 12, 16, seed=9, jitter=0.1, walls=walls, save=path)
    d = load_case(path, params=c.params)
    for k in ("nx", "ny", "nn", "N_CELL0", "N_CELL", "N_VERT", "NZ", "N_CELLexact", "N_CELLghost", "bctop",
              "bcbot", "TotalVolume"):
        assert getattr(d, k) == getattr(c, k), k
    assert sorted(d.arrays) == sorted(c.arrays)
    for k in c.arrays:
        assert c.arrays[k].dtype == d.arrays[k].dtype and np.array_equal(c.arrays[k], d.arrays[k]), k


def test_cached_setup_refuses_damaged_files(tmp_path):
    path = str(tmp_path / "case.bin")
    make_case(8, 6, 4, seed=3, save=path)
    good = open(path, "rb").read()
    cases = {"truncated": good[:len(good) // 2], "corrupt": good[:200] + bytes([good[200] ^ 1]) + good[201:],
             "foreign": b"NOTACASE" + good[8:]}
    for what, data in cases.items():
        p = str(tmp_path / (what + ".bin"))
        with open(p, "wb") as f:
            f.write(data)
        with pytest.raises(RuntimeError):
            load_case(p)
    with pytest.raises(RuntimeError):
        load_case(str(tmp_path / "missing.bin"))
