"""Shared helpers of the test-suite."""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np

from conftest import GOLDEN
from nsx.case import config_case, make_case

TINY = dict(nx=6, ny=4, nn=4, seed=7, jitter=0.1, bed=True)
STATE_MAP = {"ufn": "ufn", "vfn": "vfn", "wfn": "wfn", "pfn": "pfn", "ufv": "ufv", "vfv": "vfv", "wfv": "wfv",
             "pfv": "pfv", "U1FACE": "U1", "U2FACE": "U2", "U3FACE": "U3", "UTFACE": "UT", "UBFACE": "UB",
             "rhsuf": "rhsuf", "rhsvf": "rhsvf", "rhswf": "rhswf", "eddy": "eddy"}


def tiny_case():
    c = make_case(TINY["nx"], TINY["ny"], TINY["nn"], seed=TINY["seed"], jitter=TINY["jitter"], bed=TINY["bed"])
    rng = np.random.default_rng(3)
    p0 = np.zeros((c.N_CELL, c.NZ), order="F")
    p0[:c.N_CELL0, 1:-1] = rng.uniform(-1.0, 1.0, (c.N_CELL0, c.NZ - 2))
    return c, p0


def load_tiny():
    return dict(np.load(os.path.join(GOLDEN, "tiny_unit.npz")))


def load_steps(cfg: int):
    with open(os.path.join(GOLDEN, f"cfg{cfg}_steps.json")) as f:
        return json.load(f)


def rows_of(case, a):
    return case.N_CELLexact if a.shape[0] == case.N_CELL else a.shape[0]


def digest(case, a) -> str:
    b = np.ascontiguousarray(np.asarray(a)[:rows_of(case, a)], dtype=np.float64) + 0.0
    return hashlib.sha256(b.tobytes()).hexdigest()


def sha(a) -> str:
    a = np.ascontiguousarray(a)
    if a.dtype.kind == "f":
        a = a + 0.0
    return hashlib.sha256(a.tobytes()).hexdigest()


def setup_views(dump_or_arrays, N_CELL, N_CELLexact, Dimsurrounding):
    """the canonical part of each set-up array: rows 1..N_CELLexact of cell
    arrays, the interior layers of the PPE matrix, the used entries of the
    inverse-distance stencil"""
    out = {}
    for k, a in dump_or_arrays.items():
        a = np.asarray(a)
        if k.startswith("AmC"):
            v = a[:, 1:-1]
        elif k in ("surrounding", "weight"):
            v = a.copy()
            for j in range(v.shape[1]):
                v[Dimsurrounding <= j, j] = 0
        elif a.ndim >= 1 and a.shape[0] == N_CELL:
            v = a[:N_CELLexact]
        else:
            v = a
        out[k] = v
    return out


def load_bc():
    """the boundary-condition fixtures (tests/golden/bc_cases.json)"""
    with open(os.path.join(GOLDEN, "bc_cases.json")) as f:
        return json.load(f)


def bc_case(rec, name: str, which: str):
    """the case of boundary set `name` on the tiny mesh (which = "tiny",
    with its pressure field) or the 3-step mesh (which = "steps")"""
    from nsx.case import Params
    e = rec["sets"][name]
    m = rec["tiny"] if which == "tiny" else rec["steps_case"]
    c = make_case(m["nx"], m["ny"], m["nn"], seed=m["seed"], jitter=m["jitter"], bed=m.get("bed", False),
                  walls=tuple(e["walls"]), params=Params(TopBC=e["TopBC"]))
    if which != "tiny":
        return c
    rng = np.random.default_rng(rec["tiny_p_seed"])
    p0 = np.zeros((c.N_CELL, c.NZ), order="F")
    p0[:c.N_CELL0, 1:-1] = rng.uniform(-1.0, 1.0, (c.N_CELL0, c.NZ - 2))
    return c, p0


def interior(case, a):
    """Interior cells (i <= N_CELL0, k = 2..NZ-1) of an (N_CELL,NZ) array."""
    return np.asarray(a)[:case.N_CELL0, 1:-1]


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def ppe_residual(case, phi, rhs):
    """Relative norms (e1, e2) of b - A phi over the interior, with A the
    stored PDE_center coefficients AmC0..AmCB and b the mean-removed rhs of
    Glob_Norm(rhs,1) (bicgstab.F90:635-663).  Periodic neighbours are read
    through the ghost rows' partner (No_cp(g,3)); the wall-neighbour
    coefficients are zero, so the vertical ghost layers do not enter."""
    A = case.arrays
    n0, nz = case.N_CELL0, case.NZ
    phi = np.asarray(phi)
    ncp = A["No_cp"].astype(np.int64) - 1
    col = np.arange(n0)
    nb = []
    for j in range(3):
        jc = ncp[:n0, j].copy()
        gh = jc >= n0
        jc[gh] = ncp[jc[gh], 2]
        nb.append(jc)
    k = slice(1, nz - 1)
    Ax = (A["AmC0"][:, k] * phi[col, k] + A["AmC1"][:, k] * phi[nb[0], k] + A["AmC2"][:, k] * phi[nb[1], k]
          + A["AmC3"][:, k] * phi[nb[2], k] + A["AmCT"][:, k] * phi[col, 2:nz] + A["AmCB"][:, k] * phi[col, 0:nz - 2])
    vol = case.areaCell[:n0, None] * case.dsig[None, 1:-1]
    rhs_i = np.asarray(rhs)[:n0, 1:-1]
    avec = rhs_i.sum() / (case.TotalVolume * (case.params.sigFin - case.params.sigIni))
    b = rhs_i - avec * vol
    r = b - Ax
    return float(np.abs(r).sum() / np.abs(b).sum()), float(np.linalg.norm(r) / np.linalg.norm(b))


def vertex_coordinates(c):
    """xvt, yvt, zvt (N_VERT, NZ-1) as initial.F90:382-388 sets them, with
    the channel's Hprv = 1 and hv = 0"""
    A = c.arrays
    nk = c.NZ - 1
    xvt = np.asfortranarray(np.repeat(A["xv"][:, None], nk, axis=1))
    yvt = np.asfortranarray(np.repeat(A["yv"][:, None], nk, axis=1))
    zvt = np.asfortranarray(np.repeat(A["sigv"][None, :] * 1.0 - 0.0, c.N_VERT, axis=0))
    return xvt, yvt, zvt


def cell_coordinates(c):
    """xct, yct, zct (N_CELL, NZ) as initial.F90:372-378 sets them, with the
    channel's Hpr = 1 and h = 0"""
    A = c.arrays
    xct = np.asfortranarray(np.repeat(A["xc"][:, None], c.NZ, axis=1))
    yct = np.asfortranarray(np.repeat(A["yc"][:, None], c.NZ, axis=1))
    zct = np.asfortranarray(np.repeat(A["sig"][None, :] * 1.0 - 0.0, c.N_CELL, axis=0))
    return xct, yct, zct
