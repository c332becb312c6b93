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


def interior(case, a):
    """Interior cells (i <= N_CELL0, k = 2..NZ-1) of an (N_CELL,NZ) array."""
    return np.asarray(a)[:case.N_CELL0, 1:-1]


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
