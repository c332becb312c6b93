"""Channel cases: seeded mesh, geometry, flags, PPE matrix and initial field.

Wraps the host set-up library (host/nsx_case.c), which restates the set-up
half of the reference (src/Main/input.F90's mesh file, src/Main/
calcul_geometry.F90, src/BoundaryConditions/update_bc.F90,
src/AdvectionDiffusion/PDE.F90) for a doubly periodic channel: the synthetic
seeded meshes, or a mesh given as arrays or as the reference's data.txt.  Every array keeps the reference's Fortran layout
(column-major, 1-based connectivity, ghost cells after N_CELL0) and name.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib

# The workloads (layer count NN, quads nx*ny, viscosity, dt, steps).  Box
# Lx=2*pi, Ly=pi, H=2, walls at both sigma ends, forcing bdx=1.
CONFIGS = {
    0: dict(nx=32, ny=16, nn=32, Re=180.0, dt=1e-3, steps=10, jitter=0.0, bed=False, Ubulk=15.6),
    1: dict(nx=128, ny=64, nn=64, Re=180.0, dt=5e-4, steps=100, jitter=0.1, bed=True, Ubulk=15.6),
    2: dict(nx=256, ny=128, nn=96, Re=395.0, dt=2.5e-4, steps=50, jitter=0.1, bed=True, Ubulk=17.5),
    3: dict(nx=64, ny=32, nn=48, Re=180.0, dt=5e-4, steps=20, jitter=0.1, bed=True, Ubulk=15.6),
    4: dict(nx=512, ny=256, nn=128, Re=590.0, dt=1e-4, steps=20, jitter=0.1, bed=True, Ubulk=18.6),
}

LX = 2.0 * math.pi
LY = math.pi
H = 2.0


@dataclass
class Params:
    """Run parameters (the reference reads them from dataParameters.txt)."""
    dt: float = 1e-3
    Re: float = 180.0
    bdx: float = 1.0
    bdy: float = 0.0
    bdz: float = 0.0
    eps: float = 1e-10
    MaxIters: int = 20000
    FinRK: int = 2
    Iturbu: int = 1
    TopBC: int = 0
    BotBC: int = 0
    sigIni: float = 0.0
    sigFin: float = H


@dataclass
class Case:
    nx: int
    ny: int
    nn: int
    N_CELL0: int
    N_CELL: int
    N_VERT: int
    NZ: int
    N_CELLexact: int
    N_CELLghost: int
    bctop: int
    bcbot: int
    TotalVolume: float
    params: Params
    arrays: dict = field(default_factory=dict)

    def __getattr__(self, name):
        arrays = self.__dict__.get("arrays")
        if arrays is not None and name in arrays:
            return arrays[name]
        raise AttributeError(name)

    # reference-layout shapes
    def cell_field(self):
        return np.zeros((self.N_CELL, self.NZ), dtype=np.float64, order="F")

    def vert_field(self):
        return np.zeros((self.N_VERT, self.NZ - 1), dtype=np.float64, order="F")

    @property
    def n_interior(self) -> int:
        return self.N_CELL0 * (self.NZ - 2)


_SHAPES = {
    # name: shape builder (Fortran order)
    "No_vp": lambda c: (c.N_CELL0, 3), "No_cp": lambda c: (c.N_CELL, 3),
    "surrounding": lambda c: (c.N_VERT, 10), "weight": lambda c: (c.N_VERT, 10),
    "AmC0": lambda c: (c.N_CELL0, c.NZ), "AmC1": lambda c: (c.N_CELL0, c.NZ),
    "AmC2": lambda c: (c.N_CELL0, c.NZ), "AmC3": lambda c: (c.N_CELL0, c.NZ),
    "AmCT": lambda c: (c.N_CELL0, c.NZ), "AmCB": lambda c: (c.N_CELL0, c.NZ),
    "u0": lambda c: (c.N_CELL, c.NZ), "v0": lambda c: (c.N_CELL, c.NZ),
    "w0": lambda c: (c.N_CELL, c.NZ), "p0": lambda c: (c.N_CELL, c.NZ),
}
_FACE3 = {"dxVV", "dyVV", "dlVV", "xnn", "ynn", "dxCC", "dyCC", "dlCC", "xme", "yme", "xmo",
          "ymo", "doe", "dle1", "dle2", "dle3"}


def make_case(nx: int, ny: int, nn: int, *, seed: int = 1234, jitter: float = 0.0, bed: bool = False,
              Ubulk: float = 15.6, noise: float = 0.05, stretch: float = 2.0, params: Params | None = None,
              Lx: float = LX, Ly: float = LY, Hh: float = H, walls: tuple | None = None,
              save: str | None = None) -> Case:
    """Generate a channel case.  walls = (XPB, YPB, XBC, YBC) for a box with
    walled sides (default periodic in x and y); save: also write the
    finished set-up to this file (load_case reads it back).  Raises
    RuntimeError with the set-up's message."""
    lib = _lib.host()
    xpb, ypb, xbc, ybc = walls if walls is not None else (1, 1, 0, 1)
    ptr = lib.nsx_case_create_walls(nx, ny, nn, Lx, Ly, Hh, stretch, jitter, int(bool(bed)), Ubulk, noise,
                                    ctypes.c_uint64(seed), int(xpb), int(ypb), int(xbc), int(ybc))
    if save is not None and lib.nsx_case_status(ptr, None, 0) == 0:
        if lib.nsx_case_save(ptr, save.encode()) != 0:
            lib.nsx_case_free(ptr)
            raise OSError(f"cannot write the set-up to {save}")
    return _collect(lib, ptr, nx, ny, nn, params or Params(sigFin=Hh))


def load_case(path: str, params: Params | None = None) -> Case:
    """A set-up written by make_case(save=...) (nsx_case_load: no geometry,
    flag or matrix build).  Raises RuntimeError when the file is missing,
    foreign, truncated or corrupt."""
    lib = _lib.host()
    ptr = lib.nsx_case_load(path.encode())
    n = ctypes.c_int64()
    t = ctypes.c_int()
    g = lib.nsx_case_get(ptr, b"grid", ctypes.byref(n), ctypes.byref(t))
    grid = np.ctypeslib.as_array(ctypes.cast(g, ctypes.POINTER(ctypes.c_int32)), shape=(3,)).copy() if g else (0, 0, 0)
    return _collect(lib, ptr, int(grid[0]), int(grid[1]), int(grid[2]), params or Params())


MESH_KEYS = ("No_vp", "xv", "yv", "No_cp", "zbv", "nbe", "nbev")


def read_mesh(path: str, N_CELL0: int, N_VERT: int) -> dict:
    """The mesh file of the reference (data.txt, input.F90:147-194) as arrays
    No_vp, No_cp (N_CELL0, 3), xv, yv, zbv, nbev (N_VERT), nbe (N_CELL0)."""
    m = _mesh_arrays(N_CELL0, N_VERT)
    rc = _lib.host().nsx_read_mesh(path.encode(), N_CELL0, N_VERT, *[m[k].ctypes.data for k in MESH_KEYS])
    if rc == -1:
        raise OSError(f"cannot open {path}")
    if rc < 0:
        raise ValueError(f"{path}: record {-rc} cannot be read as the mesh file")
    return m


def write_mesh(path: str, mesh: dict) -> None:
    """Write mesh arrays (or a Case) in the reference's data.txt format."""
    get = mesh.arrays.get if isinstance(mesh, Case) else mesh.get
    n0, nv = get("nbe").shape[0], get("nbev").shape[0]
    a = {k: np.asfortranarray(get(k)[:n0] if k == "No_cp" else get(k)) for k in MESH_KEYS}
    if _lib.host().nsx_write_mesh(path.encode(), n0, nv, *[a[k].ctypes.data for k in MESH_KEYS]) != 0:
        raise OSError(f"cannot write {path}")


def _mesh_arrays(n0: int, nv: int) -> dict:
    return {"No_vp": np.zeros((n0, 3), np.int32, order="F"), "xv": np.zeros(nv), "yv": np.zeros(nv),
            "No_cp": np.zeros((n0, 3), np.int32, order="F"), "zbv": np.zeros(nv),
            "nbe": np.zeros(n0, np.int32), "nbev": np.zeros(nv, np.int32)}


def case_from_mesh(mesh: dict, nn: int, *, box=(0.0, LX, 0.0, LY), Hh: float = H, stretch: float = 2.0,
                   Ubulk: float = 15.6, noise: float = 0.05, seed: int = 1234, params: Params | None = None) -> Case:
    """calcul_geometry + adjust_bc + PDE_center (and the seeded initial field)
    on a given mesh (read_mesh's arrays) in the doubly periodic box
    (XDIni, XDFin, YDIni, YDFin).  Raises RuntimeError with the set-up's
    message."""
    n0, nv = mesh["nbe"].shape[0], mesh["nbev"].shape[0]
    a = {k: np.asfortranarray(mesh[k], dtype=np.int32 if k in ("No_vp", "No_cp", "nbe", "nbev") else np.float64)
         for k in MESH_KEYS}
    lib = _lib.host()
    ptr = lib.nsx_case_from_mesh(n0, nv, nn, *[a[k].ctypes.data for k in MESH_KEYS], *[float(b) for b in box], Hh,
                                 stretch, Ubulk, noise, ctypes.c_uint64(seed))
    return _collect(lib, ptr, 0, 0, nn, params or Params(sigFin=Hh))


def _collect(lib, ptr, nx, ny, nn, params) -> Case:
    try:
        msg = ctypes.create_string_buffer(256)
        if lib.nsx_case_status(ptr, msg, 256) != 0:
            raise RuntimeError("case set-up failed: " + msg.value.decode())
        info = (ctypes.c_int * 8)()
        lib.nsx_case_info(ptr, info)
        c = Case(nx=nx, ny=ny, nn=nn, N_CELL0=info[0], N_CELL=info[1], N_VERT=info[2], NZ=info[3],
                 N_CELLexact=info[4], N_CELLghost=info[5], bctop=info[6], bcbot=info[7],
                 TotalVolume=lib.nsx_case_total_volume(ptr), params=params)
        for a in range(lib.nsx_case_num_arrays(ptr)):
            name = lib.nsx_case_array_name(ptr, a).decode()
            n = ctypes.c_int64()
            t = ctypes.c_int()
            p = lib.nsx_case_get(ptr, name.encode(), ctypes.byref(n), ctypes.byref(t))
            ct = ctypes.c_double if t.value == 1 else ctypes.c_int32
            arr = np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ct)), shape=(n.value,)).copy()
            if name in _SHAPES:
                arr = arr.reshape(_SHAPES[name](c), order="F")
            elif name in _FACE3:
                arr = arr.reshape((c.N_CELL0, 3), order="F")
            c.arrays[name] = arr
        return c
    finally:
        lib.nsx_case_free(ptr)


def config_case(cfg: int, *, seed: int = 1234, **over) -> Case:
    """The channel case of workload `cfg` (0..4)."""
    d = dict(CONFIGS[cfg])
    d.update(over)
    p = Params(dt=d["dt"], Re=d["Re"])
    return make_case(d["nx"], d["ny"], d["nn"], seed=seed, jitter=d["jitter"], bed=d["bed"],
                     Ubulk=d["Ubulk"], params=p)


def gaussian_random_field(case: Case, seed: int, corr: float = 0.5, std: float = 0.1, mean: float = 0.0,
                          modes: int = 512) -> np.ndarray:
    """A seeded Gaussian random field at the prism centres (xc, yc, sig(k)),
    interior prisms (N_CELL0, NZ-2): mean + std * sqrt(2/M) sum_m cos(k_m . x
    + phi_m) with wave vectors k_m ~ N(0, I/corr^2) and phases phi_m ~ U(0,
    2 pi) (random Fourier features), whose covariance tends to std^2
    exp(-|r|^2 / (2 corr^2)) -- correlation length corr -- as M grows."""
    rng = np.random.default_rng(seed)
    k = rng.standard_normal((modes, 3)) / corr
    phi = rng.uniform(0.0, 2.0 * math.pi, modes)
    A = case.arrays
    x = A["xc"][:case.N_CELL0][:, None]
    y = A["yc"][:case.N_CELL0][:, None]
    z = A["sig"][1:-1][None, :]
    out = np.zeros((case.N_CELL0, case.NZ - 2))
    for m in range(modes):
        out += np.cos(k[m, 0] * x + k[m, 1] * y + k[m, 2] * z + phi[m])
    return mean + std * math.sqrt(2.0 / modes) * out
