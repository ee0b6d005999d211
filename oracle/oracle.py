"""ctypes wrapper around liboracle.so -- TEST INFRASTRUCTURE ONLY.

The oracle is the CPU restatement of the reference algorithm
(oracle/aphid_oracle.c).  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg import this module, and only as the checker or
the CPU baseline; the product (aphid_git_amd) never does.

Arrays are numpy, Fortran-ordered, with the reference's shapes
(x-edges (nx-1,ny,nz), y-edges (nx,ny-1,nz), z-edges (nx,ny,nz-1),
nodes (nx,ny,nz)).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import List, Optional, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

MU0 = np.pi * 4.0e-7
EPS0 = 8.8541878176e-12


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = C.CDLL(_LIB)
        _lib.om_op_create.restype = C.c_void_p
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def comp_shapes(nx, ny, nz) -> List[Tuple[int, int, int]]:
    return [(nx - 1, ny, nz), (nx, ny - 1, nz), (nx, ny, nz - 1), (nx, ny, nz)]


def zeros4(nx, ny, nz, dtype=np.complex128):
    return [np.zeros(s, dtype=dtype, order="F") for s in comp_shapes(nx, ny, nz)]


def metrics(c: np.ndarray):
    c = np.ascontiguousarray(c, dtype=np.float64)
    n = len(c)
    d = np.zeros(n + 1)
    dc = np.zeros(n + 1)
    lib().om_metrics(C.c_int(n), _p(c), _p(d), _p(dc))
    return d, dc


class Mesh:
    def __init__(self, xn, yn, zn):
        self.xn, self.yn, self.zn = (np.asarray(a, np.float64) for a in (xn, yn, zn))
        self.nx, self.ny, self.nz = len(self.xn), len(self.yn), len(self.zn)
        self.dx, self.dxc = metrics(self.xn)
        self.dy, self.dyc = metrics(self.yn)
        self.dz, self.dzc = metrics(self.zn)

    @property
    def shape(self):
        return self.nx, self.ny, self.nz


def sigma_hat(case) -> np.ndarray:
    nx, ny, nz = case.shape
    nc = (nx - 1) * (ny - 1) * (nz - 1)
    cells = case.cells

    def f(name):
        a = cells.get(name)
        return None if a is None else np.asfortranarray(a, dtype=np.float32)

    a, b, c, d = f("sigma"), f("eps_rel"), f("eta"), f("tau")
    if case.model == "D":
        b, c, d = f("eta"), f("tau"), f("cee")
    elif case.model in "CI":
        b = f("eps_rel")
    sig = np.zeros((nx - 1, ny - 1, nz - 1), np.complex128, order="F")
    rc = lib().om_sigma(C.c_char(case.model.encode()), C.c_double(case.omega),
                        C.c_int64(nc), _p(a), _p(b), _p(c), _p(d), _p(sig))
    if rc != 0:
        raise ValueError(f"unknown sigma model {case.model!r}")
    return sig


def k2_grids(mesh: Mesh, sig: np.ndarray, omega: float):
    nx, ny, nz = mesh.shape
    k2 = zeros4(nx, ny, nz)
    lib().om_k2(nx, ny, nz, _p(mesh.dx), _p(mesh.dy), _p(mesh.dz),
                _p(np.asfortranarray(sig)), C.c_double(omega), *[_p(a) for a in k2])
    return k2


def jacobi_diag(mesh: Mesh, k2):
    nx, ny, nz = mesh.shape
    dg = zeros4(nx, ny, nz)
    lib().om_diag(nx, ny, nz, _p(mesh.dx), _p(mesh.dy), _p(mesh.dz),
                  *[_p(a) for a in k2], *[_p(a) for a in dg])
    return dg


def rhs(case, mesh: Optional[Mesh] = None):
    mesh = mesh or Mesh(case.xn, case.yn, case.zn)
    nx, ny, nz = mesh.shape
    b = zeros4(nx, ny, nz)
    ns = len(case.sources)
    arr = np.array(case.sources, dtype=np.int32).reshape(ns, 4) if ns else np.zeros((0, 4), np.int32)
    dr, ip, jp, kp = (np.ascontiguousarray(arr[:, q]) for q in range(4))
    skipped = np.zeros(max(ns, 1), np.int32)
    rc = lib().om_rhs(case.mode, nx, ny, nz, _p(mesh.xn), _p(mesh.yn), _p(mesh.zn),
                      _p(mesh.dx), _p(mesh.dy), _p(mesh.dz), ns, _p(dr), _p(ip),
                      _p(jp), _p(kp), _p(skipped), *[_p(a) for a in b])
    if rc != 0:
        raise ValueError(f"rhs mode {case.mode} invalid for this mesh")
    return b, skipped[:ns]


class Operator:
    """K applied matrix-free (inprod_x/y/z/p)."""

    def __init__(self, mesh: Mesh, k2):
        self.mesh = mesh
        self.k2 = [np.asfortranarray(a) for a in k2]
        nx, ny, nz = mesh.shape
        self._h = lib().om_op_create(nx, ny, nz, _p(mesh.dx), _p(mesh.dy), _p(mesh.dz),
                                     _p(mesh.dxc), _p(mesh.dyc), _p(mesh.dzc),
                                     *[_p(a) for a in self.k2])

    def __del__(self):
        if getattr(self, "_h", None):
            lib().om_op_destroy(C.c_void_p(self._h))
            self._h = None

    def apply(self, u):
        u = [np.asfortranarray(a, dtype=np.complex128) for a in u]
        out = zeros4(*self.mesh.shape)
        lib().om_apply(C.c_void_p(self._h), *[_p(a) for a in u], *[_p(a) for a in out])
        return out


class QmrReport(C.Structure):
    _fields_ = [("iters", C.c_int), ("status", C.c_int), ("rnorm", C.c_double),
                ("rhsnorm", C.c_double), ("rho0", C.c_double), ("seconds", C.c_double)]


def qmr(op: Operator, b, dg, tol: float, max_it: int, nthreads: int = 1,
        x0=None, hist_cap: Optional[int] = None):
    """Returns (x, report, hist).  b is not modified."""
    x = [np.array(a, dtype=np.complex128, order="F", copy=True) for a in b]
    dg = [np.asfortranarray(a) for a in dg]
    x0p = [None] * 4 if x0 is None else [np.asfortranarray(a, dtype=np.complex128) for a in x0]
    cap = hist_cap if hist_cap is not None else max_it + 1
    hist = np.zeros(max(cap, 1))
    rep = QmrReport()
    lib().om_qmr(C.c_void_p(op._h), *[_p(a) for a in x], *[_p(a) for a in dg],
                 *[_p(a) for a in x0p], C.c_double(tol), C.c_int(max_it),
                 C.c_int(nthreads), _p(hist), C.c_int(cap), C.byref(rep))
    # iterations executed: `iters`, except on exhaustion where the reference
    # reports max_it+1 (src/aphid_qmr.f90:387,390)
    n = rep.iters if rep.status != 1 else max(max_it, 1)
    return x, rep, hist[:min(max(n, 0), cap)]


def qmr_dense(A: np.ndarray, diag: np.ndarray, b: np.ndarray, tol: float, max_it: int):
    A = np.ascontiguousarray(A, dtype=np.complex128)
    x = np.array(b, dtype=np.complex128, copy=True)
    dgv = np.ascontiguousarray(diag, dtype=np.complex128)
    hist = np.zeros(max_it + 2)
    rep = QmrReport()
    lib().om_qmr_dense(C.c_int(len(x)), _p(A), _p(dgv), _p(x), C.c_double(tol),
                       C.c_int(max_it), _p(hist), C.c_int(max_it + 2), C.byref(rep))
    n = rep.iters if rep.status != 1 else max(max_it, 1)
    return x, rep, hist[:min(max(n, 0), max_it + 2)]


class System:
    """Everything the reference sets up before qmr (src/aphid.f90:72-100)."""

    def __init__(self, case):
        self.case = case
        self.mesh = Mesh(case.xn, case.yn, case.zn)
        self.sig = sigma_hat(case)
        self.k2 = k2_grids(self.mesh, self.sig, case.omega)
        self.diag = jacobi_diag(self.mesh, self.k2)
        self.b, self.skipped = rhs(case, self.mesh)
        self.op = Operator(self.mesh, self.k2)

    def solve(self, tol=None, max_it=None, nthreads=1, x0=None):
        return qmr(self.op, self.b, self.diag, self.case.target if tol is None else tol,
                   self.case.max_it if max_it is None else max_it, nthreads, x0)
