"""Python binding of libaphid_gpu.so (include/aphid_gpu.h) via ctypes.

This mirrors the reference's interface for the hot path: the operator
apply (inprod_x/y/z/p, src/aphid.f90:803-1282), the setup it depends on
(gen_k2_*grid, gen_diag, gen_inprod_*_coefs, read_mesh's sigma models,
gen_rhs) and the QMR solve (qmr, src/aphid_qmr.f90:5-434), with the
reference's array shapes and in-place solution convention.

There is no CPU fallback: if the HIP library is missing or no GPU is
visible, constructing a :class:`Solver` raises.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os
from typing import List, Optional, Sequence, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libaphid_gpu.so")
_lib = None

F_X0 = 1
F_NO_TRUE_RES = 2
F_MIXED = 4
# meshes need nx*ny*nz < MAX_NODES (APHID_MAX_NODES: 32-bit stencil offsets)
MAX_NODES = 67108864
# scratch of aphid_apply_noalloc per GPU (APHID_NOALLOC_SCRATCH_BYTES)
NOALLOC_SCRATCH_BYTES = 64 << 20
# the stencil tiles (tk, tj) the kernels are built for (kTiles)
TILES = [(16, 16), (32, 8), (64, 4), (16, 8), (32, 4), (32, 16), (64, 8), (16, 32)]
WS_PREV = 1
WS_LSQ = 2
WS_JLSQ = 4
STATUS = {0: "converged", 1: "max_it", 2: "breakdown"}


class AphidError(RuntimeError):
    pass


class Report(C.Structure):
    _fields_ = [("iters", C.c_int), ("status", C.c_int), ("rnorm", C.c_double),
                ("rhsnorm", C.c_double), ("true_rel_res", C.c_double),
                ("seconds_qmr", C.c_double), ("seconds_total", C.c_double),
                ("hist", C.POINTER(C.c_double)), ("hist_cap", C.c_int),
                ("hist_len", C.c_int), ("outer_iters", C.c_int), ("reserved", C.c_int),
                ("hist_time", C.POINTER(C.c_double))]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise AphidError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
        _lib = C.CDLL(LIB_PATH)
        _lib.aphid_last_error.restype = C.c_char_p
        _lib.aphid_unet_create.restype = C.c_void_p
        _lib.aphid_unet_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_ulonglong]
        _lib.aphid_unet_destroy.argtypes = [C.c_void_p]
        _lib.aphid_unet_param_count.restype = C.c_longlong
        _lib.aphid_unet_param_count.argtypes = [C.c_void_p]
        _lib.aphid_unet_forward_flops.restype = C.c_double
        _lib.aphid_unet_forward_flops.argtypes = [C.c_void_p]
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def comp_shapes(nx, ny, nz) -> List[Tuple[int, int, int]]:
    return [(nx - 1, ny, nz), (nx, ny - 1, nz), (nx, ny, nz - 1), (nx, ny, nz)]


def _c128(a, shape):
    a = np.asfortranarray(a, dtype=np.complex128)
    if a.shape != tuple(shape):
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


def _check(rc, what):
    if rc < 0:
        raise AphidError(f"{what}: {lib().aphid_last_error().decode()} (code {rc})")
    return rc


@dataclasses.dataclass
class SolveResult:
    x: List[np.ndarray]
    iters: int
    status: int
    rnorm: float
    rhsnorm: float
    true_rel_res: float
    seconds_qmr: float
    seconds_total: float
    outer_iters: int
    hist: np.ndarray
    hist_time: Optional[np.ndarray] = None

    @property
    def converged(self) -> bool:
        return self.status == 0


class Solver:
    """One device context: mesh, operator data and Krylov workspace."""

    def __init__(self, device: int = -1):
        self._ctx = C.c_void_p()
        _check(lib().aphid_ctx_create(C.byref(self._ctx), C.c_int(device)), "aphid_ctx_create")
        self.shape: Optional[Tuple[int, int, int]] = None

    def close(self):
        if self._ctx:
            lib().aphid_ctx_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- setup -------------------------------------------------------------
    def set_mesh(self, xn, yn, zn):
        xn, yn, zn = (np.ascontiguousarray(a, np.float64) for a in (xn, yn, zn))
        _check(lib().aphid_set_mesh(self._ctx, len(xn), len(yn), len(zn), _p(xn), _p(yn), _p(zn)),
               "aphid_set_mesh")
        self.shape = (len(xn), len(yn), len(zn))

    def set_cells(self, model: str, omega: float, cells: dict):
        nx, ny, nz = self.shape
        cs = (nx - 1, ny - 1, nz - 1)

        def f(name):
            a = cells.get(name)
            if a is None:
                return None
            a = np.asfortranarray(a, dtype=np.float32)
            if a.shape != cs:
                raise ValueError(f"cell block {name}: expected {cs}, got {a.shape}")
            return a

        blocks = {"R": ("sigma", None, None, None), "C": ("sigma", "eps_rel", None, None),
                  "D": ("sigma", "eta", "tau", "cee"), "I": (None, "eps_rel", None, None)}[model]
        arrs = [f(n) if n else None for n in blocks]
        _check(lib().aphid_set_cells(self._ctx, C.c_char(model.encode()), C.c_double(omega),
                                     *[_p(a) for a in arrs]), "aphid_set_cells")

    def set_sigma(self, sig, omega: float):
        nx, ny, nz = self.shape
        sig = _c128(sig, (nx - 1, ny - 1, nz - 1))
        _check(lib().aphid_set_sigma(self._ctx, _p(sig), C.c_double(omega)), "aphid_set_sigma")

    def set_k2(self, k2):
        k2 = [_c128(a, s) for a, s in zip(k2, comp_shapes(*self.shape))]
        _check(lib().aphid_set_k2(self._ctx, *[_p(a) for a in k2]), "aphid_set_k2")

    def set_diag(self, dg):
        dg = [_c128(a, s) for a, s in zip(dg, comp_shapes(*self.shape))]
        _check(lib().aphid_set_diag(self._ctx, *[_p(a) for a in dg]), "aphid_set_diag")

    def _zeros4(self):
        return [np.zeros(s, np.complex128, order="F") for s in comp_shapes(*self.shape)]

    def get_k2(self):
        out = self._zeros4()
        _check(lib().aphid_get_k2(self._ctx, *[_p(a) for a in out]), "aphid_get_k2")
        return out

    def get_diag(self):
        out = self._zeros4()
        _check(lib().aphid_get_diag(self._ctx, *[_p(a) for a in out]), "aphid_get_diag")
        return out

    def setup_case(self, case):
        """read_mesh + gen_k2_* + gen_diag for a meshgen.Case; returns b."""
        self.set_mesh(case.xn, case.yn, case.zn)
        self.set_cells(case.model, case.omega, case.cells)
        b, skipped = self.build_rhs(case.mode, case.sources)
        return b

    # -- operator and right-hand side ---------------------------------------
    def apply(self, u):
        u = [_c128(a, s) for a, s in zip(u, comp_shapes(*self.shape))]
        out = self._zeros4()
        _check(lib().aphid_apply(self._ctx, *[_p(a) for a in u], *[_p(a) for a in out]),
               "aphid_apply")
        return out

    def apply_noalloc(self, u):
        """K u like apply(), through the library's per-GPU scratch array
        (no device allocation; meshes up to NOALLOC_SCRATCH_BYTES)."""
        u = [_c128(a, s) for a, s in zip(u, comp_shapes(*self.shape))]
        out = self._zeros4()
        _check(lib().aphid_apply_noalloc(self._ctx, *[_p(a) for a in u], *[_p(a) for a in out]),
               "aphid_apply_noalloc")
        return out

    def apply_prec(self, u, precision: int):
        """K u with the complex128 (0) or complex64 (1) stencil."""
        u = [_c128(a, s) for a, s in zip(u, comp_shapes(*self.shape))]
        out = self._zeros4()
        _check(lib().aphid_apply_prec(self._ctx, precision, *[_p(a) for a in u],
                                      *[_p(a) for a in out]), "aphid_apply_prec")
        return out

    def build_rhs(self, mode: int, sources: Sequence[Tuple[int, int, int, int]] = ()):
        out = self._zeros4()
        ns = len(sources)
        arr = np.array(sources, np.int32).reshape(ns, 4) if ns else np.zeros((0, 4), np.int32)
        cols = [np.ascontiguousarray(arr[:, q]) for q in range(4)]
        skipped = np.zeros(max(ns, 1), np.int32)
        _check(lib().aphid_build_rhs(self._ctx, mode, ns, *[_p(a) for a in cols], _p(skipped),
                                     *[_p(a) for a in out]), "aphid_build_rhs")
        return out, skipped[:ns]

    # -- solve ------------------------------------------------------------------
    def qmr_solve(self, b, tol: float, max_it: int, x0=None, flags: int = 0,
                  hist_cap: Optional[int] = None) -> SolveResult:
        shp = comp_shapes(*self.shape)
        x = [np.array(_c128(a, s), order="F", copy=True) for a, s in zip(b, shp)]
        x0a = [None] * 4
        if x0 is not None:
            x0a = [_c128(a, s) for a, s in zip(x0, shp)]
            flags |= F_X0
        cap = max_it + 1 if hist_cap is None else hist_cap
        hist = np.zeros(max(cap, 1))
        htime = np.zeros(max(cap, 1))
        rep = Report()
        rep.hist = hist.ctypes.data_as(C.POINTER(C.c_double))
        rep.hist_time = htime.ctypes.data_as(C.POINTER(C.c_double))
        rep.hist_cap = cap
        _check(lib().aphid_qmr_solve(self._ctx, *[_p(a) for a in x], *[_p(a) for a in x0a],
                                     C.c_double(tol), C.c_int(max_it), C.c_uint(flags),
                                     C.byref(rep)), "aphid_qmr_solve")
        return SolveResult(x, rep.iters, rep.status, rep.rnorm, rep.rhsnorm, rep.true_rel_res,
                           rep.seconds_qmr, rep.seconds_total, rep.outer_iters,
                           hist[:rep.hist_len].copy(), htime[:rep.hist_len].copy())

    def load_rhs(self, b, x0=None):
        shp = comp_shapes(*self.shape)
        b = [_c128(a, s) for a, s in zip(b, shp)]
        x0a = [None] * 4 if x0 is None else [_c128(a, s) for a, s in zip(x0, shp)]
        _check(lib().aphid_load_rhs(self._ctx, *[_p(a) for a in b], *[_p(a) for a in x0a]),
               "aphid_load_rhs")

    def solve_resident(self, tol: float, max_it: int, flags: int = 0,
                       hist_cap: int = 0) -> SolveResult:
        hist = np.zeros(max(hist_cap, 1))
        rep = Report()
        rep.hist = hist.ctypes.data_as(C.POINTER(C.c_double))
        rep.hist_cap = hist_cap
        _check(lib().aphid_solve_resident(self._ctx, C.c_double(tol), C.c_int(max_it),
                                          C.c_uint(flags), C.byref(rep)), "aphid_solve_resident")
        return SolveResult([], rep.iters, rep.status, rep.rnorm, rep.rhsnorm, rep.true_rel_res,
                           rep.seconds_qmr, rep.seconds_total, rep.outer_iters,
                           hist[:rep.hist_len].copy())

    def warm_start(self, net: Optional["UNet"] = None, features=None, prev: bool = False,
                   lsq: bool = False, jlsq: bool = False) -> float:
        """Initial guess for solve_resident(..., flags=F_X0) from the network
        and/or the current solution (prev), amplitudes by the Galerkin
        condition (or least squares on the residual; jlsq: least squares on
        the Jacobi-weighted residual, network only); returns
        ||b - K x0|| / ||b||."""
        r0 = C.c_double()
        f = None if features is None else np.ascontiguousarray(features, np.float32)
        flags = (WS_PREV if prev else 0) | (WS_LSQ if lsq else 0) | (WS_JLSQ if jlsq else 0)
        _check(lib().aphid_warm_start(self._ctx, None if net is None else C.c_void_p(net._n),
                                      _p(f), C.c_uint(flags), C.byref(r0)), "aphid_warm_start")
        return r0.value

    def gen_fields(self, omega: float):
        """B = curl A on faces, E = -i omega (A - grad Phi) on edges of the
        resident solution: [Bx, By, Bz, Ex, Ey, Ez] in the reference's shapes."""
        nx, ny, nz = self.shape
        shp = [(nx, ny - 1, nz - 1), (nx - 1, ny, nz - 1), (nx - 1, ny - 1, nz),
               (nx - 1, ny, nz), (nx, ny - 1, nz), (nx, ny, nz - 1)]
        out = [np.zeros(s, np.complex128, order="F") for s in shp]
        _check(lib().aphid_gen_fields(self._ctx, C.c_double(omega), *[_p(a) for a in out]),
               "aphid_gen_fields")
        return out

    def get_sigma(self) -> np.ndarray:
        nx, ny, nz = self.shape
        out = np.zeros((nx - 1, ny - 1, nz - 1), np.complex128, order="F")
        _check(lib().aphid_get_sigma(self._ctx, _p(out)), "aphid_get_sigma")
        return out

    def fetch_solution(self):
        out = self._zeros4()
        _check(lib().aphid_fetch_solution(self._ctx, *[_p(a) for a in out]),
               "aphid_fetch_solution")
        return out

    # -- tuning / timing ------------------------------------------------------
    def residual_loss_grad(self, net: "UNet", features, jacobi: bool = True):
        """L = ||W (b - K x0)||^2 / ||W b||^2 of the network's fitted guess
        on this context's operator and resident b (W = 1/diag if jacobi,
        else 1), and dL/dparams."""
        g = np.zeros(net.n_params, np.float32)
        loss = C.c_double()
        _check(lib().aphid_warm_residual_loss_grad(self._ctx, C.c_void_p(net._n),
                                                   _p(net._f(features, 3)),
                                                   C.c_uint(WS_JLSQ if jacobi else WS_LSQ),
                                                   C.byref(loss), _p(g)),
               "aphid_warm_residual_loss_grad")
        return loss.value, g

    def train_residual(self, net: "UNet", sigmas, omegas, steps: int, batch: int = 4,
                       lr: float = 1e-3, seed: int = 0, jacobi: bool = True):
        """Residual training over the network's stored samples: sigmas
        [nsamples][cells] (model R, Fortran order per sample), omegas
        [nsamples]; this context holds the mesh and the shared b.  Returns
        (mean loss per step, seconds)."""
        sig = np.ascontiguousarray(
            np.stack([np.asarray(s, np.float32).ravel(order="F") for s in sigmas]), np.float32)
        om = np.ascontiguousarray(omegas, np.float64)
        losses = np.zeros(max(steps, 1), np.float32)
        secs = C.c_double()
        _check(lib().aphid_warm_train_residual(self._ctx, C.c_void_p(net._n), _p(sig), _p(om),
                                               C.c_uint(WS_JLSQ if jacobi else WS_LSQ),
                                               steps, batch, C.c_float(lr), C.c_ulonglong(seed),
                                               _p(losses), C.byref(secs)),
               "aphid_warm_train_residual")
        return losses[:steps], secs.value

    def get_tile(self, precision: int = 0):
        t = [C.c_int(), C.c_int(), C.c_int()]
        _check(lib().aphid_get_tile(self._ctx, precision, *[C.byref(v) for v in t]), "aphid_get_tile")
        return tuple(v.value for v in t)

    def set_tile(self, tk: int, tj: int, chunk: int, precision: int = 0):
        _check(lib().aphid_set_tile(self._ctx, precision, tk, tj, chunk), "aphid_set_tile")

    def time_apply(self, precision: int = 0, reps: int = 20) -> float:
        s = C.c_double()
        _check(lib().aphid_time_apply(self._ctx, precision, reps, C.byref(s)), "aphid_time_apply")
        return s.value

    def time_copy(self, nbytes: float, reps: int = 10) -> float:
        """Seconds per plain device copy of nbytes (reads nbytes, writes nbytes)."""
        t = C.c_double()
        _check(lib().aphid_time_copy(self._ctx, C.c_double(nbytes), reps, C.byref(t)),
               "aphid_time_copy")
        return t.value

    def set_reproducible(self, on: bool = True):
        """Complex128 solves whose iterates do not depend on the stencil's
        launch shape (eps from fixed-order segment partials)."""
        _check(lib().aphid_set_reproducible(self._ctx, int(bool(on))), "aphid_set_reproducible")

    def set_compact_k2(self, on: bool = True) -> bool:
        """Compact complex128 k^2 in the stencil (imaginary parts as two
        pair arrays, k2p's real part as a one-byte offset of its bit pattern,
        2-D tables of the edge grids' real parts; used only where exact);
        returns whether it is in use."""
        a = C.c_int()
        _check(lib().aphid_set_compact_k2(self._ctx, int(bool(on)), C.byref(a)),
               "aphid_set_compact_k2")
        return bool(a.value)

    def time_traffic(self, reps: int = 20) -> float:
        """Seconds per launch of the tuned stencil tile's loads and stores
        alone (no arithmetic): the floor its access pattern sets."""
        t = C.c_double()
        _check(lib().aphid_time_traffic(self._ctx, reps, C.byref(t)), "aphid_time_traffic")
        return t.value

    def time_iteration(self, precision: int = 0, reps: int = 10):
        """[p-update, stencil+eps, ||t-beta v||, fused update, iteration] seconds."""
        out = np.zeros(5)
        _check(lib().aphid_time_iteration(self._ctx, precision, reps, _p(out)),
               "aphid_time_iteration")
        return out


class SysReport(C.Structure):
    _fields_ = [("iters", C.c_int), ("status", C.c_int), ("rnorm", C.c_double),
                ("rhsnorm", C.c_double), ("true_rel_res", C.c_double), ("r0_rel", C.c_double)]


class BatchSolver:
    """Many independent systems in one launch per QMR phase (the on-GPU
    replacement of the reference's MPI farm, src/aphid_parallel.f90)."""

    def __init__(self, device: int = -1):
        self._b = C.c_void_p()
        _check(lib().aphid_batch_create(C.byref(self._b), C.c_int(device)), "aphid_batch_create")
        self.op_shapes: List[Tuple[int, int, int]] = []
        self.sys_ops: List[int] = []

    def close(self):
        if self._b:
            lib().aphid_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add_operator(self, xn, yn, zn, model: str, omega: float, cells: dict) -> int:
        xn, yn, zn = (np.ascontiguousarray(a, np.float64) for a in (xn, yn, zn))
        nx, ny, nz = len(xn), len(yn), len(zn)
        cs = (nx - 1, ny - 1, nz - 1)
        blocks = {"R": ("sigma", None, None, None), "C": ("sigma", "eps_rel", None, None),
                  "D": ("sigma", "eta", "tau", "cee"), "I": (None, "eps_rel", None, None)}[model]
        arrs = []
        for n in blocks:
            a = None if n is None else np.asfortranarray(cells[n], dtype=np.float32)
            if a is not None and a.shape != cs:
                raise ValueError(f"cell block {n}: expected {cs}, got {a.shape}")
            arrs.append(a)
        rc = _check(lib().aphid_batch_add_operator(self._b, nx, ny, nz, _p(xn), _p(yn), _p(zn),
                                                    C.c_char(model.encode()), C.c_double(omega),
                                                    *[_p(a) for a in arrs]),
                    "aphid_batch_add_operator")
        self.op_shapes.append((nx, ny, nz))
        return rc

    def add_case_operator(self, case) -> int:
        return self.add_operator(case.xn, case.yn, case.zn, case.model, case.omega, case.cells)

    def add_system(self, op: int, b, x0=None) -> int:
        shp = comp_shapes(*self.op_shapes[op])
        b = [_c128(a, s) for a, s in zip(b, shp)]
        x0a = [None] * 4 if x0 is None else [_c128(a, s) for a, s in zip(x0, shp)]
        rc = _check(lib().aphid_batch_add_system(self._b, op, *[_p(a) for a in b],
                                                  *[_p(a) for a in x0a]), "aphid_batch_add_system")
        self.sys_ops.append(op)
        return rc

    def solve(self, tol: float, max_it: int, flags: int = 0):
        ns = len(self.sys_ops)
        reps = (SysReport * ns)()
        secs = C.c_double()
        _check(lib().aphid_batch_solve(self._b, C.c_double(tol), C.c_int(max_it), C.c_uint(flags),
                                       reps, C.byref(secs)), "aphid_batch_solve")
        return [dict(iters=r.iters, status=r.status, rnorm=r.rnorm, rhsnorm=r.rhsnorm,
                     true_rel_res=r.true_rel_res, r0_rel=r.r0_rel) for r in reps], secs.value

    def solution(self, sys: int):
        out = [np.zeros(s, np.complex128, order="F")
               for s in comp_shapes(*self.op_shapes[self.sys_ops[sys]])]
        _check(lib().aphid_batch_get_solution(self._b, sys, *[_p(a) for a in out]),
               "aphid_batch_get_solution")
        return out

    def set_guess(self, sys: int, from_sys: int):
        _check(lib().aphid_batch_set_guess(self._b, sys, from_sys), "aphid_batch_set_guess")

    def set_sweep_guess(self, sys: int, from_sys: Sequence[int]):
        """Fitted system: its initial guess is fitted (Galerkin) to the
        solutions of from_sys (systems on the same mesh, solved in earlier
        waves; they may be fitted systems themselves, without a cycle)."""
        arr = np.ascontiguousarray(list(from_sys), np.int32)
        _check(lib().aphid_batch_set_sweep_guess(self._b, sys, len(arr),
                                                 _p(arr) if len(arr) else None),
               "aphid_batch_set_sweep_guess")

    def set_history(self, cap: int):
        """Keep ||r|| of the first `cap` iterations of every system on
        later complex128 solves (0: none)."""
        _check(lib().aphid_batch_set_history(self._b, C.c_int(cap)), "aphid_batch_set_history")

    def history(self, sys: int, cap: int):
        """(||r|| per iteration, seconds at which each was read) of system
        sys from the last solve, at most cap entries."""
        h = np.zeros(max(cap, 0), np.float64)
        t = np.zeros(max(cap, 0), np.float64)
        n = C.c_int()
        _check(lib().aphid_batch_get_history(self._b, sys, _p(h) if cap else None,
                                             _p(t) if cap else None, C.c_int(cap), C.byref(n)),
               "aphid_batch_get_history")
        return h[:n.value].copy(), t[:n.value].copy()


class UNet:
    """The warm-start network (aphid_unet_*): a 3-level 3-D U-Net on a cell
    grid of (nx, ny, nz) cells, trained and evaluated on the GPU."""

    # buffer indices of aphid_unet_activation (the network's tensor list)
    BUFFERS = ["in", "a1", "cat1", "p1", "a2", "cat2", "p2", "a3", "b3", "r3",
               "d2a", "d2b", "r2", "d1a", "d1b", "out"]
    CHANNELS = [8, 16, 32, 16, 32, 64, 32, 64, 64, 32, 32, 32, 16, 16, 16, 8]
    LEVELS = [0, 0, 0, 1, 1, 1, 2, 2, 2, 2, 1, 1, 1, 0, 0, 0]

    def __init__(self, cells: Tuple[int, int, int], seed: int = 0, device: int = 0):
        self.cells = tuple(int(c) for c in cells)
        n = lib().aphid_unet_create(device, *self.cells, C.c_ulonglong(seed))
        if not n:
            raise AphidError(f"aphid_unet_create: {lib().aphid_last_error().decode()}")
        self._n = n
        self.n_params = int(lib().aphid_unet_param_count(C.c_void_p(n)))
        self.V = int(np.prod(self.cells))

    def close(self):
        if self._n:
            lib().aphid_unet_destroy(C.c_void_p(self._n))
            self._n = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def forward_flops(self) -> float:
        return float(lib().aphid_unet_forward_flops(C.c_void_p(self._n)))

    def _f(self, a, ch):
        a = np.ascontiguousarray(a, np.float32)
        if a.size != self.V * ch:
            raise ValueError(f"expected {self.V}x{ch} values, got {a.shape}")
        return a

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.n_params, np.float32)
        _check(lib().aphid_unet_get_params(C.c_void_p(self._n), _p(out)), "aphid_unet_get_params")
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.n_params:
            raise ValueError("wrong parameter count")
        _check(lib().aphid_unet_set_params(C.c_void_p(self._n), _p(p)), "aphid_unet_set_params")

    def forward(self, features) -> np.ndarray:
        out = np.zeros((self.V, 8), np.float32)
        _check(lib().aphid_unet_forward(C.c_void_p(self._n), _p(self._f(features, 3)), _p(out)),
               "aphid_unet_forward")
        return out

    def loss_grad(self, features, target):
        g = np.zeros(self.n_params, np.float32)
        loss = C.c_double()
        _check(lib().aphid_unet_loss_grad(C.c_void_p(self._n), _p(self._f(features, 3)),
                                          _p(self._f(target, 8)), C.byref(loss), _p(g)),
               "aphid_unet_loss_grad")
        return loss.value, g

    def activation(self, name: str, grad: bool = False) -> np.ndarray:
        i = self.BUFFERS.index(name)
        lv = self.LEVELS[i]
        n = int(np.prod([c >> lv for c in self.cells])) * self.CHANNELS[i]
        out = np.zeros(n, np.float32)
        _check(lib().aphid_unet_activation(C.c_void_p(self._n), i, int(grad), _p(out)),
               "aphid_unet_activation")
        return out.reshape(-1, self.CHANNELS[i])

    def reserve(self, capacity: int):
        _check(lib().aphid_unet_reserve(C.c_void_p(self._n), capacity), "aphid_unet_reserve")

    def add_sample(self, features, target) -> int:
        return _check(lib().aphid_unet_add_sample(C.c_void_p(self._n), _p(self._f(features, 3)),
                                                  _p(self._f(target, 8))), "aphid_unet_add_sample")

    def train(self, steps: int, batch: int = 4, lr: float = 1e-3, seed: int = 0):
        losses = np.zeros(max(steps, 1), np.float32)
        secs = C.c_double()
        _check(lib().aphid_unet_train(C.c_void_p(self._n), steps, batch, C.c_float(lr),
                                      C.c_ulonglong(seed), _p(losses), C.byref(secs)),
               "aphid_unet_train")
        return losses[:steps], secs.value

    def time(self, reps: int = 5):
        f, s = C.c_double(), C.c_double()
        _check(lib().aphid_unet_time(C.c_void_p(self._n), reps, C.byref(f), C.byref(s)),
               "aphid_unet_time")
        return f.value, s.value


EXPORTS = ["aphid_ctx_create", "aphid_ctx_destroy", "aphid_last_error", "aphid_set_mesh",
           "aphid_set_cells", "aphid_set_sigma", "aphid_set_k2", "aphid_set_diag", "aphid_get_k2",
           "aphid_get_diag", "aphid_apply", "aphid_apply_noalloc", "aphid_apply_prec", "aphid_build_rhs", "aphid_qmr_solve", "aphid_get_tile",
           "aphid_set_tile", "aphid_time_apply", "aphid_load_rhs", "aphid_solve_resident",
           "aphid_fetch_solution", "aphid_time_iteration", "aphid_time_copy", "aphid_time_traffic", "aphid_set_reproducible",
           "aphid_set_compact_k2",
           "aphid_batch_create", "aphid_batch_destroy", "aphid_batch_add_operator",
           "aphid_batch_add_system", "aphid_batch_solve", "aphid_batch_get_solution",
           "aphid_batch_set_guess", "aphid_batch_set_sweep_guess", "aphid_batch_set_history",
           "aphid_batch_get_history",
           "aphid_unet_create", "aphid_unet_destroy", "aphid_unet_param_count",
           "aphid_unet_forward_flops", "aphid_unet_get_params", "aphid_unet_set_params",
           "aphid_unet_forward", "aphid_unet_loss_grad", "aphid_unet_activation",
           "aphid_unet_reserve", "aphid_unet_add_sample", "aphid_unet_train", "aphid_unet_time",
           "aphid_warm_start", "aphid_warm_residual_loss_grad", "aphid_warm_train_residual",
           "aphid_gen_fields", "aphid_get_sigma"]
