"""Learned warm start: network inputs and targets on the cell grid, and the
generate -> train -> evaluate loop.

The network (aphid_unet_* in libaphid_gpu.so) sees, per cell of an
(nx-1, ny-1, nz-1) cell grid (z fastest):
  0: (log10 sigma + 4) / 4 (air at 1e-8 S/m -> -1, earth 1e-4..1 -> 0..1);
  1: log10(h / delta) / 3 clipped to [-1, 1], where delta = sqrt(2/(omega mu0
     sigma)) is the skin depth and h the cell's geometric-mean size: the
     frequency as the solver sees it, cell by cell;
  2: source mask, 1 on cells touching an unknown with a nonzero right-hand
     side.
It predicts 8 channels per cell: re, im of Ax, Ay, Az, Phi at the unknowns
the cell owns (the edges leaving its lower corner node, and that node),
scaled by 1/rms(x) over all unknowns.  The solver maps the prediction back
onto the staggered unknowns and fits one complex amplitude per component by
a Galerkin condition (aphid_warm_start), so the network only has to get the
field's shape right.
"""
from __future__ import annotations

import time
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import gpu, meshgen

MU0 = 4e-7 * np.pi


def cell_sigma(case: meshgen.Case) -> np.ndarray:
    """Real conductivity per cell, [nx-1, ny-1, nz-1] (model R)."""
    if case.model != "R":
        raise ValueError("features are defined for the R (real sigma) model")
    return np.asarray(case.cells["sigma"], np.float64)


def features(case: meshgen.Case, b: Sequence[np.ndarray]) -> np.ndarray:
    """[V][3] float32 network input of a case and its right-hand side b."""
    sig = cell_sigma(case)
    dx, dy, dz = np.diff(case.xn), np.diff(case.yn), np.diff(case.zn)
    h = np.cbrt(dx[:, None, None] * dy[None, :, None] * dz[None, None, :])
    delta = np.sqrt(2.0 / (case.omega * MU0 * sig))
    f0 = (np.log10(sig) + 4.0) / 4.0
    f1 = np.clip(np.log10(h / delta) / 3.0, -1.0, 1.0)
    f2 = source_mask(b, sig.shape)
    return np.stack([f0, f1, f2], axis=-1).reshape(-1, 3).astype(np.float32)


def source_mask(b: Sequence[np.ndarray], cshape) -> np.ndarray:
    """1 on cells touching a nonzero right-hand-side unknown."""
    cx, cy, cz = cshape
    m = np.zeros(cshape)
    for c, a in enumerate(b):
        for i, j, k in zip(*np.nonzero(a)):
            lo, hi = [], []
            for ax, p, n in ((0, i, cx), (1, j, cy), (2, k, cz)):
                if ax == c:
                    lo.append(p), hi.append(p)
                else:
                    lo.append(max(p - 1, 0)), hi.append(min(p, n - 1))
            m[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1] = 1.0
    return m


def target(x: Sequence[np.ndarray]) -> np.ndarray:
    """[V][8] float32 target of a solution x (4 components, reference
    shapes): cell (i,j,k) holds the unknowns at its lower corner node -- the
    x-, y-, z-edges leaving node (i,j,k) and Phi there -- over rms(x)."""
    nx, ny, nz = x[3].shape
    rms = np.sqrt(sum(float(np.sum(np.abs(a) ** 2)) for a in x) / sum(a.size for a in x))
    chans = []
    for c in range(4):
        m = np.asarray(x[c])[:nx - 1, :ny - 1, :nz - 1] / (rms if rms > 0 else 1.0)
        chans += [m.real, m.imag]
    return np.stack(chans, axis=-1).reshape(-1, 8).astype(np.float32)


def random_case(ncell, seed: int, freq_range=(1e-2, 1e5), **kw) -> meshgen.Case:
    """A config-1-family random block model at a log-uniform frequency."""
    rng = np.random.default_rng(seed + 7919)
    f = float(10 ** rng.uniform(np.log10(freq_range[0]), np.log10(freq_range[1])))
    return meshgen.random_block_case(ncell, seed, f, **kw)


def solve_case(s: gpu.Solver, case: meshgen.Case, tol: float, max_it: int,
               net: Optional[gpu.UNet] = None, feat=None):
    """Set up `case` on solver s and solve it from x0 = 0 (net None) or from
    the network's warm start.  Returns (result, r0_rel, features, b)."""
    b = s.setup_case(case)
    if feat is None:
        feat = features(case, b)
    s.load_rhs(b)
    r0, t_ws = 1.0, 0.0
    flags = 0
    if net is not None:
        t = time.perf_counter()
        r0 = s.warm_start(net, feat)
        t_ws = time.perf_counter() - t
        flags = gpu.F_X0
    r = s.solve_resident(tol, max_it, flags)
    r.seconds_warm_start = t_ws
    return r, r0, feat, b


def build_dataset(net: gpu.UNet, ncell, seeds: Sequence[int], tol: float = 1e-8,
                  max_it: int = 20000, log=None) -> Dict[str, float]:
    """Solve one random case per seed from zero and add (features, target)."""
    s = gpu.Solver(0)
    t0 = time.time()
    its = []
    for q, sd in enumerate(seeds):
        case = random_case(ncell, sd)
        r, _, feat, _ = solve_case(s, case, tol, max_it)
        x = s.fetch_solution()
        net.add_sample(feat, target(x))
        its.append(r.iters)
        if log and (q + 1) % max(1, len(seeds) // 10) == 0:
            log(f"[data] {q + 1}/{len(seeds)} mean iters {np.mean(its):.0f} {time.time() - t0:.1f}s")
    s.close()
    return {"samples": len(seeds), "seconds": time.time() - t0, "mean_iters": float(np.mean(its))}


def evaluate(net: Optional[gpu.UNet], ncell, seeds: Sequence[int], tol: float = 1e-8,
             max_it: int = 20000) -> List[dict]:
    """Iterations and time to tol from x0 = 0 and from the warm start."""
    s = gpu.Solver(0)
    out = []
    for sd in seeds:
        case = random_case(ncell, sd)
        r0z, _, feat, _ = solve_case(s, case, tol, max_it)
        rec = {"seed": sd, "freq": case.freq, "iters_zero": r0z.iters,
               "sec_zero": r0z.seconds_total, "true_zero": r0z.true_rel_res}
        if net is not None:
            rw, r0, _, _ = solve_case(s, case, tol, max_it, net=net, feat=feat)
            rec.update(iters_warm=rw.iters, sec_warm=rw.seconds_total + rw.seconds_warm_start,
                       sec_warm_start=rw.seconds_warm_start, r0=r0,
                       true_warm=rw.true_rel_res, status_warm=rw.status)
        out.append(rec)
    s.close()
    return out


def cells_to_unknowns(out: np.ndarray, nodes) -> List[np.ndarray]:
    """Host restatement of the solver's mapping of a cell-grid prediction
    [V][8] onto the staggered unknowns (reference shapes): each unknown
    takes the cell whose lower corner node it leaves from (edges on the last
    planes take the nearest cell across); Phi on boundary nodes is 0."""
    nx, ny, nz = nodes
    cx, cy, cz = nx - 1, ny - 1, nz - 1
    o = np.asarray(out, np.float64).reshape(cx, cy, cz, 8)
    res = []
    for c in range(4):
        f = o[..., 2 * c] + 1j * o[..., 2 * c + 1]
        shp = [nx, ny, nz]
        if c < 3:
            shp[c] -= 1
        ii = np.minimum(np.arange(shp[0]), cx - 1)
        jj = np.minimum(np.arange(shp[1]), cy - 1)
        kk = np.minimum(np.arange(shp[2]), cz - 1)
        v = f[np.ix_(ii, jj, kk)]
        if c == 3:
            v[0, :, :] = v[-1, :, :] = 0
            v[:, 0, :] = v[:, -1, :] = 0
            v[:, :, 0] = v[:, :, -1] = 0
        res.append(np.asfortranarray(v))
    return res
