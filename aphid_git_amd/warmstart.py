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


def protocol(cells, ntrain: int, nhold: int, mse_steps: int, res_steps: int, batch: int = 4,
             lr: float = 2e-3, tol: float = 1e-8, max_it: int = 20000, seed0: int = 1000,
             hold0: int = 500000, log=None) -> dict:
    """The learned warm start end to end on random-block models of one cell
    grid (frequencies log-uniform in 1e-2..1e5 Hz, the same surface dipole):

    1. ntrain solver-made training samples: each model solved from x0 = 0
       to tol, its features and normalised solution stored on the GPU;
    2. two networks trained on the GPU over those models:
       "field": mse_steps Adam steps on the field (MSE against the solver's
       solutions), the guess fitted by the Galerkin condition as before;
       "residual": res_steps on the residual of the network's own guess,
       ||b - K x0||^2 / ||b||^2 after the least-squares fit of its four
       component amplitudes (aphid_warm_train_residual), fitted the same way;
    3. nhold held-out models, each solved from x0 = 0 and from each
       network's guess (the true residual as the judge: never worse than
       x0 = 0): iterations and seconds to tol.
    """
    t_all = time.time()
    s = gpu.Solver(0)
    nets = {"field": gpu.UNet(cells, seed=0), "residual": gpu.UNet(cells, seed=1)}
    for n in nets.values():
        n.reserve(ntrain)
    zeros = np.zeros((int(np.prod(cells)), 8), np.float32)
    sigmas, omegas, its = [], [], []
    b0 = None
    t0 = time.time()
    for q in range(ntrain):
        case = random_case(cells, seed0 + q)
        r, _, feat, b = solve_case(s, case, tol, max_it)
        if b0 is None:
            b0 = b
        nets["field"].add_sample(feat, target(s.fetch_solution()))
        nets["residual"].add_sample(feat, zeros)
        sigmas.append(np.asarray(case.cells["sigma"], np.float32))
        omegas.append(case.omega)
        its.append(r.iters)
        if log and (q + 1) % max(1, ntrain // 10) == 0:
            log(f"[protocol] data {q + 1}/{ntrain} mean iters {np.mean(its):.0f} "
                f"{time.time() - t0:.0f}s")
    t_data = time.time() - t0
    train = {}
    l_f, t_f = nets["field"].train(mse_steps, batch=batch, lr=lr, seed=1)
    train["field"] = {"steps": mse_steps, "seconds": t_f, "loss_first": float(l_f[:5].mean()),
                      "loss_last": float(l_f[-20:].mean())}
    s.setup_case(random_case(cells, seed0))  # the mesh; b is the same dipole for every case
    s.load_rhs(b0)
    l_r, t_r = s.train_residual(nets["residual"], sigmas, omegas, res_steps, batch=batch, lr=lr,
                                seed=2, jacobi=False)
    train["residual"] = {"steps": res_steps, "seconds": t_r, "loss_first": float(l_r[:5].mean()),
                         "loss_last": float(l_r[-20:].mean())}
    if log:
        log(f"[protocol] trained: {train}")
    del sigmas
    ev = []
    for q in range(nhold):
        case = random_case(cells, hold0 + q)
        rz, _, feat, b = solve_case(s, case, tol, max_it)
        e = {"freq": case.freq, "iters_zero": rz.iters, "sec_zero": rz.seconds_total,
             "ok_zero": rz.status == 0 and rz.true_rel_res <= tol}
        for name, net in nets.items():
            s.load_rhs(b)
            t = time.perf_counter()
            r0 = s.warm_start(net, feat, lsq=(name == "residual"))
            t_ws = time.perf_counter() - t
            rw = s.solve_resident(tol, max_it, gpu.F_X0)
            e[name] = {"iters": rw.iters, "sec": rw.seconds_total + t_ws, "r0": r0,
                       "ok": rw.status == 0 and rw.true_rel_res <= tol}
        ev.append(e)
        if log and (q + 1) % max(1, nhold // 8) == 0:
            log(f"[protocol] held-out {q + 1}/{nhold}")
    s.close()
    for n in nets.values():
        n.close()
    iz = float(np.mean([e["iters_zero"] for e in ev]))
    sz = float(np.mean([e["sec_zero"] for e in ev]))
    out = {"cells": list(cells), "train_samples": ntrain, "held_out": nhold,
           "data_seconds": t_data, "data_iters_mean": float(np.mean(its)),
           "iters_zero_mean": iz, "seconds_zero_mean": sz,
           "all_zero_converged": all(e["ok_zero"] for e in ev)}
    for name in nets:
        iw = float(np.mean([e[name]["iters"] for e in ev]))
        sw = float(np.mean([e[name]["sec"] for e in ev]))
        out[name] = {**train[name], "iters_warm_mean": iw, "iters_ratio": iw / iz,
                     "seconds_warm_mean": sw, "seconds_ratio": sw / sz,
                     "r0_median": float(np.median([e[name]["r0"] for e in ev])),
                     "r0_min": float(np.min([e[name]["r0"] for e in ev])),
                     "all_warm_converged": all(e[name]["ok"] for e in ev),
                     "per_decade": _per_decade(ev, name)}
    out["seconds_total"] = time.time() - t_all
    return out


def _per_decade(ev, name):
    out = {}
    for e in ev:
        d = int(np.floor(np.log10(e["freq"])))
        o = out.setdefault(f"1e{d}", {"n": 0, "iters_zero": 0, "iters_warm": 0})
        o["n"] += 1
        o["iters_zero"] += e["iters_zero"]
        o["iters_warm"] += e[name]["iters"]
    for o in out.values():
        o["ratio"] = o["iters_warm"] / o["iters_zero"]
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
