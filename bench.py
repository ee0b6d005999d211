"""Benchmark: converged A-Phi Jacobi-QMR solves per second on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload: the 128x128x96-cell stretched mesh (129x129x97 nodes,
M = 6,415,041 complex unknowns), 0.01 S/m earth with 20 seeded resistive /
conductive blocks under 16 air layers, +X electric dipole at the surface
centre, 1 Hz, complex128, Jacobi-QMR to relative residual 1e-8.  A step is
one complete solve from x0 = 0 with the operator and right-hand side already
resident in HBM; a step counts as converged when the recomputed true
residual ||b - K x|| / ||b|| is <= 1e-8.  Prints one JSON line.

Also measured: the stencil (matvec) kernel and every kernel of one QMR
iteration with HIP events on the library's stream (roofline against the
8 TB/s HBM peak), and a CPU baseline: the reference's own Fortran/OpenMP
QMR (compiled from its sources into oracle/_ref) timed on this host on a
bounded sample of the same system.

Further parts (--parts, default all), reported as extra JSON fields:
  batch  the 33^3-node model family at 64 frequencies (1e-2..1e5 Hz) x 3
         dipoles = 192 systems: one batched solve vs one solve at a time;
  unet   the warm-start network on a 96x96x80 cell grid: forward and
         forward+backward time, TFLOP/s against the 2.5 PFLOP/s dense bf16
         matrix-core peak;
  warm   the learned warm start's protocol on 16^3 cells: solver-made
         training set, two networks trained on the GPU (field MSE; the
         residual of the network's own guess), iterations vs x0 = 0 on
         held-out models.  --parts warm5 runs it at 96x96x80 cells with
         2000 training and 256 held-out models (about 15 minutes).
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Vector passes (reads + writes of one 4*nn complex vector) of a QMR
# iteration as the library schedules it: p-update 7 on average (v, 1/diag,
# p read and p written: 4; the previous iteration's d update: 2; its x
# update, every second iteration: 2), stencil 3 (u, k^2, K u),
# ||t - beta v|| 2, fused update 7.5 on average (t, s, r, v, 1/diag read,
# s and v written; r written every second iteration).  The reference's
# statements make 22 passes per iteration.
PUPDATE_PASSES = 7.0
UPDATE_PASSES = 7.5
ITERATION_PASSES = PUPDATE_PASSES + 3 + 2 + UPDATE_PASSES
BF16_PEAK_TFLOPS = 2500.0
TOL = 1e-8
MAX_IT = 100000
PARTS = ("solve", "config4", "shard", "batch", "unet", "warm", "cpu")
EXTRA_PARTS = ("batch3b", "warm5", "prec4", "prec4sweep")  # not default (minutes of GPU time)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_cpu():
    """lscpu model name and the host's logical CPU count."""
    model = ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=30).stdout
        m = re.search(r"Model name:\s*(.+)", out)
        model = m.group(1).strip() if m else ""
    except (OSError, subprocess.SubprocessError):
        pass
    return {"model_name": model, "nproc": os.cpu_count()}


def _ref_sample(harness, case, cmd: str, threads: int, sample_its: int):
    """Run the reference harness's QMR (cmd qmr1 = serial qmr, qmr4 =
    qmr_mthread) for sample_its iterations on the case; seconds per
    iteration from its qmr.iter clock (column 3, refreshed every 10
    iterations, src/aphid_qmr.f90:376-381) between iterations 10 and
    sample_its."""
    from aphid_git_amd import meshgen

    with tempfile.TemporaryDirectory() as d:
        c2 = meshgen.Case(case.xn, case.yn, case.zn, case.model, case.cells, case.freq,
                          case.mode, case.sources, case.target, sample_its, case.comment)
        meshgen.write_case(c2, d)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close")
        t0 = time.time()
        r = subprocess.run([harness, cmd], cwd=d, env=env, capture_output=True, text=True,
                           timeout=900)
        wall = time.time() - t0
        if r.returncode != 0:
            raise RuntimeError(r.stderr[-500:])
        it = np.loadtxt(os.path.join(d, "qmr.iter"), ndmin=2)
        t10 = it[it[:, 0] == 10][0, 2]
        tN = it[it[:, 0] == sample_its][0, 2]
        m = re.search(r"Time spent in QMR \[s\]:\s*([0-9.Ee+-]+)", r.stdout)
    return (tN - t10) / (sample_its - 10), wall, (float(m.group(1)) if m else None)


def cpu_baseline(case, gpu_iters: float, sample_its: int = 30, serial_its: int = 20):
    """The reference's own QMR on this host, both legal thread counts
    (src/aphid_serial.f90:9-16): qmr_mthread with OMP_NUM_THREADS=4 (the
    reported value) and the serial qmr with OMP_NUM_THREADS=1; or the oracle
    port when the reference harness is not built."""
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    M = case.n_unknowns
    cpu = host_cpu()
    if os.path.exists(harness):
        s4, wall4, q4 = _ref_sample(harness, case, "qmr4", 4, sample_its)
        s1, wall1, q1 = _ref_sample(harness, case, "qmr1", 1, serial_its)
        return {
            "value": 1.0 / (gpu_iters * s4),
            "unit": "converged solves/s",
            "cores": 4,
            "kind": "reference",
            "sample": (f"reference qmr_mthread (OMP_NUM_THREADS=4, OMP_PROC_BIND=close) on the "
                       f"same 129x129x97 system, {sample_its} iterations; {s4:.4f} s/iteration "
                       f"from its qmr.iter clock (iterations 10-{sample_its}), extrapolated to the "
                       f"GPU's {gpu_iters:.0f} iterations per converged solve; harness wall "
                       f"{wall4:.1f} s" + (f", its 'Time spent in QMR' {q4:.2f} s" if q4 else "")),
            "s_per_iteration": s4,
            "ns_per_unknown_iteration": 1e9 * s4 / M,
            "serial": {"value": 1.0 / (gpu_iters * s1), "cores": 1, "s_per_iteration": s1,
                       "ns_per_unknown_iteration": 1e9 * s1 / M,
                       "sample": f"reference serial qmr (OMP_NUM_THREADS=1), {serial_its} "
                                 f"iterations, s/iteration over iterations 10-{serial_its}; "
                                 f"harness wall {wall1:.1f} s"},
            "host_cpu": cpu,
            "host": os.uname().nodename,
        }
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    S = O.System(case)
    out = {"unit": "converged solves/s", "cores": 4, "kind": "port", "host_cpu": cpu}
    for nt, its in ((4, sample_its), (1, serial_its)):
        x, rep, hist = O.qmr(S.op, S.b, S.diag, 0.0, its, nthreads=nt)
        s_it = rep.seconds / its
        d = {"value": 1.0 / (gpu_iters * s_it), "cores": nt, "s_per_iteration": s_it,
             "sample": f"oracle port, {nt} thread(s), {its} iterations, extrapolated to "
                       f"{gpu_iters:.0f} iterations"}
        if nt == 4:
            out.update(d)
        else:
            out["serial"] = d
    return out


def bench_batch():
    """Config-1 family x 64 frequencies x 3 dipoles: batched vs sequential."""
    from aphid_git_amd import gpu, meshgen

    base = meshgen.config1()
    freqs = np.logspace(-2, 5, 64)
    rng = np.random.default_rng(3)
    nx, ny, nz = base.shape
    pos = [(int(rng.integers(4, nx - 4)), int(rng.integers(4, ny - 4))) for _ in range(3)]
    srcs = [(1, pos[0][0], pos[0][1], 9), (2, pos[1][0], pos[1][1], 9), (3, pos[2][0], pos[2][1], 9)]
    single = gpu.Solver(0)
    single.set_mesh(base.xn, base.yn, base.zn)
    bs = [single.build_rhs(5, [q])[0] for q in srcs]
    # one system at a time (warm the tuner and graphs on the first)
    single.set_cells("R", 2 * np.pi * freqs[0], base.cells)
    single.load_rhs(bs[0])
    single.solve_resident(TOL, 20000)
    t0 = time.perf_counter()
    conv_seq, its_seq = 0, []
    for f in freqs:
        single.set_cells("R", 2 * np.pi * f, base.cells)
        for q in range(3):
            single.load_rhs(bs[q])
            r = single.solve_resident(TOL, 20000)
            conv_seq += int(r.status == 0 and r.true_rel_res <= TOL)
            its_seq.append(r.iters)
    t_seq = time.perf_counter() - t0
    single.close()
    B = gpu.BatchSolver(0)
    t0 = time.perf_counter()
    ops = [B.add_operator(base.xn, base.yn, base.zn, "R", 2 * np.pi * f, base.cells) for f in freqs]
    for o in ops:
        for q in range(3):
            B.add_system(o, bs[q])
    t_setup = time.perf_counter() - t0
    B.solve(TOL, 20000)  # warm: graph capture and tiles
    t0 = time.perf_counter()
    reps, secs = B.solve(TOL, 20000)
    t_b = time.perf_counter() - t0
    x_one = [B.solution(q) for q in range(len(reps))]
    B.close()
    conv_b = sum(1 for r in reps if r["status"] == 0 and r["true_rel_res"] <= TOL)
    # the combined path: batched, complex128 (mixed precision loses on this
    # family), and a frequency sweep of L levels: every 2^(L-1)-th frequency
    # from x0 = 0, then each other frequency k from the solutions at k - d
    # and k + d (same dipole), d the lowest set bit of k, with
    # Galerkin-fitted amplitudes on its own operator.  L = 2: even
    # frequencies from zero, odd ones from their neighbours.  The fastest
    # schedule is the combined figure
    def sweep(levels):
        B = gpu.BatchSolver(0)
        ops = [B.add_operator(base.xn, base.yn, base.zn, "R", 2 * np.pi * f, base.cells)
               for f in freqs]
        for o in ops:
            for q in range(3):
                B.add_system(o, bs[q])
        fitted = {}
        for k in range(len(freqs)):
            if k % (1 << (levels - 1)) == 0:
                continue
            d = k & -k
            for q in range(3):
                nb = [k - d] + ([k + d] if k + d < len(freqs) else [])
                B.set_sweep_guess(3 * k + q, [3 * kk + q for kk in nb])
                fitted[3 * k + q] = d
        B.solve(TOL, 20000)  # warm
        t0 = time.perf_counter()
        rs, _ = B.solve(TOL, 20000)
        t = time.perf_counter() - t0
        diff = 0.0
        for q in range(len(rs)):
            a = np.concatenate([v.ravel() for v in B.solution(q)])
            w = np.concatenate([v.ravel() for v in x_one[q]])
            diff = max(diff, float(np.linalg.norm(a - w) / np.linalg.norm(w)))
        B.close()
        conv = sum(1 for r in rs if r["status"] == 0 and r["true_rel_res"] <= TOL)
        odd = [q for q, d in fitted.items() if d == 1]
        row = {"converged": conv, "seconds": t, "solves_per_s": conv / t,
               "speedup_vs_sequential_c128": (conv / t) / (conv_seq / t_seq),
               "speedup_vs_batched_one_wave": (conv / t) / (conv_b / t_b),
               "iters_mean": float(np.mean([r["iters"] for r in rs])),
               "iters_total": int(sum(r["iters"] for r in rs)),
               "iters_fitted_from_neighbours_mean": float(np.mean([rs[q]["iters"] for q in odd])),
               "iters_same_systems_from_zero_mean": float(np.mean([reps[q]["iters"] for q in odd])),
               "r0_fitted_from_neighbours_median": float(np.median([rs[q]["r0_rel"] for q in odd])),
               "max_field_rel_diff_vs_one_wave": diff}
        row["levels"] = levels
        row["by_distance"] = {
            str(d): {"systems": len(qs), "iters_mean": float(np.mean([rs[q]["iters"] for q in qs])),
                     "r0_median": float(np.median([rs[q]["r0_rel"] for q in qs]))}
            for d in sorted(set(fitted.values()))
            for qs in [[q for q, dd in fitted.items() if dd == d]]}
        return row

    sws = {L: sweep(L) for L in (2, 3, 4)}
    best = min(sws.values(), key=lambda r: r["seconds"])
    return {"workload": "33x33x33-node model family (config 1), 64 frequencies 1e-2..1e5 Hz x "
                        "3 unit dipoles (x, y, z) at seeded surface nodes = 192 systems, rel-res 1e-8",
            "systems": len(reps), "converged_batched": conv_b, "converged_sequential": conv_seq,
            "seconds_batched": t_b, "seconds_sequential": t_seq, "seconds_batch_setup": t_setup,
            "solves_per_s_batched": conv_b / t_b, "solves_per_s_sequential": conv_seq / t_seq,
            "speedup": (conv_b / t_b) / (conv_seq / t_seq),
            "iters_mean": float(np.mean([r["iters"] for r in reps])),
            "iters_max": int(max(r["iters"] for r in reps)),
            "iters_mean_sequential": float(np.mean(its_seq)),
            "combined": {
                "what": "batched x complex128 x frequency sweep of %d levels (every %d-th "
                        "frequency from x0 = 0, each other one from a Galerkin fit of the "
                        "solutions d frequencies below and above it, d the lowest set bit of "
                        "its index)" % (best["levels"], 1 << (best["levels"] - 1)),
                **best},
            "sweeps_by_levels": {str(L): r for L, r in sws.items()}}


def bench_batch3b(max_it: int = 10000):
    """48 random-block models, 16 each on 64^3, 96^3 and 128^3 cells, each at
    one log-uniform frequency in 1e-1..1e4 Hz: one packed batched solve vs
    one solve at a time."""
    from aphid_git_amd import gpu, meshgen

    rng = np.random.default_rng(35)
    cases = []
    for n in (64, 96, 128):
        for q in range(16):
            f = float(10 ** rng.uniform(-1, 4))
            cases.append(meshgen.random_block_case((n, n, n), 3500 + 100 * n + q, f))
    m_tot = sum(c.n_unknowns for c in cases)
    single = gpu.Solver(0)
    rhs = []
    for c in cases:
        rhs.append(single.setup_case(c))
    single.load_rhs(rhs[-1])  # the solver holds the last case's mesh
    single.solve_resident(TOL, 50)  # warm the tuner
    seq = []
    t_seq = 0.0  # solve time only: each case's setup is excluded, as for the batch
    for q, (c, b) in enumerate(zip(cases, rhs)):
        single.setup_case(c)
        single.load_rhs(b)
        r = single.solve_resident(TOL, max_it)
        t_seq += r.seconds_total
        seq.append(r)
        log(f"[bench] batch3b sequential {q}: {c.shape} f {c.freq:.3g} iters {r.iters} "
            f"status {r.status} {r.seconds_total:.2f}s")
    single.close()
    B = gpu.BatchSolver(0)
    ops = [B.add_case_operator(c) for c in cases]  # every operator before the systems
    for o, b in zip(ops, rhs):
        B.add_system(o, b)
    B.solve(TOL, 20)  # warm: graph capture and tiles
    t0 = time.perf_counter()
    reps, secs = B.solve(TOL, max_it)
    t_b = time.perf_counter() - t0
    B.close()
    conv_b = sum(1 for r in reps if r["status"] == 0 and r["true_rel_res"] <= TOL)
    conv_s = sum(1 for r in seq if r.status == 0 and r.true_rel_res <= TOL)
    return {"workload": "16 each of 64^3, 96^3, 128^3-cell random-block models (50 m cells), "
                        "one log-uniform frequency each in 1e-1..1e4 Hz, +X surface dipole",
            "systems": len(cases), "unknowns_total": int(m_tot),
            "converged_batched": conv_b, "converged_sequential": conv_s,
            "seconds_batched": t_b, "seconds_sequential": t_seq,
            "speedup": (conv_b / t_b) / max(conv_s / t_seq, 1e-30),
            "iters_by_size": {str(n): [int(r["iters"]) for r in reps[16 * q:16 * q + 16]]
                              for q, n in enumerate((64, 96, 128))}}


def bench_warm5(ntrain=2000, nhold=256, mse_steps=2000, res_steps=1500):
    """The learned warm start's protocol at its cell grid (96x96x80 cells):
    2000 solver-made training solutions, two networks trained on the GPU
    (field MSE against them; the residual of the network's own guess), 256
    held-out models solved from x0 = 0 and from each network's guess.
    About 15 minutes."""
    from aphid_git_amd import warmstart

    out = warmstart.protocol((96, 96, 80), ntrain, nhold, mse_steps, res_steps, log=log)
    out["workload"] = ("random-block models on 96x96x80 cells (97x97x81 nodes), frequencies "
                       "log-uniform 1e-2..1e5 Hz, +X surface dipole, rel-res 1e-8")
    return out


def bench_prec4(freq: float = 1e4, max_it: int = 300000, max_it_mixed: int = 100000,
                stall: int = 1000):
    """Config 4 (192x192x160 cells, 23.9M unknowns, contrast 1e5) at one
    frequency: the complex128 solve to rel-res 1e-8 (capped at max_it) and
    mixed precision (complex64 stencil and vectors, FP64 scalars and
    refinement; inner solves stop on stagnation) capped at max_it_mixed."""
    from aphid_git_amd import gpu, meshgen

    os.environ.setdefault("APHID_PROGRESS", "5000")  # read once, by the first solve
    case = meshgen.config4(freq=freq)
    s = gpu.Solver(0)
    b = s.setup_case(case)
    out = {"workload": f"config4: 192x192x160-cell stretched mesh, 40 blocks, log10 sigma in [-5, 0], "
                       f"+X dipole, {freq:g} Hz, rel-res 1e-8", "unknowns": case.n_unknowns,
           "max_it": max_it, "max_it_mixed": max_it_mixed}
    s.load_rhs(b)
    t0 = time.perf_counter()
    r = s.solve_resident(TOL, max_it)
    out["c128"] = {"seconds": time.perf_counter() - t0, "iters": r.iters, "status": r.status,
                   "true_rel_res": r.true_rel_res}
    log(f"[bench] prec4 c128: {out['c128']}")
    max_it = max_it_mixed
    x128 = np.concatenate([a.ravel() for a in s.fetch_solution()])
    os.environ["APHID_MIXED_STALL"] = str(stall)
    s.load_rhs(b)
    t0 = time.perf_counter()
    r = s.solve_resident(TOL, max_it, gpu.F_MIXED)
    del os.environ["APHID_MIXED_STALL"]
    x = np.concatenate([a.ravel() for a in s.fetch_solution()])
    out["mixed"] = {"seconds": time.perf_counter() - t0, "iters_c64": r.iters,
                    "sweeps": r.outer_iters, "status": r.status, "true_rel_res": r.true_rel_res,
                    "field_rel_diff_vs_c128": float(np.linalg.norm(x - x128) / np.linalg.norm(x128)),
                    "stall_window": stall}
    log(f"[bench] prec4 mixed: {out['mixed']}")
    s.close()
    return out


def bench_prec4_sweep(cap: int = 10000, stall: int = 1000):
    """Config 4 at its 8 frequencies (log-spaced 1e-1..1e4 Hz) with a fixed
    budget of `cap` iterations each: the complex128 solve and mixed
    precision (complex64 inner solves capped at the same total, stagnation
    window `stall`, FP64 refinement), their true residuals, seconds, and
    the field difference between the two."""
    from aphid_git_amd import gpu, meshgen

    s = gpu.Solver(0)
    rows = []
    for f in np.logspace(-1, 4, 8):
        case = meshgen.config4(freq=float(f))
        b = s.setup_case(case)
        s.load_rhs(b)
        t0 = time.perf_counter()
        r = s.solve_resident(TOL, cap)
        t128 = time.perf_counter() - t0
        x128 = np.concatenate([a.ravel() for a in s.fetch_solution()])
        os.environ["APHID_MIXED_STALL"] = str(stall)
        s.load_rhs(b)
        t0 = time.perf_counter()
        rm = s.solve_resident(TOL, cap, gpu.F_MIXED)
        tm = time.perf_counter() - t0
        del os.environ["APHID_MIXED_STALL"]
        xm = np.concatenate([a.ravel() for a in s.fetch_solution()])
        row = {"freq": float(f),
               "c128": {"iters": r.iters, "status": r.status, "true_rel_res": r.true_rel_res,
                        "seconds": t128, "ms_per_iteration": 1e3 * t128 / max(min(r.iters, cap), 1)},
               "mixed": {"iters_c64": rm.iters, "sweeps": rm.outer_iters, "status": rm.status,
                         "true_rel_res": rm.true_rel_res, "seconds": tm,
                         "ms_per_iteration": 1e3 * tm / max(min(rm.iters, cap), 1)},
               "field_rel_diff": float(np.linalg.norm(xm - x128) / np.linalg.norm(x128))}
        rows.append(row)
        log(f"[bench] prec4 sweep {f:.3g} Hz: {row}")
    s.close()
    return {"workload": "config4: 192x192x160-cell stretched mesh, 40 blocks, log10 sigma in "
                        "[-5, 0], +X dipole, 8 frequencies log-spaced 1e-1..1e4 Hz",
            "iteration_cap": cap, "stall_window": stall, "frequencies": rows}


def bench_config4():
    """Config 4 (192x192x160 cells, 23.9M unknowns): the HBM figures of the
    complex128 apply and iteration, and of the complex64 apply, on a live
    Krylov state (300 iterations of the config's own right-hand side)."""
    from aphid_git_amd import gpu, meshgen

    case = meshgen.config4()
    s = gpu.Solver(0)
    b = s.setup_case(case)
    s.load_rhs(b)
    s.solve_resident(TOL, 300)
    t128 = s.time_apply(0, 20)
    kt = s.time_iteration(0, 10)
    t64 = s.time_apply(1, 20)
    tiles = {"c128": list(s.get_tile(0)), "c64": list(s.get_tile(1))}
    s.close()
    M = case.n_unknowns

    def rate(nbytes, t):
        return {"seconds": t, "algorithmic_bytes": nbytes, "GBps": nbytes / t / 1e9,
                "frac": nbytes / t / 1e9 / HBM_PEAK_GBS}

    return {"workload": "config4: 192x192x160-cell stretched mesh, 40 blocks, +X dipole, 1 Hz",
            "unknowns": M, "tiles": tiles,
            "apply_c128": rate(3 * M * 16, t128), "iteration_c128": rate(ITERATION_PASSES * M * 16, kt[4]),
            "apply_c64": rate(3 * M * 8, t64)}


def bench_shard(nshard: int = 8):
    """One GPU's share of config 2 under the x-slab decomposition of a mesh
    too large for one GPU: an interior slab of ceil(129/nshard) x-planes
    plus one halo plane per side (129/8 -> 17 + 2 = 19 planes), as its own
    mesh, with the same y/z extent and cells.  Times the apply and one QMR
    iteration on a live Krylov state (a capped solve of a random b).  The
    halo exchange and the scalar all-reduces are not run (one GPU)."""
    from aphid_git_amd import gpu, meshgen

    base = meshgen.config2()
    nx = len(base.xn)
    w = -(-nx // nshard) + 2
    i0 = (nx - w) // 2
    cells = {k: np.asfortranarray(v[i0:i0 + w - 1]) for k, v in base.cells.items()}
    case = meshgen.Case(base.xn[i0:i0 + w].copy(), base.yn, base.zn, base.model, cells, base.freq,
                        5, [], base.target, base.max_it, "config-2 x-slab")
    s = gpu.Solver(0)
    s.setup_case(case)
    rng = np.random.default_rng(11)
    b = [np.asfortranarray(rng.standard_normal(sh) + 1j * rng.standard_normal(sh))
         for sh in gpu.comp_shapes(*case.shape)]
    s.load_rhs(b)
    s.solve_resident(1e-12, 300)
    t_apply = s.time_apply(0, 20)
    kt = s.time_iteration(0, 10)
    tile = s.get_tile(0)
    s.close()
    nxs, ny, nz = case.shape
    M = 4 * nxs * ny * nz  # stored entries (the slab's unknowns plus padding)
    halo = 2 * 4 * ny * nz * 16
    return {"workload": f"config-2 x-slab {nxs}x{ny}x{nz} nodes (1/{nshard} of 129 x-planes + "
                        f"1 halo plane per side), complex128",
            "entries": M, "tile": list(tile),
            "apply_us": t_apply * 1e6, "apply_GBps": 3 * M * 16 / t_apply / 1e9,
            "iteration_us": kt[4] * 1e6, "iteration_GBps": ITERATION_PASSES * M * 16 / kt[4] / 1e9,
            "halo_exchange_bytes_per_apply": halo,
            "full_mesh_iteration_share_us": None}


def bench_unet(reps: int = 5):
    from aphid_git_amd import gpu

    net = gpu.UNet((96, 96, 80), seed=0)
    f, st = net.time(reps)
    F = net.forward_flops
    net.close()
    return {"workload": "3-level 3-D U-Net 16/32/64 channels on a 96x96x80 cell grid, bf16 "
                        "weights/activations, FP32 accumulation, batch 1",
            "params": net.n_params, "forward_gflop": F / 1e9, "forward_ms": 1e3 * f,
            "forward_tflops": F / f / 1e12, "train_step_ms": 1e3 * st,
            "train_step_tflops": 3 * F / st / 1e12,
            "roofline": {"bound": "mfma", "achieved": F / f / 1e12, "peak": BF16_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": F / f / 1e12 / BF16_PEAK_TFLOPS}}


def bench_warm(ncell=(16, 16, 16), ntrain=64, nhold=16, mse_steps=400, res_steps=400):
    """The learned warm start's protocol on a small grid (seconds): what the
    trained network gains over x0 = 0 on held-out models."""
    from aphid_git_amd import warmstart

    out = warmstart.protocol(ncell, ntrain, nhold, mse_steps, res_steps)
    out["workload"] = (f"random-block models on {ncell[0]}x{ncell[1]}x{ncell[2]} cells, "
                       f"frequencies log-uniform 1e-2..1e5 Hz, +X surface dipole, rel-res 1e-8")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-its", type=int, default=30)
    ap.add_argument("--max-it", type=int, default=MAX_IT,
                    help="iteration cap per solve (profiling runs use a small cap; "
                         "a capped solve does not count as converged)")
    ap.add_argument("--parts", default=",".join(PARTS),
                    help="comma-separated subset of " + ",".join(PARTS + EXTRA_PARTS))
    args = ap.parse_args()
    parts = set(args.parts.split(","))
    if args.no_cpu:
        parts.discard("cpu")
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); systems of this solve "
              f"shard one per GPU with no communication, so N GPUs scale the single-GPU rate.",
              file=sys.stderr)
        sys.exit(2)

    from aphid_git_amd import gpu, meshgen

    case = meshgen.config2()
    M = case.n_unknowns
    nx, ny, nz = case.shape
    out = {
        "metric": "converged solves/s at rel-res 1e-8 (128x128x96 grid); matvec HBM GB/s vs peak",
        "value": None,
        "unit": "converged solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex128 (f64)",
        "data": "synthetic (seeded block model, config2)",
        "config": {"workload": "config2: 128x128x96-cell stretched mesh, 64x64x48 core at 25 m, "
                               "20 random blocks, +X dipole, 1 Hz, Jacobi-QMR to rel-res 1e-8",
                   "nodes": [nx, ny, nz], "unknowns": M, "tol": TOL,
                   "value_definition": "converged solves per second of wall time over the timed "
                                       "steps, each a complete solve from x0 = 0 with the "
                                       "operator and b resident in HBM (aphid_solve_resident); "
                                       "converged = true ||b - K x|| / ||b|| <= 1e-8. "
                                       "end_to_end adds setup, upload and download"},
    }
    iters = None
    if "solve" in parts:
        log(f"[bench] config2 {case.shape} nodes, M={M}")
        s = gpu.Solver(0)
        b = s.setup_case(case)
        s.load_rhs(b)
        for w in range(args.warmup):
            r = s.solve_resident(TOL, args.max_it)
            log(f"[bench] warmup {w}: status {r.status} iters {r.iters} true {r.true_rel_res:.2e} "
                f"{r.seconds_qmr:.3f}s")
        results = []
        t0 = time.perf_counter()  # each solve returns after its stream has drained
        for k in range(args.steps):
            r = s.solve_resident(TOL, args.max_it)
            results.append(r)
            log(f"[bench] step {k}: status {r.status} iters {r.iters} true {r.true_rel_res:.2e} "
                f"{r.seconds_qmr:.3f}s")
        elapsed = time.perf_counter() - t0
        converged = sum(1 for r in results if r.status == 0 and r.true_rel_res <= TOL)
        iters = float(np.mean([r.iters for r in results]))

        # kernel timing (HIP events on the library's stream)
        t_apply = s.time_apply(0, 20)
        t_traffic = s.time_traffic(20)
        kt = s.time_iteration(0, 10)
        tile = s.get_tile(0)
        k2_compact = s.set_compact_k2(True)  # the default; reports whether it is in use
        s_bytes = 16
        B_mv = 3 * M * s_bytes
        per_kernel = {
            "p_update": (PUPDATE_PASSES * M * s_bytes, kt[0]),
            "stencil_eps": (B_mv, kt[1]),
            "vt_norm": (2 * M * s_bytes, kt[2]),
            "fused_update": (UPDATE_PASSES * M * s_bytes, kt[3]),
        }
        kernels = {k: {"algorithmic_bytes": int(bb), "seconds": t, "GBps": bb / t / 1e9,
                       "frac": bb / t / 1e9 / HBM_PEAK_GBS} for k, (bb, t) in per_kernel.items()}
        dom = max(per_kernel, key=lambda k: per_kernel[k][1])
        copy_bytes = float(1 << 30)
        t_copy = s.time_copy(copy_bytes, 10)
        copy_probe = {"bytes_read": copy_bytes, "bytes_written": copy_bytes, "seconds": t_copy,
                      "GBps": 2 * copy_bytes / t_copy / 1e9,
                      "frac_of_peak": 2 * copy_bytes / t_copy / 1e9 / HBM_PEAK_GBS}
        it_bytes = ITERATION_PASSES * M * s_bytes
        traffic = None
        tf = os.path.join(ROOT, "profiles", "traffic_config2.json")
        if os.path.exists(tf):
            with open(tf) as f:
                tj = json.load(f)
            # PMC traffic of the same kernel and tile only
            if list(tj.get("tile", [])) == list(tile):
                traffic = tj.get("stencil_bytes_per_launch")
        # the roofline entry times the stencil kernel as the iteration
        # launches it (with the eps = p^T K p epilogue, inside the iteration's
        # back-to-back kernel sequence, without the one-block finalize kernel
        # that follows it: the duration rocprofv3 reports for k_apply); the
        # plain apply alone is reported beside it
        t_st = kt[1]
        out.update({
            "value": converged / elapsed,
            "ms_per_step": 1e3 * elapsed / args.steps,
            "roofline": {"bound": "hbm", "kernel": "stencil (k_apply with the eps epilogue, "
                                                   "in the QMR iteration; finalize excluded)",
                         "achieved": B_mv / t_st / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": B_mv / t_st / 1e9 / HBM_PEAK_GBS, "traffic": traffic,
                         "traffic_source": "profiles/traffic_config2.json: rocprofv3 --pmc "
                                           "FETCH_SIZE and WRITE_SIZE of the same kernel and "
                                           "tile, scaled by a copy kernel's counts",
                         "algorithmic_bytes_per_launch": B_mv, "seconds_per_launch": t_st,
                         "plain_apply_seconds": t_apply,
                         "plain_apply_frac": B_mv / t_apply / 1e9 / HBM_PEAK_GBS,
                         # the same tile's loads and stores without arithmetic,
                         # repeated like the plain apply: the floor its access
                         # pattern sets
                         "pattern_ceiling": {
                             "seconds": t_traffic,
                             "GBps": B_mv / t_traffic / 1e9,
                             "frac_of_peak": B_mv / t_traffic / 1e9 / HBM_PEAK_GBS,
                             "plain_apply_over_ceiling": t_traffic / t_apply,
                             "definition": "aphid_time_traffic: the tuned march tile's loads "
                                           "(own columns of the 4 fields and of k^2 as the "
                                           "stencil reads it, halo cells) and 4 stores per "
                                           "node, no LDS, barriers or arithmetic, same "
                                           "occupancy"},
                         "tile": {"tk": tile[0], "tj": tile[1], "chunk": tile[2]},
                         "k2_compact": k2_compact,
                         "k2_compact_note": "with the compact k^2 (model R) the stencil reads "
                                            "33 bytes of k^2 per node instead of 64: the "
                                            "imaginary parts of k2x, k2p, k2y, k2z as two "
                                            "16-byte pairs, a one-byte offset of k2p's real "
                                            "part from a common bit pattern, and the edge "
                                            "grids' real parts from 2-D tables (the same "
                                            "values, checked bit for bit), so it moves about "
                                            "2.52*M*16 of the 3*M*16 algorithmic bytes; "
                                            "achieved and frac count the algorithmic bytes, "
                                            "traffic is what the counters measured"},
            "roofline_dominant_kernel": {"kernel": dom, **kernels[dom], "peak": HBM_PEAK_GBS},
            "kernels": kernels,
            "iteration": {"seconds": kt[4], "algorithmic_bytes": it_bytes,
                          "GBps": it_bytes / kt[4] / 1e9,
                          "frac": it_bytes / kt[4] / 1e9 / HBM_PEAK_GBS},
            "hbm_copy_probe": copy_probe,
        })
        out["config"].update({"iterations_per_solve": iters, "converged_steps": converged,
                              "true_rel_res_max": max(r.true_rel_res for r in results)})
        # end to end (not `value`): host download of the solution, then a
        # steady-state re-setup of the case (metrics, k^2, diag, RHS) and
        # the host upload of b, beside the resident solve's wall time
        t0 = time.perf_counter()
        s.fetch_solution()
        t_d2h = time.perf_counter() - t0
        t0 = time.perf_counter()
        b2 = s.setup_case(case)
        t_setup = time.perf_counter() - t0
        t0 = time.perf_counter()
        s.load_rhs(b2)
        t_h2d = time.perf_counter() - t0
        t_solve = elapsed / max(args.steps, 1)
        total = t_setup + t_h2d + t_solve + t_d2h
        out["end_to_end"] = {"setup_s": t_setup, "h2d_s": t_h2d, "solve_s": t_solve, "d2h_s": t_d2h,
                             "total_s": total,
                             "converged_solves_per_s": (converged / max(args.steps, 1)) / total,
                             "note": "host arrays in Fortran order; permutation to the device "
                                     "layout on the device; not part of `value`"}
        log(f"[bench] end to end: {out['end_to_end']}")
        s.close()
    if "config4" in parts:
        log("[bench] config 4 apply and iteration ...")
        out["config4_kernels"] = bench_config4()
        log(f"[bench] config4: {out['config4_kernels']}")
    if "shard" in parts:
        log("[bench] one GPU's share of config 2 (x-slab of 8) ...")
        out["shard_config2_1of8"] = bench_shard()
        if "iteration" in out:
            out["shard_config2_1of8"]["full_mesh_iteration_share_us"] = out["iteration"]["seconds"] * 1e6 / 8
        log(f"[bench] shard: {out['shard_config2_1of8']}")
    if "batch" in parts:
        log("[bench] batch: 192 systems of the config-1 family ...")
        out["batch_config3a"] = bench_batch()
        log(f"[bench] batch: {out['batch_config3a']}")
    if "prec4" in parts:
        log("[bench] precision study on config 4 ...")
        out["precision_config4"] = bench_prec4()
    if "prec4sweep" in parts:
        log("[bench] precision study on config 4, 8 frequencies ...")
        out["precision_config4_sweep"] = bench_prec4_sweep()
    if "warm5" in parts:
        log("[bench] learned warm start, 96x96x80 cells ...")
        out["warm_start_96x96x80"] = bench_warm5()
        log(f"[bench] warm5: {out['warm_start_96x96x80']}")
    if "batch3b" in parts:
        log("[bench] batch: 48 uneven random-block systems ...")
        out["batch_config3b"] = bench_batch3b()
        log(f"[bench] batch3b: {out['batch_config3b']}")
    if "unet" in parts:
        log("[bench] warm-start network at 96x96x80 cells ...")
        out["unet"] = bench_unet()
        log(f"[bench] unet: {out['unet']}")
    if "warm" in parts:
        log("[bench] learned warm start, 16^3 cells ...")
        out["warm_start"] = bench_warm()
        log(f"[bench] warm: {out['warm_start']}")
    cpu = None
    if "cpu" in parts and iters is not None:
        log("[bench] CPU baseline (reference, 4 threads and serial) ...")
        try:
            cpu = cpu_baseline(case, iters, args.cpu_its)
        except Exception as e:  # reported, never fatal
            cpu = {"error": str(e)[:300]}
        log(f"[bench] cpu: {cpu}")
    out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
