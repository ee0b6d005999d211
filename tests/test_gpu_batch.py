"""Batched solve (many systems per launch) against single-system solves and
the CPU oracle: an uneven pack of meshes, shared operators, per-system
convergence and warm restarts."""
import numpy as np
import pytest

from aphid_git_amd import meshgen
from conftest import flat4

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu_mod():
    from aphid_git_amd import gpu as G

    return G


def _uneven_cases():
    cs = [meshgen.random_block_case((14, 12, 10), seed=1, freq=30.0),
          meshgen.random_block_case((20, 18, 16), seed=2, freq=3.0),
          meshgen.random_block_case((10, 22, 12), seed=3, freq=300.0)]
    out = []
    for c in cs:
        nx, ny, nz = c.shape
        k0 = c.sources[0][3]
        out.append((c, [(1, nx // 2, ny // 2, k0), (2, nx // 2 + 1, ny // 2 - 1, k0)]))
    return out


def test_uneven_pack_matches_oracle_and_single_solves(gpu_mod, oracle):
    cases = _uneven_cases()
    single = gpu_mod.Solver()
    with gpu_mod.BatchSolver() as B:
        want = []
        for c, srcs in cases:
            op = B.add_case_operator(c)
            single.set_mesh(c.xn, c.yn, c.zn)
            for src in srcs:
                b, _ = single.build_rhs(5, [src])
                want.append((c, src, b))
        sys_ids = []
        opi = 0
        for c, srcs in cases:
            for src in srcs:
                pass
        # systems in operator order
        k = 0
        for o, (c, srcs) in enumerate(cases):
            for src in srcs:
                sys_ids.append(B.add_system(o, want[k][2]))
                k += 1
        reps, secs = B.solve(1e-9, 5000)
        print(f"[batch] {len(reps)} systems in {secs:.3f}s; iters "
              f"{[r['iters'] for r in reps]}", flush=True)
        for q, (c, src, b) in enumerate(want):
            r = reps[q]
            assert r["status"] == 0 and r["true_rel_res"] < 1.05e-9
            x = flat4(B.solution(q))
            S = oracle.System(meshgen.Case(c.xn, c.yn, c.zn, c.model, c.cells, c.freq, 5, [src]))
            xo, rep, _ = S.solve(tol=1e-9, max_it=5000)
            xo = flat4(xo)
            assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-7
            single.set_mesh(c.xn, c.yn, c.zn)
            single.set_cells(c.model, c.omega, c.cells)
            rs = single.qmr_solve(b, 1e-9, 5000, hist_cap=0)
            assert abs(r["iters"] - rs.iters) <= max(3, 0.05 * rs.iters)
    single.close()


@pytest.mark.parametrize("max_it", [5, 6])
def test_batched_iterate_after_odd_and_even_iteration_counts(gpu_mod, oracle, max_it):
    """The batched update stores x and r on even iterations only and the
    solve completes them for systems that stopped after an odd one: the
    iterate after max_it iterations is the oracle's (a missing or doubled
    d_k would be an O(1) error)."""
    c = _uneven_cases()[1][0]
    with gpu_mod.BatchSolver() as B:
        single = gpu_mod.Solver()
        single.set_mesh(c.xn, c.yn, c.zn)
        b, _ = single.build_rhs(5, c.sources)
        single.close()
        o = B.add_case_operator(c)
        B.add_system(o, b)
        reps, _ = B.solve(1e-30, max_it)
        assert reps[0]["status"] == 1
        S = oracle.System(c)
        xo, rep, _ = S.solve(tol=1e-30, max_it=max_it)
        xo = flat4(xo)
        assert np.linalg.norm(flat4(B.solution(0)) - xo) / np.linalg.norm(xo) < 1e-11


def test_both_batched_stencils_match_the_oracle(gpu_mod, oracle, monkeypatch):
    """The x-march over flat chunks (default) and the one-node-per-thread
    batched stencil (APHID_BATCH_DIRECT=1, the fallback for long z-rows)
    solve the same uneven pack to the oracle's fields."""
    cases = _uneven_cases()
    sols = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("APHID_BATCH_DIRECT", mode)
        with gpu_mod.BatchSolver() as B:
            single = gpu_mod.Solver()
            rhs = []
            for o, (c, srcs) in enumerate(cases):
                B.add_case_operator(c)  # every operator before the first system
                single.set_mesh(c.xn, c.yn, c.zn)
                rhs += [(o, single.build_rhs(5, [src])[0]) for src in srcs]
            single.close()
            for o, b in rhs:
                B.add_system(o, b)
            k = len(rhs)
            reps, _ = B.solve(1e-9, 5000)
            assert all(r["status"] == 0 and r["true_rel_res"] < 1.05e-9 for r in reps)
            sols[mode] = [flat4(B.solution(q)) for q in range(k)]
    q = 0
    for c, srcs in cases:
        for src in srcs:
            S = oracle.System(meshgen.Case(c.xn, c.yn, c.zn, c.model, c.cells, c.freq, 5, [src]))
            xo = flat4(S.solve(tol=1e-9, max_it=5000)[0])
            for mode in ("0", "1"):
                assert np.linalg.norm(sols[mode][q] - xo) / np.linalg.norm(xo) < 1e-7
            q += 1


@pytest.mark.parametrize("nzc,seed", [(50, 11), (100, 12), (180, 13), (240, 14)])
def test_batched_stencil_window_strides_match_the_oracle(gpu_mod, oracle, nzc, seed):
    """The batched x-march is compiled for four LDS window strides, picked by
    the batch's longest z-row (up to 40, 72, 136 and 228 nodes). Meshes of
    51, 101 and 181 z-nodes take the last three; 241 z-nodes is past the
    largest window and takes the one-node-per-thread stencil. Each batch
    solves to the oracle's fields."""
    c = meshgen.random_block_case((5, 4, nzc), seed=seed, freq=30.0)
    nx, ny, nz = c.shape
    k0 = c.sources[0][3]
    srcs = [(1, nx // 2, ny // 2, k0), (3, nx // 2, ny // 2 + 1, k0 + 2)]
    single = gpu_mod.Solver()
    single.set_mesh(c.xn, c.yn, c.zn)
    rhs = [single.build_rhs(5, [src])[0] for src in srcs]
    single.close()
    with gpu_mod.BatchSolver() as B:
        op = B.add_case_operator(c)
        for b in rhs:
            B.add_system(op, b)
        reps, _ = B.solve(1e-9, 5000)
        assert all(r["status"] == 0 and r["true_rel_res"] < 1.05e-9 for r in reps)
        sols = [flat4(B.solution(q)) for q in range(len(srcs))]
    for src, xs in zip(srcs, sols):
        S = oracle.System(meshgen.Case(c.xn, c.yn, c.zn, c.model, c.cells, c.freq, 5, [src]))
        xo = flat4(S.solve(tol=1e-9, max_it=5000)[0])
        assert np.linalg.norm(xs - xo) / np.linalg.norm(xo) < 1e-7


def test_shared_operators_frequency_sweep_and_warm_restart(gpu_mod):
    """config-1 family: 3 frequencies x 3 dipoles (+X, +Y, +Z at the surface
    centre) share one operator per frequency; then a warm restart from the
    converged solutions stops at once."""
    base = meshgen.config1()
    freqs = [1.0, 10.0, 100.0]
    srcs = [(1, 17, 17, 9), (2, 17, 17, 9), (3, 17, 17, 9)]
    single = gpu_mod.Solver()
    single.set_mesh(base.xn, base.yn, base.zn)
    with gpu_mod.BatchSolver() as B:
        ops = [B.add_operator(base.xn, base.yn, base.zn, "R", 2 * np.pi * f, base.cells) for f in freqs]
        bs = []
        for o in ops:
            for src in srcs:
                b, _ = single.build_rhs(5, [src])
                bs.append(b)
                B.add_system(o, b)
        reps, secs = B.solve(1e-8, 5000)
        print(f"[batch sweep] 9 systems {secs:.3f}s iters {[r['iters'] for r in reps]}", flush=True)
        for q, r in enumerate(reps):
            assert r["status"] == 0 and r["true_rel_res"] < 1e-8
            f = freqs[q // 3]
            single.set_cells("R", 2 * np.pi * f, base.cells)
            rs = single.qmr_solve(bs[q], 1e-8, 5000, hist_cap=0)
            x, xs = flat4(B.solution(q)), flat4(rs.x)
            assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-6
        reps2, _ = B.solve(1e-8, 5000, flags=gpu_mod.F_X0)
        assert all(r["status"] == 0 and r["iters"] <= 2 for r in reps2)
        # a neighbouring frequency's solution is a valid (if poorer) guess
        B.set_guess(3, 0)
        B.set_guess(4, 1)
        reps3, _ = B.solve(1e-8, 5000, flags=gpu_mod.F_X0)
        assert all(r["status"] == 0 and r["true_rel_res"] < 1e-8 for r in reps3)
    single.close()


def test_batched_mixed_precision_meets_fp64_accuracy(gpu_mod):
    """complex64 inner batch + per-system FP64 refinement reaches rel-res 1e-8
    on the FP64 operator, and its fields are as close to a tight (1e-11)
    complex128 solve as the plain complex128 path at 1e-8 is.  Two solutions
    at residual 1e-8 differ by up to cond(K)*1e-8 (a few 1e-6 at 0.1 Hz), so
    the bar is the plain path's own error, stated per system."""
    base = meshgen.config1()
    freqs = [0.1, 10.0, 1000.0]
    srcs = [(1, 17, 17, 9), (3, 12, 20, 9)]
    single = gpu_mod.Solver()
    single.set_mesh(base.xn, base.yn, base.zn)
    bs = [single.build_rhs(5, [s])[0] for s in srcs]
    single.close()
    sols = {}
    for key, tol, flags in (("tight", 1e-11, 0), ("c128", 1e-8, 0), ("mixed", 1e-8, gpu_mod.F_MIXED)):
        with gpu_mod.BatchSolver() as B:
            ops = [B.add_operator(base.xn, base.yn, base.zn, "R", 2 * np.pi * f, base.cells)
                   for f in freqs]
            for o in ops:
                for b in bs:
                    B.add_system(o, b)
            reps, secs = B.solve(tol, 20000, flags=flags)
            res = ["%.1e" % r["true_rel_res"] for r in reps]
            print(f"[batch {key}] {secs:.3f}s iters {[r['iters'] for r in reps]} true {res}",
                  flush=True)
            assert all(r["status"] == 0 and r["true_rel_res"] < tol for r in reps)
            sols[key] = [flat4(B.solution(q)) for q in range(len(reps))]
    for q, ref in enumerate(sols["tight"]):
        e128 = np.linalg.norm(sols["c128"][q] - ref) / np.linalg.norm(ref)
        emix = np.linalg.norm(sols["mixed"][q] - ref) / np.linalg.norm(ref)
        print(f"  system {q}: field error vs tight FP64 solve: c128 {e128:.1e}  mixed {emix:.1e}")
        assert emix <= max(1e-6, 3.0 * e128)


def test_large_systems_take_the_single_path_in_place(gpu_mod, monkeypatch):
    """Systems at or above APHID_BATCH_SEQ_MIN unknowns are solved one at a
    time on the single-system path, in place on the batch's vectors; the
    rest share the batched launches.  Two large systems share each operator
    (the captured iteration graph follows the system's vectors).  Large
    systems reproduce Solver.qmr_solve exactly (same kernels and launch
    shapes); small ones match it as the batched path does; a warm restart
    from the batch's own solutions re-converges within a few iterations on
    both paths."""
    base = meshgen.config1()                      # 33^3 nodes: 4 nn = 143,748
    monkeypatch.setenv("APHID_BATCH_SEQ_MIN", "100000")
    small = _uneven_cases()[:2]                   # 4 nn < 100,000: batched
    freqs = [3.0, 300.0]
    srcs = [(1, 17, 17, 9), (3, 12, 20, 9)]
    single = gpu_mod.Solver()
    with gpu_mod.BatchSolver() as B:
        plan = []                                 # (case-like, omega, model, cells, b)
        ops_big = [B.add_operator(base.xn, base.yn, base.zn, "R", 2 * np.pi * f, base.cells)
                   for f in freqs]
        ops_small = [B.add_case_operator(c) for c, _ in small]
        single.set_mesh(base.xn, base.yn, base.zn)
        for o, f in zip(ops_big, freqs):
            for src in srcs:
                b, _ = single.build_rhs(5, [src])
                B.add_system(o, b)
                plan.append((base, 2 * np.pi * f, "R", base.cells, b, True))
        for o, (c, ss) in zip(ops_small, small):
            single.set_mesh(c.xn, c.yn, c.zn)
            b, _ = single.build_rhs(5, [ss[0]])
            B.add_system(o, b)
            plan.append((c, c.omega, c.model, c.cells, b, False))
        reps, secs = B.solve(1e-9, 5000)
        print(f"[batch split] {len(reps)} systems {secs:.3f}s iters {[r['iters'] for r in reps]}",
              flush=True)
        for q, (c, om, model, cells, b, big) in enumerate(plan):
            r = reps[q]
            assert r["status"] == 0 and r["true_rel_res"] < 1.05e-9, (q, r)
            single.set_mesh(c.xn, c.yn, c.zn)
            single.set_cells(model, om, cells)
            rs = single.qmr_solve(b, 1e-9, 5000, hist_cap=0)
            x, xs = flat4(B.solution(q)), flat4(rs.x)
            err = np.linalg.norm(x - xs) / np.linalg.norm(xs)
            if big:
                assert r["iters"] == rs.iters and err <= 1e-13, (q, r["iters"], rs.iters, err)
            else:
                assert abs(r["iters"] - rs.iters) <= max(3, 0.05 * rs.iters) and err < 1e-6
        # a first solve stops on the recurrence residual, so the true residual
        # it leaves may sit just above the target: a restart then needs a few
        # iterations, far fewer than from zero
        reps2, _ = B.solve(1e-9, 5000, flags=gpu_mod.F_X0)
        assert all(r["status"] == 0 and r["true_rel_res"] < 1e-9 and r["iters"] <= 10
                   for r in reps2), reps2
    single.close()
