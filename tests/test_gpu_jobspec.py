"""Multi-frequency / multi-source jobs (the reference's MPI farm,
src/aphid_parallel.f90) as one batched solve: every system equals its
single-system solve, and the per-process directories are written."""
import os

import numpy as np
import pytest

from aphid_git_amd import gpu, jobspec, meshgen

pytestmark = pytest.mark.gpu


def _job(tmp_path, stype, freqs, arrays):
    case = meshgen.random_block_case((12, 12, 12), 21, 1.0)
    head = ["=" * 20, "test job", "=" * 20]
    job = jobspec.Job(stype, 1e-10, 20000, "R", ["R        !! conductivity model"], head,
                      freqs, arrays, case.xn, case.yn, case.zn, case.cells)
    d = str(tmp_path / stype)
    jobspec.write_job(job, d)
    return d, job


@pytest.mark.parametrize("stype", ["CS", "MT"])
def test_job_systems_match_single_solves(tmp_path, stype):
    arrays = [[(1, 7, 7, 4)], [(2, 6, 7, 4), (2, 6, 8, 4)]] if stype == "CS" else []
    freqs = [0.3, 30.0, 3000.0] if stype == "CS" else [1.0, 100.0]
    d, job = _job(tmp_path, stype, freqs, arrays)
    reports, secs, wall = jobspec.run_job(d)
    systems = job.systems()
    assert len(reports) == len(systems)
    s = gpu.Solver(0)
    s.set_mesh(job.xn, job.yn, job.zn)
    for p, (mode, f, dips) in enumerate(systems):
        assert reports[p]["status"] == 0 and reports[p]["true_rel_res"] <= 1e-10
        td = os.path.join(d, jobspec.tmp_dir_name(p, len(systems)))
        raw = np.fromfile(os.path.join(td, "aphid_solution.bin"), np.complex128)
        s.set_cells("R", 2 * np.pi * f, job.cells)
        b, _ = s.build_rhs(mode, dips)
        s.load_rhs(b)
        r = s.solve_resident(1e-10, 20000)
        assert r.status == 0
        x = np.concatenate([a.ravel(order="F") for a in s.fetch_solution()])
        err = np.linalg.norm(raw - x) / np.linalg.norm(x)
        assert err < 1e-6, (p, err)
        lines = open(os.path.join(td, "aphid.input")).read().splitlines()
        assert int(lines[3].split()[0]) == mode
        assert float(lines[6].split()[0]) == pytest.approx(f, rel=1e-6)
    s.close()


@pytest.mark.parametrize("sweep", [False, True])
@pytest.mark.parametrize("stype", ["CS", "MT"])
def test_job_systems_match_the_oracle(tmp_path, oracle, stype, sweep):
    """Every system of a job, solved by the batched HIP path, against the
    CPU oracle's complex128 QMR on the same (mode, frequency, source): the
    MT plane-wave modes 2 and 3 (src/aphid.f90:588-636) and CS dipole arrays
    (mode 5), 12^3 cells, rel-res 1e-10 on both sides.  Fields within 1e-7
    (relative 2-norm over all unknowns).  With `sweep` the frequencies of
    a source (in ascending order) are solved in the waves of
    jobspec.sweep_guesses, each fitted one from solutions at neighbouring
    frequencies: its initial guess is better than x0 = 0 and its fields
    meet the same bar.  Each tmp-N/qmr.iter holds one line per iteration
    (src/aphid_qmr.f90:381); from x0 = 0 its first 20 residual norms equal
    the oracle's history to 1e-6 and its length is within 10% of the
    oracle's iteration count."""
    arrays = [[(1, 7, 7, 4)], [(-3, 6, 7, 4), (2, 6, 8, 4)]] if stype == "CS" else []
    freqs = [30.0, 0.3, 3000.0, 3.0] if stype == "CS" else [1.0, 100.0, 1e4, 10.0]
    d, job = _job(tmp_path, stype, freqs, arrays)
    reports, _, _ = jobspec.run_job(d, sweep=sweep)
    second = jobspec.sweep_guesses(job.freqs, job.n_sources) if sweep else {}
    for p in second:
        assert reports[p]["r0_rel"] < 1.0, (p, reports[p])
    systems = job.systems()
    assert len(reports) == len(systems)
    for p, (mode, f, dips) in enumerate(systems):
        assert reports[p]["status"] == 0
        td = os.path.join(d, jobspec.tmp_dir_name(p, len(systems)))
        x = np.fromfile(os.path.join(td, "aphid_solution.bin"), np.complex128)
        case = meshgen.Case(job.xn, job.yn, job.zn, job.model, job.cells, f, mode, list(dips),
                            job.tol, job.max_it)
        xo, rep, ohist = oracle.System(case).solve(tol=job.tol, max_it=job.max_it)
        assert rep.status == 0
        it = np.loadtxt(os.path.join(td, "qmr.iter"), ndmin=2)
        assert it.shape == (reports[p]["iters"], 3)
        assert np.array_equal(it[:, 0], np.arange(1, len(it) + 1))
        assert np.all(np.diff(it[:, 2]) >= 0)
        if p not in second:
            m = min(20, len(ohist), len(it))
            rel = np.abs(it[:m, 1] - ohist[:m]) / ohist[:m]
            assert rel.max() < 1e-6, (p, rel.max())
            assert abs(len(it) - rep.iters) <= max(3, 0.1 * rep.iters), (len(it), rep.iters)
        xo = np.concatenate([a.ravel(order="F") for a in xo])
        err = np.linalg.norm(x - xo) / np.linalg.norm(xo)
        print(f"[job {stype} sweep={sweep}] system {p} mode {mode} f {f:g} Hz: iters "
              f"{reports[p]['iters']} (oracle {rep.iters}), r0 {reports[p]['r0_rel']:.2e}, "
              f"field err {err:.2e}", flush=True)
        assert err < 1e-7, (p, err)


PAR_EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "aphid_git_amd", "aphid_gpu_parallel")


@pytest.mark.parametrize("sweep", [False, True])
@pytest.mark.parametrize("stype", ["CS", "MT"])
def test_fortran_job_driver_matches_run_job(tmp_path, stype, sweep):
    """aphid_gpu_parallel (Fortran over iso_c_binding) writes the same
    per-process directories as run_job: the same batched solve, so the same
    iteration counts and the same solutions; with -s both run the same
    frequency sweep."""
    import subprocess
    if not os.path.exists(PAR_EXE):
        pytest.fail("aphid_gpu_parallel is not built")
    arrays = [[(1, 7, 7, 4)], [(-3, 6, 7, 4), (2, 6, 8, 4)]] if stype == "CS" else []
    freqs = [0.5, 50.0] if stype == "CS" else [2.0, 200.0, 2e4]
    d, job = _job(tmp_path, stype, freqs, arrays)
    py_out = str(tmp_path / "py")
    reports, _, _ = jobspec.run_job(d, out_dir=py_out, sweep=sweep)
    r = subprocess.run([PAR_EXE, "-w", d] + (["-s"] if sweep else []), capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    n = len(job.systems())
    assert f"Solved {n:6d} systems" in r.stdout
    for p, (mode, f, dips) in enumerate(job.systems()):
        name = jobspec.tmp_dir_name(p, n)
        fd, pd = os.path.join(d, name), os.path.join(py_out, name)
        xf = np.fromfile(os.path.join(fd, "aphid_solution.bin"), np.complex128)
        xp = np.fromfile(os.path.join(pd, "aphid_solution.bin"), np.complex128)
        assert xf.shape == xp.shape
        assert np.linalg.norm(xf - xp) <= 1e-12 * np.linalg.norm(xp), p
        itf = np.loadtxt(os.path.join(fd, "qmr.iter"), ndmin=2)
        itp = np.loadtxt(os.path.join(pd, "qmr.iter"), ndmin=2)
        assert itf.shape == itp.shape == (reports[p]["iters"], 3)
        assert np.array_equal(itf[:, 0], np.arange(1, len(itf) + 1))
        assert np.allclose(itf[:, 1], itp[:, 1], rtol=1e-9, atol=0)
        assert itf[-1, 1] < job.tol * reports[p]["rhsnorm"]
        assert np.all(np.diff(itf[:, 2]) >= 0) and np.all(itf[:, 2] >= 0)
        inp = open(os.path.join(fd, "aphid.input")).read().splitlines()
        assert int(inp[3].split()[0]) == mode
        assert float(inp[6].split()[0]) == pytest.approx(f, rel=1e-6)
        assert inp[4].split()[0] == repr(job.tol) and int(inp[5].split()[0]) == job.max_it
        assert inp[7].strip().startswith("R")
        src = os.path.join(fd, "aphid.source")
        if stype == "CS":
            got = [ln.split() for ln in open(src).read().splitlines()]
            assert got == [[meshgen.DIR_NAMES[dr], str(i), str(j), str(k)]
                           for (dr, i, j, k) in dips]
        else:
            assert not os.path.exists(src)
