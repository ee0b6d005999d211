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
