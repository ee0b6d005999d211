"""The multi-frequency / multi-source job files of the reference's MPI
driver (src/aphid_parallel.f90:40-193), parsed on the CPU.  The sample
files under tests/golden/mpi_sample are the reference's own
parallel/aphid_mpi.{freq,input,source}."""
import os
import shutil

import numpy as np
import pytest

from aphid_git_amd import jobspec, meshgen
from conftest import GOLDEN

SAMPLE = os.path.join(GOLDEN, "mpi_sample")


def _with_mesh(tmp_path, input_type=None):
    d = tmp_path / "job"
    shutil.copytree(SAMPLE, d)
    case = meshgen.random_block_case((8, 8, 8), 1, 1.0)
    meshgen.write_mesh(str(d / "aphid.mesh"), case.xn, case.yn, case.zn, "R", case.cells)
    if input_type:
        lines = (d / "aphid_mpi.input").read_text().splitlines()
        lines[3] = input_type + "        !! mode of operation"
        (d / "aphid_mpi.input").write_text("\n".join(lines) + "\n")
    return str(d), case


def test_reference_sample_mt_job(tmp_path):
    d, case = _with_mesh(tmp_path)
    job = jobspec.read_job(d)
    assert job.source_type == "MT"
    assert job.tol == 1e-18 and job.max_it == 2000 and job.model == "R"
    # the count line says 2: the third listed frequency is not read
    assert job.freqs == [0.01, 0.1]
    sys = job.systems()
    # process p: frequency p mod NFREQ; x-polarisation (mode 2) first
    assert [(m, f) for m, f, _ in sys] == [(2, 0.01), (2, 0.1), (3, 0.01), (3, 0.1)]
    np.testing.assert_array_equal(job.xn, case.xn)
    np.testing.assert_array_equal(job.cells["sigma"], case.cells["sigma"])


def test_reference_sample_cs_job(tmp_path):
    d, _ = _with_mesh(tmp_path, "CS")
    job = jobspec.read_job(d)
    assert job.source_type == "CS"
    assert job.arrays == [[(1, 50, 50, 51)], [(1, 50, 50, 51), (1, 52, 50, 51)],
                          [(2, 50, 50, 51), (2, 50, 52, 51)]]
    sys = job.systems()
    assert len(sys) == 6
    assert all(m == 5 for m, _, _ in sys)
    assert sys[3] == (5, 0.1, job.arrays[1])


def test_write_read_round_trip(tmp_path):
    d, _ = _with_mesh(tmp_path, "CS")
    job = jobspec.read_job(d)
    job.freqs = [0.5, 2.0, 8.0]
    d2 = str(tmp_path / "out")
    jobspec.write_job(job, d2)
    back = jobspec.read_job(d2)
    assert back.freqs == job.freqs and back.arrays == job.arrays
    assert back.tol == job.tol and back.max_it == job.max_it and back.model == job.model
    np.testing.assert_array_equal(back.cells["sigma"], job.cells["sigma"])


@pytest.mark.parametrize("p,n,name", [(0, 4, "tmp-1"), (9, 12, "tmp-10"), (0, 12, "tmp-01"),
                                      (191, 192, "tmp-192"), (4, 192, "tmp-005")])
def test_process_directory_names(p, n, name):
    assert jobspec.tmp_dir_name(p, n) == name


def test_sweep_schedule_pairs_neighbouring_frequencies():
    """Per source, the frequencies in ascending order (process p =
    s*NFREQ + f, src/aphid_parallel.f90 process order), also when the job
    lists its frequencies unsorted.  Two levels: every second one starts
    from its neighbours.  Three (the default): every fourth from zero, the
    ones halfway between from those two places away, the rest from their
    neighbours."""
    from aphid_git_amd import jobspec

    freqs = [30.0, 0.3, 3000.0, 3.0, 300.0]  # ascending: 1 3 0 4 2
    g = jobspec.sweep_guesses(freqs, 2, levels=2)
    assert g == {3: [1, 0], 4: [0, 2], 8: [6, 5], 9: [5, 7]}
    g = jobspec.sweep_guesses(freqs, 2)
    assert g == {3: [1, 0], 0: [1, 2], 4: [0, 2], 8: [6, 5], 5: [6, 7], 9: [5, 7]}
    assert jobspec.sweep_guesses([5.0], 3) == {}
    assert jobspec.sweep_guesses([2.0, 1.0], 1) == {0: [1]}
    with pytest.raises(ValueError):
        jobspec.sweep_guesses(freqs, 1, levels=1)
