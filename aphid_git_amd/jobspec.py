"""Multi-frequency / multi-source jobs of the reference's MPI driver, run as
one batched solve on one GPU.

The reference (src/aphid_parallel.f90:40-193, man/aphid.1:168-207) reads
  aphid_mpi.freq    3 comment lines, NFREQ, then NFREQ frequencies [Hz];
  aphid_mpi.input   3 comment lines, the job type ('MT' plane waves or 'CS'
                    dipole arrays), then the serial aphid.input lines without
                    the mode and frequency lines (QMR target, max iterations,
                    conductivity model, ...);
  aphid_mpi.source  (CS only) 3 comment lines, NSRC, 1 comment line, then per
                    dipole array 3 comment lines, N, N dipoles '{+/-}{X/Y/Z}
                    i j k', 1 separator line;
  aphid.mesh        as for the serial program;
and starts one MPI process per (source, frequency): process p (0-based)
takes frequency p mod NFREQ and source p div NFREQ (for MT: x- then
y-polarisation, serial modes 2 and 3; for CS: array p div NFREQ, mode 5).
Each process runs in directory tmp-<p+1> (zero-padded to the digits of the
process count) from a generated aphid.input.

Here every (source, frequency) is one system of a BatchSolver: one operator
per frequency (the mesh and model are shared), and one solve for all.
run_job writes, per system, the generated aphid.input, qmr.iter
(iteration, residual norm, seconds, one line per iteration as each of the
reference's processes writes it, src/aphid_qmr.f90:381; at most
HISTORY_ENTRIES // n_systems lines) and aphid_solution.bin (A_x, A_y,
A_z, Phi in the reference's shapes, raw complex128 in Fortran order, as the
serial driver aphid_gpu_serial writes).
"""
from __future__ import annotations

import dataclasses
import os
import time
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import gpu, meshgen

Dipole = Tuple[int, int, int, int]  # (signed direction code, i, j, k), 1-based


def _num(t: str) -> float:
    return float(t.replace("d", "e").replace("D", "e"))


@dataclasses.dataclass
class Job:
    source_type: str              # 'MT' or 'CS'
    tol: float
    max_it: int
    model: str
    input_tail: List[str]         # aphid_mpi.input lines after the max_it line
    input_head: List[str]         # its first 3 lines
    freqs: List[float]
    arrays: List[List[Dipole]]    # CS dipole arrays (empty for MT)
    xn: np.ndarray
    yn: np.ndarray
    zn: np.ndarray
    cells: dict

    @property
    def n_sources(self) -> int:
        return 2 if self.source_type == "MT" else len(self.arrays)

    def systems(self) -> List[Tuple[int, float, List[Dipole]]]:
        """(serial mode, frequency, dipoles) of process p = s*NFREQ + f."""
        out = []
        for s in range(self.n_sources):
            for f in self.freqs:
                if self.source_type == "MT":
                    out.append((2 + s, f, []))
                else:
                    out.append((5, f, self.arrays[s]))
        return out


def read_job(d: str) -> Job:
    with open(os.path.join(d, "aphid_mpi.freq")) as f:
        fl = f.read().splitlines()
    nf = int(fl[3].split()[0])
    freqs = [_num(fl[4 + q].split()[0]) for q in range(nf)]
    with open(os.path.join(d, "aphid_mpi.input")) as f:
        il = f.read().splitlines()
    stype = il[3].split()[0][:2]
    if stype not in ("MT", "CS"):
        raise ValueError(f"aphid_mpi.input: job type must be MT or CS, got {il[3]!r}")
    tol = _num(il[4].split()[0])
    max_it = int(il[5].split()[0])
    model = il[6].strip()[0]
    arrays: List[List[Dipole]] = []
    if stype == "CS":
        with open(os.path.join(d, "aphid_mpi.source")) as f:
            sl = f.read().splitlines()
        ns = int(sl[3].split()[0])
        pos = 5
        for _ in range(ns):
            pos += 3
            nd = int(sl[pos].split()[0])
            pos += 1
            arr = []
            for q in range(nd):
                t = sl[pos + q].split()
                arr.append((meshgen.DIR_CODES[t[0]], int(t[1]), int(t[2]), int(t[3])))
            arrays.append(arr)
            pos += nd + 1
    xn, yn, zn, cells = meshgen.read_mesh(os.path.join(d, "aphid.mesh"), model)
    return Job(stype, tol, max_it, model, il[6:], il[:3], freqs, arrays, xn, yn, zn, cells)


def write_job(job: Job, d: str) -> None:
    """The four input files of a job (the inverse of read_job)."""
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "aphid_mpi.freq"), "w") as f:
        f.write("=" * 72 + "\nAPHID: source frequency array [Hz]\n" + "=" * 72 + "\n")
        f.write(f"{len(job.freqs)}            ! number of frequencies\n")
        for v in job.freqs:
            f.write(f"{v!r}\n")
    with open(os.path.join(d, "aphid_mpi.input"), "w") as f:
        f.write("\n".join(job.input_head) + "\n")
        f.write(f"{job.source_type}        !! mode of operation\n{job.tol!r}  !! QMR target\n"
                f"{job.max_it}  !! max QMR iterates\n")
        f.write("\n".join(job.input_tail) + "\n")
    if job.source_type == "CS":
        with open(os.path.join(d, "aphid_mpi.source"), "w") as f:
            f.write("=" * 72 + "\nAPHID: dipole arrays\n" + "=" * 72 + "\n")
            f.write(f"{len(job.arrays)}              !  number of dipole arrays\n\n")
            for arr in job.arrays:
                f.write("=" * 72 + "\ndipole array definition\n" + "=" * 72 + "\n")
                f.write(f"{len(arr)}              ! number of elementary dipoles\n")
                for (dr, i, j, k) in arr:
                    f.write(f"{meshgen.DIR_NAMES[dr]} {i} {j} {k}\n")
                f.write("\n")
    meshgen.write_mesh(os.path.join(d, "aphid.mesh"), job.xn, job.yn, job.zn, job.model, job.cells)


def tmp_dir_name(p: int, nprocs: int) -> str:
    """The reference's per-process directory: tmp-<p+1>, zero-padded to
    the digits of the process count (src/aphid_parallel.f90:143-161)."""
    if nprocs > 999:
        raise ValueError("the reference's directory scheme stops at 999 processes")
    digits = 1 if nprocs <= 9 else 2 if nprocs <= 99 else 3
    return f"tmp-{p + 1:0{digits}d}"


def sweep_guesses(freqs: Sequence[float], n_sources: int, levels: int = 3) -> Dict[int, List[int]]:
    """The frequency-sweep schedule of a job: per source, the frequencies in
    ascending order, q = 0, 1, ...; every 2^(levels-1)-th one starts from
    x0 = 0, and each other one (system p = s*NFREQ + f) from the solutions
    d places below and above it, d the lowest set bit of q (the batch
    solves it in the wave after theirs).  levels = 2: every second
    frequency from its two neighbours.  Returns {fitted system: source
    systems}."""
    if levels < 2:
        raise ValueError("a sweep has at least 2 levels")
    nf = len(freqs)
    order = sorted(range(nf), key=lambda f: freqs[f])
    out = {}
    for s in range(n_sources):
        for q in range(1, nf):
            if q % (1 << (levels - 1)) == 0:
                continue
            d = q & -q
            nb = [order[q - d]] + ([order[q + d]] if q + d < nf else [])
            out[s * nf + order[q]] = [s * nf + f for f in nb]
    return out


# residual history kept over all systems of a job (256 MB of FP64)
HISTORY_ENTRIES = 1 << 25


def run_job(d: str, device: int = 0, out_dir: Optional[str] = None, write: bool = True,
            flags: int = 0, sweep: bool = False):
    """Solve every (source, frequency) of the job in d as one batched solve
    (sweep: the frequencies of each source, in ascending order, are solved
    in the waves of sweep_guesses, all but every fourth from a fit of
    solutions at neighbouring frequencies).  Returns (per-system
    reports, device seconds, wall seconds of the solve)."""
    job = read_job(d)
    out_dir = out_dir or d
    systems = job.systems()
    rhs = gpu.Solver(device)
    rhs.set_mesh(job.xn, job.yn, job.zn)
    B = gpu.BatchSolver(device)
    ops = [B.add_operator(job.xn, job.yn, job.zn, job.model, 2 * np.pi * f, job.cells)
           for f in job.freqs]
    nf = len(job.freqs)
    for p, (mode, f, dips) in enumerate(systems):
        b, skipped = rhs.build_rhs(mode, dips)
        B.add_system(ops[p % nf], b)
    rhs.close()
    if sweep:
        for p, src in sweep_guesses(job.freqs, job.n_sources).items():
            B.set_sweep_guess(p, src)
    hcap = max(1, min(job.max_it, HISTORY_ENTRIES // len(systems))) if write else 0
    B.set_history(hcap)
    t0 = time.perf_counter()
    reports, secs = B.solve(job.tol, job.max_it, flags=flags)
    wall = time.perf_counter() - t0
    if write:
        for p, (mode, f, dips) in enumerate(systems):
            td = os.path.join(out_dir, tmp_dir_name(p, len(systems)))
            os.makedirs(td, exist_ok=True)
            with open(os.path.join(td, "aphid.input"), "w") as fo:
                fo.write(job.input_head[0] + "\n")
                fo.write("APhiD input file generated from ../aphid_mpi.input\n")
                fo.write(job.input_head[2] + "\n")
                fo.write(f"{mode:4d}   !! operational mode\n{job.tol!r}  !! QMR target residual\n"
                         f"{job.max_it}  !! max QMR iterates\n{f:14.7e}  !! frequency [Hz]\n")
                fo.write("\n".join(job.input_tail) + "\n")
            if dips:
                with open(os.path.join(td, "aphid.source"), "w") as fo:
                    for (dr, i, j, k) in dips:
                        fo.write(f"{meshgen.DIR_NAMES[dr]} {i} {j} {k}\n")
            hist, htime = B.history(p, hcap)
            with open(os.path.join(td, "qmr.iter"), "w") as fo:
                fo.writelines(f"{i + 1:8d} {h:24.16e} {t:24.16e}\n"
                              for i, (h, t) in enumerate(zip(hist, htime)))
            x = B.solution(p)
            with open(os.path.join(td, "aphid_solution.bin"), "wb") as fo:
                for a in x:
                    np.asfortranarray(a, np.complex128).ravel(order="F").tofile(fo)
    B.close()
    return reports, secs, wall
