"""Write the golden fixtures in tests/golden/*.npz from the REFERENCE.

Run in the build container only (it needs /root/reference):

    python tests/golden/make_golden.py

It builds the reference's own Fortran sources with oracle/Makefile (target
`ref`, outputs in oracle/_ref/), writes each synthetic case in the reference
file formats, runs oracle/_ref/ref_harness in that case directory and stores
the harness's full-precision dumps together with the case inputs.  The
fixtures are data only: inputs and the reference's outputs.
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from aphid_git_amd import meshgen  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")


def comp_shapes(nx, ny, nz):
    return [(nx - 1, ny, nz), (nx, ny - 1, nz), (nx, ny, nz - 1), (nx, ny, nz)]


def run(d, *args, threads=1):
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    r = subprocess.run([HARNESS, *args], cwd=d, capture_output=True, text=True, env=env)
    if r.returncode != 0:
        raise RuntimeError(f"harness {args} failed: {r.stderr[-400:]}")
    return r.stdout


def read_setup(d, shape):
    nx, ny, nz = shape
    raw = open(os.path.join(d, "ref_setup.bin"), "rb").read()
    off = 0

    def take(n, dt):
        nonlocal off
        a = np.frombuffer(raw, dtype=dt, count=n, offset=off)
        off += a.nbytes
        return a.copy()

    dims = take(3, np.int32)
    assert tuple(dims) == (nx, ny, nz)
    out = {}
    out["dx"], out["dy"], out["dz"] = take(nx + 1, "f8"), take(ny + 1, "f8"), take(nz + 1, "f8")
    out["dxc"], out["dyc"], out["dzc"] = take(nx, "f8"), take(ny, "f8"), take(nz, "f8")
    out["sig"] = take((nx - 1) * (ny - 1) * (nz - 1), "c16")
    for name in ("k2", "diag", "rhs"):
        for c, s in zip("xyzp", comp_shapes(nx, ny, nz)):
            out[f"{name}_{c}"] = take(int(np.prod(s)), "c16")
    assert off == len(raw)
    return out


def case_arrays(case):
    a = {"xn": case.xn, "yn": case.yn, "zn": case.zn,
         "model": np.array(case.model), "freq": np.array(case.freq),
         "mode": np.array(case.mode), "target": np.array(case.target),
         "max_it": np.array(case.max_it),
         "sources": np.array(case.sources, dtype=np.int32).reshape(-1, 4)}
    for k, v in case.cells.items():
        a["cell_" + k] = np.asarray(v, np.float32).ravel(order="F")
    return a


def golden_case(name, case, napply=3, qmr=False, sample=None, setup=True):
    shape = case.shape
    with tempfile.TemporaryDirectory() as d:
        meshgen.write_case(case, d)
        out = case_arrays(case)
        txt = run(d, "setup")
        st = read_setup(d, shape)
        # dipoles the reference skipped: it prints "<n> error: ..." for each
        skipped = np.zeros(len(case.sources), np.int32)
        for ln in txt.splitlines():
            t = ln.split()
            if len(t) >= 2 and t[1] == "error:":
                skipped[int(t[0]) - 1] = 1
        out["skipped"] = skipped
        if setup:
            out.update(st)
        else:  # large mesh: keep only the non-zero right-hand-side entries
            for c in "xyzp":
                v = st[f"rhs_{c}"]
                nzi = np.nonzero(v)[0]
                out[f"rhsnz_{c}_idx"] = nzi.astype(np.int64)
                out[f"rhsnz_{c}_val"] = v[nzi]
        if napply:
            rng = np.random.default_rng(1234)
            ins = []
            with open(os.path.join(d, "ref_apply_in.bin"), "wb") as f:
                for _ in range(napply):
                    for s in comp_shapes(*shape):
                        n = int(np.prod(s))
                        v = rng.standard_normal(n) + 1j * rng.standard_normal(n)
                        ins.append(v)
                        f.write(v.astype(np.complex128).tobytes())
            run(d, "apply", str(napply))
            outs = np.frombuffer(open(os.path.join(d, "ref_apply_out.bin"), "rb").read(), "c16")
            out["apply_in"] = np.concatenate(ins)
            out["apply_out"] = outs.copy()
        if qmr:
            txt = run(d, "qmr1")
            it = np.loadtxt(os.path.join(d, "qmr.iter"), ndmin=2)
            sol = np.frombuffer(open(os.path.join(d, "ref_qmr_sol.bin"), "rb").read(), "c16")
            conv = [ln for ln in txt.splitlines() if "converged in" in ln]
            out["qmr_iters"] = np.array(int(conv[0].split()[-2]) if conv else -1)
            out["qmr_hist"] = it[:, 1].copy()
            lens = [int(np.prod(s)) for s in comp_shapes(*shape)]
            offs = np.cumsum([0] + lens)
            out["qmr_comp_norms"] = np.array([np.linalg.norm(sol[offs[c]:offs[c + 1]]) for c in range(4)])
            if sample is None:
                out["qmr_sol"] = sol.copy()
            else:
                rng = np.random.default_rng(77)
                idx = np.sort(rng.choice(len(sol), size=int(sample * len(sol)), replace=False))
                out["qmr_sol_idx"] = idx.astype(np.int64)
                out["qmr_sol_val"] = sol[idx].copy()
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path) / 1024:.0f} KiB)")


VTK_FILES = ("sigma.vtk", "ref_a_phi.vtk", "Bx.vtk", "By.vtk", "Bz.vtk", "Ex.vtk", "Ey.vtk",
             "Ez.vtk", "EM_grid.vtk")


def golden_fields(name, case):
    """The reference's serial QMR, gen_fields and VTK writers on one case:
    its solution, the six field arrays (raw, in order Bx By Bz Ex Ey Ez) and
    the text of every VTK file it writes (dump_a_phi's unit 11 captured as
    ref_a_phi.vtk)."""
    with tempfile.TemporaryDirectory() as d:
        meshgen.write_case(case, d)
        out = case_arrays(case)
        run(d, "fields")
        out["qmr_sol"] = np.frombuffer(open(os.path.join(d, "ref_qmr_sol.bin"), "rb").read(),
                                       "c16").copy()
        out["fields"] = np.frombuffer(open(os.path.join(d, "ref_fields.bin"), "rb").read(),
                                      "c16").copy()
        for fn in VTK_FILES:
            out["vtk_" + fn.replace(".", "_")] = np.array(open(os.path.join(d, fn)).read())
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path) / 1024:.0f} KiB)")


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    T = meshgen.tiny_case
    # operator/setup goldens for all four conductivity models + QMR on two
    golden_case("tiny_R_7x6x5", T(7, 6, 5, seed=1, model="R", freq=3.0), qmr=True)
    golden_case("tiny_C_9x8x7", T(9, 8, 7, seed=2, model="C", freq=2.0e3), qmr=True)
    golden_case("tiny_D_8x7x9", T(8, 7, 9, seed=3, model="D", freq=50.0))
    golden_case("tiny_I_6x7x8", T(6, 7, 8, seed=4, model="I", freq=1.0e6))
    # right-hand-side modes 1-3, and mode 5 with skipped (out-of-bounds /
    # boundary) dipoles
    golden_case("rhs_mode1", T(7, 7, 6, seed=5, mode=1), napply=0)
    golden_case("rhs_mode2", T(6, 7, 5, seed=6, mode=2), napply=0)
    golden_case("rhs_mode3", T(7, 6, 5, seed=7, mode=3), napply=0)
    golden_case("rhs_mode5_edges",
                T(6, 6, 6, seed=8, sources=[(1, 6, 1, 1), (1, 5, 6, 6), (-1, 1, 2, 2),
                                            (-1, 6, 6, 6), (2, 1, 1, 1), (2, 1, 6, 1),
                                            (-2, 2, 1, 3), (-2, 2, 6, 3), (3, 1, 1, 6),
                                            (3, 2, 2, 5), (-3, 3, 3, 1), (-3, 3, 3, 6),
                                            (1, 0, 2, 2), (2, 2, 2, 9)]),
                napply=0)
    # mode 4 (fixed square loop at indices 51+-20) needs >= 72x72x51 nodes
    big = T(73, 73, 52, seed=9, mode=4, stretch=1.05)
    big.cells = {"sigma": np.full((72, 72, 51), 0.01, np.float32)}  # rhs ignores sigma
    golden_case("rhs_mode4", big, napply=0, setup=False)
    # config-1 family: the 32^3 air/half-space model at 10 Hz, rel-res 1e-8
    golden_case("config1_10Hz", meshgen.config1(), napply=0, qmr=True, sample=0.01, setup=False)
    # post-processing: gen_fields and the VTK writers
    golden_fields("fields_R_7x6x5", T(7, 6, 5, seed=1, model="R", freq=3.0))
    golden_fields("fields_C_8x7x6", T(8, 7, 6, seed=12, model="C", freq=300.0))


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "fields":  # only the post-processing goldens
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
        golden_fields("fields_R_7x6x5", meshgen.tiny_case(7, 6, 5, seed=1, model="R", freq=3.0))
        golden_fields("fields_C_8x7x6", meshgen.tiny_case(8, 7, 6, seed=12, model="C", freq=300.0))
    else:
        main()
