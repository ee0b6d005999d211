"""Post-processing on the GPU: gen_fields (src/aphid.f90:2173-2252) and the
VTK files of the Fortran driver against the reference's own output
(tests/golden/fields_*.npz, made by oracle/ref_harness 'fields')."""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, load_golden, split4

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import fields_ref  # noqa: E402

pytestmark = pytest.mark.gpu
NAMES = ["fields_R_7x6x5", "fields_C_8x7x6"]


def _setup(name):
    from aphid_git_amd import gpu

    case, g = load_golden(os.path.join(GOLDEN, name + ".npz"))
    s = gpu.Solver(0)
    b = s.setup_case(case)
    return s, case, g, b


@pytest.mark.parametrize("name", NAMES)
def test_fields_of_the_reference_solution(name):
    s, case, g, b = _setup(name)
    nx, ny, nz = case.shape
    x, _ = split4(g["qmr_sol"], nx, ny, nz)
    s.load_rhs(b, x0=x)
    got = s.gen_fields(case.omega)
    ref = fields_ref.split(g["fields"], fields_ref.field_shapes(nx, ny, nz))
    for a, r in zip(got, ref):
        assert a.shape == r.shape
        assert np.abs(a - r).max() <= 1e-13 * np.abs(r).max()
    s.close()


@pytest.mark.parametrize("name", ["tiny_R_7x6x5", "tiny_C_9x8x7", "tiny_D_8x7x9", "tiny_I_6x7x8"])
def test_sigma_hat_matches_the_reference(name):
    s, case, g, b = _setup(name)
    sig = s.get_sigma().ravel(order="F")
    np.testing.assert_allclose(sig, g["sig"], rtol=1e-15, atol=0)
    s.close()


def _vtk_compare(got_txt, ref_txt, what):
    gl, rl = got_txt.splitlines(), ref_txt.splitlines()
    assert len(gl) == len(rl), (what, len(gl), len(rl))
    block = []
    for a, r in zip(gl, rl):
        if r[:1].isalpha() or r.startswith("#"):
            assert a.rstrip() == r.rstrip(), (what, a, r)
            continue
        va = np.array([float(t) for t in a.split()])
        vr = np.array([float(t) for t in r.split()])
        assert va.shape == vr.shape, what
        block.append((va, vr))
    # values: both solves stop at the case's residual target and agree to
    # ~1e-7 of the field, then 7 printed digits: bound by each block's max
    for va, vr in block:
        scale = np.abs(vr).max() if vr.size else 0.0
        assert np.all(np.abs(va - vr) <= 2e-6 * np.abs(vr) + 1e-6 * scale), what


@pytest.mark.parametrize("name", NAMES)
def test_driver_writes_the_reference_vtk_files(tmp_path, name):
    from aphid_git_amd import meshgen

    case, g = load_golden(os.path.join(GOLDEN, name + ".npz"))
    d = str(tmp_path / "run")
    meshgen.write_case(case, d)
    exe = os.path.join(ROOT, "aphid_git_amd", "aphid_gpu_serial")
    r = subprocess.run([exe, "-w", d], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    sol = np.fromfile(os.path.join(d, "aphid_solution.bin"), np.complex128)
    ref = g["qmr_sol"]
    assert np.linalg.norm(sol - ref) <= 1e-7 * np.linalg.norm(ref)
    pairs = [("sigma.vtk", "sigma.vtk"), ("aphid_a_phi.vtk", "ref_a_phi.vtk")] + \
            [(f, f) for f in ("Bx.vtk", "By.vtk", "Bz.vtk", "Ex.vtk", "Ey.vtk", "Ez.vtk", "EM_grid.vtk")]
    for mine, theirs in pairs:
        got = open(os.path.join(d, mine)).read()
        ref_txt = str(g["vtk_" + theirs.replace(".", "_")])
        if mine == "sigma.vtk":  # no solve involved: the same text
            assert got == ref_txt
        _vtk_compare(got, ref_txt, mine)
