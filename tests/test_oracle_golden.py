"""The CPU oracle against golden vectors written by the reference itself
(compiled from src/aphid_mod.f90, src/aphid.f90, src/aphid_qmr.f90 by
oracle/Makefile and driven by oracle/ref_harness.f90; see
tests/golden/make_golden.py)."""
import os

import numpy as np
import pytest

from conftest import golden_files, load_golden, relerr, split4

SETUP = [p for p in golden_files("tiny_")]
RHS = [p for p in golden_files("rhs_")]


@pytest.mark.parametrize("path", SETUP, ids=os.path.basename)
def test_setup_matches_reference(oracle, path):
    case, z = load_golden(path)
    S = oracle.System(case)
    nx, ny, nz = case.shape
    # metrics (src/aphid.f90:341-361): same IEEE operations -> bitwise
    assert np.array_equal(S.mesh.dx, z["dx"]) and np.array_equal(S.mesh.dxc[1:], z["dxc"])
    assert np.array_equal(S.mesh.dy, z["dy"]) and np.array_equal(S.mesh.dyc[1:], z["dyc"])
    assert np.array_equal(S.mesh.dz, z["dz"]) and np.array_equal(S.mesh.dzc[1:], z["dzc"])
    # sigma-hat from FP32 inputs (src/aphid.f90:379-449)
    tol_sig = 1e-14 if case.model == "D" else 0.0
    assert relerr(S.sig.ravel(order="F"), z["sig"]) <= tol_sig
    for c, name in enumerate("xyzp"):
        assert relerr(S.k2[c].ravel(order="F"), z[f"k2_{name}"]) < 1e-14
        assert relerr(S.diag[c].ravel(order="F"), z[f"diag_{name}"]) < 1e-14
        assert relerr(S.b[c].ravel(order="F"), z[f"rhs_{name}"]) == 0.0


@pytest.mark.parametrize("path", SETUP, ids=os.path.basename)
def test_apply_matches_reference(oracle, path):
    case, z = load_golden(path)
    S = oracle.System(case)
    nx, ny, nz = case.shape
    ins, outs = z["apply_in"], z["apply_out"]
    off = 0
    while off < len(ins):
        u, n = split4(ins[off:], nx, ny, nz)
        ref, _ = split4(outs[off:], nx, ny, nz)
        got = S.op.apply(u)
        for c in range(4):
            assert relerr(got[c], ref[c]) < 1e-13, f"component {c}"
        off += n


@pytest.mark.parametrize("path", [p for p in SETUP if "qmr_hist" in np.load(p)],
                         ids=os.path.basename)
def test_qmr_matches_reference(oracle, path):
    case, z = load_golden(path)
    S = oracle.System(case)
    x, rep, hist = S.solve(nthreads=1)
    ref_hist = z["qmr_hist"]
    assert rep.status == 0
    # identical arithmetic until rounding differences grow (Lanczos)
    m = 20
    assert np.max(np.abs(hist[:m] - ref_hist[:m]) / ref_hist[:m]) < 1e-10
    # iteration counts agree to within a few percent
    assert abs(rep.iters - int(z["qmr_iters"])) <= max(3, 0.1 * int(z["qmr_iters"]))
    sol = np.concatenate([a.ravel(order="F") for a in x])
    assert np.linalg.norm(sol - z["qmr_sol"]) / np.linalg.norm(z["qmr_sol"]) < 1e-7


@pytest.mark.parametrize("nthreads", [1, 4])
def test_config1_qmr_matches_reference(oracle, nthreads):
    """32^3 air/half-space at 10 Hz, rel-res 1e-8 (the reference takes 364
    iterations); the 4-thread component split gives the same iterates."""
    case, z = load_golden(golden_files("config1_")[0])
    S = oracle.System(case)
    x, rep, hist = S.solve(nthreads=nthreads)
    ref_hist = z["qmr_hist"]
    assert rep.status == 0
    assert np.max(np.abs(hist[:30] - ref_hist[:30]) / ref_hist[:30]) < 1e-10
    assert abs(rep.iters - int(z["qmr_iters"])) <= 0.05 * int(z["qmr_iters"])
    sol = np.concatenate([a.ravel(order="F") for a in x])
    idx, val = z["qmr_sol_idx"], z["qmr_sol_val"]
    assert np.linalg.norm(sol[idx] - val) / np.linalg.norm(val) < 1e-6
    norms = [np.linalg.norm(a) for a in x]
    # Ay vanishes by symmetry (rounding noise ~1e-21 of the total)
    np.testing.assert_allclose(norms, z["qmr_comp_norms"], rtol=1e-6,
                               atol=1e-12 * max(norms))


@pytest.mark.parametrize("path", RHS, ids=os.path.basename)
def test_rhs_modes_match_reference(oracle, path):
    case, z = load_golden(path)
    mesh = oracle.Mesh(case.xn, case.yn, case.zn)
    b, skipped = oracle.rhs(case, mesh)
    for c, name in enumerate("xyzp"):
        flat = b[c].ravel(order="F")
        if f"rhs_{name}" in z:
            assert np.array_equal(flat, z[f"rhs_{name}"])
        else:
            idx = z[f"rhsnz_{name}_idx"]
            assert np.array_equal(np.nonzero(flat)[0], idx)
            assert np.array_equal(flat[idx], z[f"rhsnz_{name}_val"])
    # the oracle skips exactly the dipoles the reference reported skipping
    assert skipped.tolist() == z["skipped"].tolist()


def test_config2_oracle_solve_is_pinned_to_the_reference():
    """The config-2 fixture (tests/golden/make_config2_golden.py): the
    oracle's 4-thread solve of the 128x128x96-cell benchmark system follows
    the reference's own serial qmr over the first 30 iterations (residual
    history within 1e-10 relative, then the two round differently, as any
    Lanczos process does), converges, and its stored solution meets the
    tolerance by its recomputed true residual."""
    path = os.path.join(os.path.dirname(golden_files()[0]), "config2_1Hz.npz")
    if not os.path.exists(path):
        pytest.skip("tests/golden/config2_1Hz.npz not generated yet")
    z = dict(np.load(path))
    idx, val = z["oracle_hist_idx"], z["oracle_hist_val"]
    first = val[idx < 30]
    assert len(first) == 30
    assert np.max(np.abs(first - z["ref_hist30"]) / z["ref_hist30"]) < 1e-10
    assert int(z["oracle_status"]) == 0
    assert float(z["oracle_true_rel_res"]) < 1.05e-8
    assert val[-1] < 1e-8 * float(z["oracle_rhsnorm"])
    assert np.linalg.norm(z["sol_val"]) > 0 and len(z["sol_idx"]) == len(z["sol_val"])
