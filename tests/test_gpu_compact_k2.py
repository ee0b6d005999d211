"""The stencil's compact k^2 (aphid_set_compact_k2): for models whose
sigma-hat has one imaginary part in every cell (model R), the real parts of
k2x, k2y and k2z come from 2-D tables, the real part of k2p from a one-byte
offset of its bit pattern, and a 32-byte record of the four imaginary parts
per node.  The operator's values must not change: the compact and the whole
k^2 give the same applies and the same solves, bit for bit, and the library
keeps the whole grids wherever the tables would not reproduce them."""
import numpy as np
import pytest

from aphid_git_amd import meshgen
from conftest import flat4, relerr, shapes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    from aphid_git_amd import gpu as G

    s = G.Solver()
    yield s
    s.close()


def _rand4(shape, seed):
    rng = np.random.default_rng(seed)
    return [np.asfortranarray(rng.standard_normal(s) + 1j * rng.standard_normal(s))
            for s in shapes(*shape)]


def _bitwise(a, b):
    a = np.ascontiguousarray(np.asarray(a).ravel(order="F"))
    b = np.ascontiguousarray(np.asarray(b).ravel(order="F"))
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_compact_k2_applies_are_bitwise_the_whole_k2_applies_on_config2(gpu):
    """Config 2 (129x129x97 nodes, model R): the tuned tile and two others,
    each with the compact and the whole k^2, on all 6,456,708 entries."""
    case = meshgen.config2()
    print(f"[compact] config2 {case.shape}", flush=True)
    gpu.setup_case(case)
    assert gpu.set_compact_k2(True), "model R must take the compact k^2"
    kp = np.ascontiguousarray(np.asarray(gpu.get_k2()[3]).real).view(np.int64)
    spread = int(kp.max() - kp.min())
    print(f"[compact] config2 k2p real parts span {spread} units in the last place", flush=True)
    assert spread <= 255
    u = _rand4(case.shape, 3)
    tiles = [None, (32, 8, 12), (16, 16, 8)]
    for t in tiles:
        if t:
            gpu.set_tile(*t)
        tile = gpu.get_tile(0)
        assert tile[2] > 0, "the march variant is the one with a compact form"
        assert gpu.set_compact_k2(True)
        a = flat4(gpu.apply(u))
        assert not gpu.set_compact_k2(False)
        b = flat4(gpu.apply(u))
        d = np.max(np.abs(a - b)) / np.max(np.abs(b))
        print(f"[compact] tile {tile}: max rel diff {d:.1e}", flush=True)
        assert _bitwise(a, b), f"tile {tile}: compact and whole k^2 differ ({d:.1e})"
    gpu.set_compact_k2(True)


def test_compact_k2_solve_is_bitwise_the_whole_k2_solve(gpu):
    """Config 1 to rel-res 1e-8 with and without the compact k^2: the same
    iteration count, residual history and solution bits."""
    case = meshgen.config1()
    b = gpu.setup_case(case)
    assert gpu.set_compact_k2(True)
    r1 = gpu.qmr_solve(b, 1e-8, 5000)
    assert not gpu.set_compact_k2(False)
    r0 = gpu.qmr_solve(b, 1e-8, 5000)
    gpu.set_compact_k2(True)
    print(f"[compact] config1: {r1.iters} / {r0.iters} iterations", flush=True)
    assert r1.status == r0.status == 0
    assert r1.iters == r0.iters
    assert np.array_equal(r1.hist, r0.hist)
    for c in range(4):
        assert _bitwise(np.asarray(r1.x[c]), np.asarray(r0.x[c]))


@pytest.mark.parametrize("model,active", [("R", True), ("C", False), ("D", False), ("I", False)])
def test_compact_k2_only_where_the_tables_hold(gpu, oracle, model, active):
    """Models C, D and I have per-cell imaginary parts: the check keeps the
    whole grids, and the apply still matches the oracle."""
    case = meshgen.tiny_case(9, 8, 7, seed=11, model=model, freq=40.0)
    gpu.set_mesh(case.xn, case.yn, case.zn)
    gpu.set_cells(case.model, case.omega, case.cells)
    assert gpu.set_compact_k2(True) == active
    S = oracle.System(case)
    u = _rand4(case.shape, 5)
    got, want = gpu.apply(u), S.op.apply(u)
    for c in range(4):
        assert relerr(got[c], want[c]) < 1e-13


def test_compact_k2_from_host_k2(gpu, oracle):
    """k^2 given through aphid_set_k2: the oracle's model-R grids keep the
    structure (compact), the same grids with one edge real part changed do
    not, nor do grids whose k2p real parts are spread wider than 256 units
    in the last place."""
    case = meshgen.random_block_case((14, 12, 10), seed=9, freq=7.0)
    gpu.set_mesh(case.xn, case.yn, case.zn)
    S = oracle.System(case)
    k2 = [np.array(a, copy=True) for a in S.k2]
    gpu.set_k2(k2)
    assert gpu.set_compact_k2(True)
    u = _rand4(case.shape, 2)
    a = flat4(gpu.apply(u))
    k2[1][3, 4, 5] += 1e-3 * abs(k2[1][3, 4, 5])
    gpu.set_k2(k2)
    assert not gpu.set_compact_k2(True)
    b = flat4(gpu.apply(u))
    assert relerr(b, a) > 0  # the changed entry is used
    k2[1][3, 4, 5] = S.k2[1][3, 4, 5]
    k2[3][2, 2, 2] *= 1.0 + 1e-12
    gpu.set_k2(k2)
    assert not gpu.set_compact_k2(True)
    c = flat4(gpu.apply(u))
    assert relerr(c, a) > 0


def test_compact_k2_follows_the_frequency(gpu):
    """k2p's real part depends on omega: a context rebuilt at another
    frequency must use the new value (the applies stay bitwise those of the
    whole k^2)."""
    for freq in (3.0, 250.0, 3.0e4):
        case = meshgen.random_block_case((12, 10, 9), seed=4, freq=freq)
        gpu.set_mesh(case.xn, case.yn, case.zn)
        gpu.set_cells(case.model, case.omega, case.cells)
        assert gpu.set_compact_k2(True), freq
        u = _rand4(case.shape, 6)
        a = flat4(gpu.apply(u))
        assert not gpu.set_compact_k2(False)
        b = flat4(gpu.apply(u))
        gpu.set_compact_k2(True)
        assert _bitwise(a, b), freq
