"""Argument guards of the C ABI that protect the kernels' addressing, and
the autotuner's cache file.  Each check returns an error code (or ignores a
bad cache line) before any kernel runs on the data in question."""
import os

import numpy as np
import pytest

from aphid_git_amd import meshgen
from conftest import relerr

pytestmark = pytest.mark.gpu


def _axis(n):
    return np.arange(n, dtype=np.float64) * 10.0


def test_mesh_past_the_32bit_offset_limit_is_rejected():
    """410^3 nodes: 4*nn*16 bytes = 4.4 GB >= 2^32, which the stencils'
    32-bit byte offsets cannot address.  Both the single-system context and
    the batch reject it with APHID_E_ARG before allocating anything; 406^3
    nodes (the largest cube below the limit) passes the check."""
    import ctypes as C

    from aphid_git_amd import gpu as G

    assert 406 ** 3 < G.MAX_NODES <= 407 ** 3
    with G.Solver() as s:
        with pytest.raises(G.AphidError, match="too large") as e:
            s.set_mesh(_axis(410), _axis(410), _axis(410))
        assert "code -1" in str(e.value)
        # the context stays usable
        case = meshgen.tiny_case(7, 6, 5, seed=1)
        s.setup_case(case)
    lib = G.lib()
    with G.BatchSolver() as B:
        x = _axis(410)
        sig = np.zeros(1, np.float32)
        rc = lib.aphid_batch_add_operator(B._b, 410, 410, 410, x.ctypes.data_as(C.c_void_p),
                                          x.ctypes.data_as(C.c_void_p),
                                          x.ctypes.data_as(C.c_void_p), C.c_char(b"R"),
                                          C.c_double(1.0), sig.ctypes.data_as(C.c_void_p),
                                          None, None, None)
        assert rc == -1
        assert b"APHID_MAX_NODES" in lib.aphid_last_error()


def test_bad_tune_cache_lines_are_ignored(tmp_path, oracle):
    """A tile line in APHID_TUNE_CACHE that aphid_set_tile would refuse (a
    tile the kernels do not have, a negative chunk) is skipped: the tuner
    times the candidates instead, and the apply still matches the oracle.
    A valid line is used as written."""
    from aphid_git_amd import gpu as G

    bad = meshgen.random_block_case((10, 12, 16), seed=21, freq=7.0)   # 11x13x17 nodes
    good = meshgen.random_block_case((12, 10, 14), seed=22, freq=7.0)  # 13x11x15 nodes
    path = tmp_path / "tune.txt"
    path.write_text("11 13 17 0 7 7 12\n"
                    "11 13 17 0 64 4 -1\n"
                    "13 11 15 0 32 8 4\n")
    old = os.environ.get("APHID_TUNE_CACHE")
    os.environ["APHID_TUNE_CACHE"] = str(path)
    try:
        for case, want in ((bad, None), (good, (32, 8, 4))):
            with G.Solver() as s:
                s.setup_case(case)
                rng = np.random.default_rng(3)
                u = [rng.standard_normal(a) + 1j * rng.standard_normal(a)
                     for a in G.comp_shapes(*case.shape)]
                got = s.apply(u)
                tile = s.get_tile(0)
                print(f"[tune cache] {case.shape} tile {tile}", flush=True)
                if want is None:
                    assert tuple(tile[:2]) in G.TILES and tile[2] >= 0
                else:
                    assert tuple(tile) == want
                ref = oracle.System(case).op.apply(u)
                for c in range(4):
                    assert relerr(got[c], ref[c]) < 1e-13
    finally:
        if old is None:
            del os.environ["APHID_TUNE_CACHE"]
        else:
            os.environ["APHID_TUNE_CACHE"] = old


def test_traffic_hook_times_and_leaves_the_solve_alone():
    """aphid_time_traffic (the tuned stencil tile's loads and stores without
    arithmetic) returns a positive time per launch and touches only the
    spare vector: the resident solution and a following solve are those of
    a context that never ran it.  Config 1's mesh tunes to the direct
    variant, for which the hook reports APHID_E_STATE; both contexts then
    take the 64x4 march with 12-plane chunks."""
    from aphid_git_amd import gpu as G

    case = meshgen.config1()
    with G.Solver() as s, G.Solver() as ref:
        b = s.setup_case(case)
        ref.setup_case(case)
        s.time_apply(0, 1)  # tunes the tile
        if s.get_tile(0)[2] == 0:
            with pytest.raises(G.AphidError, match="direct variant"):
                s.time_traffic(1)
        s.set_tile(64, 4, 12)
        ref.set_tile(64, 4, 12)
        r1 = s.qmr_solve(b, 1e-8, 5000, hist_cap=0)
        rr = ref.qmr_solve(b, 1e-8, 5000, hist_cap=0)
        s.load_rhs(b)
        ref.load_rhs(b)
        a = s.solve_resident(1e-8, 5000)
        t = s.time_traffic(5)
        ta = s.time_apply(0, 5)
        print(f"[traffic] config1: {t * 1e6:.2f} us per launch, apply {ta * 1e6:.2f} us", flush=True)
        assert np.isfinite(t) and t > 0
        x = np.concatenate([v.ravel() for v in s.fetch_solution()])
        c = ref.solve_resident(1e-8, 5000)
        xr = np.concatenate([v.ravel() for v in ref.fetch_solution()])
        assert a.iters == c.iters and np.array_equal(x, xr)
        assert r1.iters == rr.iters
