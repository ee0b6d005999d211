"""aphid_apply_noalloc: the operator apply with its device staging in the
library's per-GPU scratch array instead of the context's vectors.  Its
output equals aphid_apply's on the same context (floating point, 1e-13
relative to each component's largest entry, the apply's own bar against
the oracle), on two calls in a row with the second mesh smaller, and it
leaves the device's free memory unchanged (checked when the process maps
a single HIP runtime)."""
import ctypes as C
import os

import numpy as np
import pytest

from aphid_git_amd import meshgen
from conftest import relerr

pytestmark = pytest.mark.gpu


def _rand4(shape, seed):
    from aphid_git_amd import gpu as G

    rng = np.random.default_rng(seed)
    return [np.asfortranarray(rng.standard_normal(s) + 1j * rng.standard_normal(s))
            for s in G.comp_shapes(*shape)]


def _hip_runtime():
    """The HIP runtime the library runs on, when the process maps exactly
    one (torch bundles its own copy: once both are mapped, a bare dlopen of
    libamdhip64 may return the one the library does not use)."""
    paths = set()
    with open("/proc/self/maps") as f:
        for ln in f:
            t = ln.split()
            if len(t) >= 6 and os.path.basename(t[5]).startswith("libamdhip64.so"):
                paths.add(os.path.realpath(t[5]))
    return C.CDLL(paths.pop()) if len(paths) == 1 else None


def _free_bytes(hip):
    if hip is None:
        return None
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def test_noalloc_apply_matches_apply_twice_second_smaller():
    from aphid_git_amd import gpu as G

    cases = [meshgen.random_block_case((30, 26, 22), seed=31, freq=40.0),  # 31x27x23 nodes
             meshgen.tiny_case(12, 9, 11, seed=32, model="C", freq=500.0)]
    for n, case in enumerate(cases):
        with G.Solver() as s:
            s.setup_case(case)
            u = _rand4(case.shape, 40 + n)
            want = s.apply(u)
            hip = _hip_runtime()
            f0 = _free_bytes(hip)
            got = s.apply_noalloc(u)
            f1 = _free_bytes(hip)
            errs = [relerr(g, w) for g, w in zip(got, want)]
            print(f"[noalloc] call {n + 1}: {case.shape} nodes, max rel diff {max(errs):.2e}, "
                  f"free bytes {f0} -> {f1}", flush=True)
            assert max(errs) < 1e-13
            # device memory unchanged (checked where one HIP runtime is mapped)
            assert f1 == f0


def test_noalloc_apply_on_an_untuned_context_and_size_limit():
    """First call of a fresh context (no tuned tile yet) uses the default
    tile and still matches aphid_apply; a mesh past the scratch size is an
    APHID_E_ARG before anything runs."""
    from aphid_git_amd import gpu as G

    case = meshgen.random_block_case((16, 14, 12), seed=33, freq=3.0)
    with G.Solver() as s:
        s.setup_case(case)
        u = _rand4(case.shape, 50)
        got = s.apply_noalloc(u)
        want = s.apply(u)
        for c in range(4):
            assert relerr(got[c], want[c]) < 1e-13
    n = 78  # 9 * 78^3 * 16 bytes > 64 MiB
    assert 9 * n ** 3 * 16 > G.NOALLOC_SCRATCH_BYTES >= 9 * 77 ** 3 * 16
    big = meshgen.random_block_case((n - 1, n - 1, n - 1), seed=34, freq=3.0)
    with G.Solver() as s:
        s.setup_case(big)
        with pytest.raises(G.AphidError, match="too large") as e:
            s.apply_noalloc(_rand4(big.shape, 51))
        assert "code -1" in str(e.value)
