"""The C-ABI library loads and exports every symbol include/aphid_gpu.h
declares (no compute calls: this runs without a GPU)."""
import os
import re

import pytest

from conftest import ROOT


def declared():
    txt = open(os.path.join(ROOT, "include", "aphid_gpu.h")).read()
    pat = r"^\s*(?:int|void|double|long long|const char \*|aphid_unet \*)\s*(aphid_\w+)\s*\("
    return sorted(set(re.findall(pat, txt, re.M)))


def test_header_declares_the_binding_surface():
    from aphid_git_amd import gpu

    assert declared() == sorted(gpu.EXPORTS)


def test_library_exports_every_declared_symbol():
    from aphid_git_amd import gpu

    if not os.path.exists(gpu.LIB_PATH):
        pytest.skip("libaphid_gpu.so not built")
    lib = gpu.lib()
    for name in declared():
        assert hasattr(lib, name), name


def test_no_gpu_is_an_error_code_not_a_crash():
    import torch

    from aphid_git_amd import gpu

    if not os.path.exists(gpu.LIB_PATH):
        pytest.skip("libaphid_gpu.so not built")
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(gpu.AphidError):
        gpu.Solver()
