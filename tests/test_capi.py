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


def test_new_entry_points_reject_bad_arguments_without_a_device():
    """Argument checks of the network, warm-start and post-processing entry
    points return error codes before touching a device."""
    import ctypes as C

    from aphid_git_amd import gpu

    if not os.path.exists(gpu.LIB_PATH):
        pytest.skip("libaphid_gpu.so not built")
    lib = gpu.lib()
    with pytest.raises(gpu.AphidError, match="multiples of 4"):
        gpu.UNet((6, 8, 8))
    assert lib.aphid_unet_param_count(None) == -1
    assert lib.aphid_unet_forward(None, None, None) == -1
    assert lib.aphid_unet_train(None, 1, 1, C.c_float(1e-3), C.c_ulonglong(0), None, None) == -1
    assert lib.aphid_warm_start(None, None, None, C.c_uint(0), None) == -1
    assert lib.aphid_gen_fields(None, C.c_double(1.0), None, None, None, None, None, None) == -1
    assert lib.aphid_get_sigma(None, None) == -1
    assert b"NULL" in lib.aphid_last_error() or b"argument" in lib.aphid_last_error()
