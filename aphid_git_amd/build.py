"""Build recipe of the native parts, in-tree.

    python -m aphid_git_amd.build

compiles libaphid_gpu.so (hipcc, gfx950) from aphid_git_amd/csrc and the
Fortran driver (amdflang) from aphid_git_amd/fortran.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
FSRC = os.path.join(HERE, "fortran")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FC = os.environ.get("APHID_FC", "/opt/rocm/bin/amdflang")
ARCH = os.environ.get("APHID_ARCH", "gfx950")

HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
             # packed-FP32 instructions are kept off for all device code
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


SOURCES = ("aphid_gpu.hip", "unet.hip")


def build_hip(verbose_resources: bool = False) -> str:
    """Each translation unit to an object (in parallel), then one shared library."""
    out = os.path.join(HERE, "libaphid_gpu.so")
    tmp = out + ".tmp"
    bdir = os.path.join(HERE, "_obj")
    os.makedirs(bdir, exist_ok=True)
    procs, objs = [], []
    for src in SOURCES:
        obj = os.path.join(bdir, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC, *HIP_FLAGS, "-c", "-o", obj, os.path.join(CSRC, src)]
        if verbose_resources:
            cmd.insert(1, "-Rpass-analysis=kernel-resource-usage")
        print("+", " ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        objs.append(obj)
    if any(p.wait() != 0 for p in procs):
        raise RuntimeError("hipcc failed")
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, out)
    return out


FPROGRAMS = (("aphid_gpu_serial", ("aphid_vtk.f90", "aphid_driver.f90")),
             ("aphid_gpu_parallel", ("aphid_parallel_driver.f90",)))


def build_fortran() -> str:
    """Fortran drivers: the iso_c_binding module, then the aphid_gpu_serial
    (one case) and aphid_gpu_parallel (multi-frequency / multi-source job)
    programs.  Returns the serial program's path."""
    if not os.path.exists(FSRC) or not os.path.exists(FC):
        return ""
    bdir = os.path.join(HERE, "fortran", "_mod")
    os.makedirs(bdir, exist_ok=True)
    mod = os.path.join(FSRC, "aphid_gpu_mod.f90")
    if not os.path.exists(mod):
        return ""
    mobj = os.path.join(bdir, "aphid_gpu_mod.o")
    _run([FC, "-O2", "-J" + bdir, "-c", mod, "-o", mobj])
    link = ["-L" + HERE, "-laphid_gpu", "-Wl,-rpath,$ORIGIN", "-L/opt/rocm/lib",
            "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"]
    for name, files in FPROGRAMS:
        srcs = [os.path.join(FSRC, f) for f in files]
        if not all(os.path.exists(s) for s in srcs):
            continue
        _run([FC, "-O2", "-J" + bdir, "-I" + bdir, mobj, *srcs, "-o",
              os.path.join(HERE, name), *link])
    return os.path.join(HERE, "aphid_gpu_serial")


def main():
    build_hip(verbose_resources="-v" in sys.argv)
    build_fortran()


if __name__ == "__main__":
    main()
