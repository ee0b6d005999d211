import glob
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def golden_files(prefix=""):
    return sorted(glob.glob(os.path.join(GOLDEN, prefix + "*.npz")))


def load_golden(path):
    """Golden fixture -> (Case, dict of arrays)."""
    from aphid_git_amd import meshgen

    z = dict(np.load(path, allow_pickle=False))
    nx, ny, nz = len(z["xn"]), len(z["yn"]), len(z["zn"])
    cells = {k[5:]: z[k].reshape((nx - 1, ny - 1, nz - 1), order="F")
             for k in z if k.startswith("cell_")}
    case = meshgen.Case(z["xn"], z["yn"], z["zn"], str(z["model"]), cells,
                        float(z["freq"]), int(z["mode"]),
                        [tuple(int(v) for v in row) for row in z["sources"]],
                        float(z["target"]), int(z["max_it"]))
    return case, z


def shapes(nx, ny, nz):
    return [(nx - 1, ny, nz), (nx, ny - 1, nz), (nx, ny, nz - 1), (nx, ny, nz)]


def split4(flat, nx, ny, nz):
    out, o = [], 0
    for s in shapes(nx, ny, nz):
        n = int(np.prod(s))
        out.append(np.asarray(flat[o:o + n]).reshape(s, order="F"))
        o += n
    return out, o


def flat4(arrs):
    return np.concatenate([np.asarray(a).ravel(order="F") for a in arrs])


def relerr(a, b):
    a, b = np.asarray(a).ravel(), np.asarray(b).ravel()
    den = np.max(np.abs(b)) if b.size else 1.0
    return float(np.max(np.abs(a - b)) / den) if b.size and den > 0 else float(np.max(np.abs(a - b), initial=0.0))


@pytest.fixture(scope="session")
def oracle():
    import oracle as O

    O.lib()
    return O
