"""Pins of the post-processing restatement (oracle/fields_ref.py) against
the reference's own gen_fields output and against what the mathematics
fixes (no GPU)."""
import os

import numpy as np
import pytest

import fields_ref
from conftest import GOLDEN, load_golden, split4


@pytest.mark.parametrize("name", ["fields_R_7x6x5", "fields_C_8x7x6"])
def test_restatement_reproduces_the_reference_fields(name):
    case, g = load_golden(os.path.join(GOLDEN, name + ".npz"))
    xn, yn, zn = g["xn"], g["yn"], g["zn"]
    nx, ny, nz = len(xn), len(yn), len(zn)
    x, _ = split4(g["qmr_sol"], nx, ny, nz)
    omega = 2 * np.pi * float(g["freq"])
    got = fields_ref.gen_fields(x, xn, yn, zn, omega)
    ref = fields_ref.split(g["fields"], fields_ref.field_shapes(nx, ny, nz))
    for a, r in zip(got, ref):
        scale = np.abs(r).max()
        assert np.abs(a - r).max() <= 1e-13 * scale


def test_discrete_div_curl_and_curl_grad_vanish():
    rng = np.random.default_rng(5)
    nx, ny, nz = 6, 7, 5
    xn, yn, zn = (np.cumsum(rng.uniform(0.5, 2.0, n)) for n in (nx, ny, nz))
    shp = [(nx - 1, ny, nz), (nx, ny - 1, nz), (nx, ny, nz - 1), (nx, ny, nz)]
    x = [rng.standard_normal(s) + 1j * rng.standard_normal(s) for s in shp]
    bx, by, bz, *_ = fields_ref.gen_fields(x, xn, yn, zn, 1.0)
    div = ((bx[1:] - bx[:-1]) / np.diff(xn)[:, None, None]
           + (by[:, 1:] - by[:, :-1]) / np.diff(yn)[None, :, None]
           + (bz[:, :, 1:] - bz[:, :, :-1]) / np.diff(zn)[None, None, :])
    assert np.abs(div).max() <= 1e-12 * max(np.abs(bx).max(), 1.0)
    # A = grad(f) on the edges gives B = 0; then E = -i w (grad f - grad Phi)
    f = rng.standard_normal((nx, ny, nz))
    g = [np.diff(f, axis=0) / np.diff(xn)[:, None, None], np.diff(f, axis=1) / np.diff(yn)[None, :, None],
         np.diff(f, axis=2) / np.diff(zn)[None, None, :], f.astype(complex)]
    out = fields_ref.gen_fields(g, xn, yn, zn, 2.0)
    for b in out[:3]:
        assert np.abs(b).max() <= 1e-12 * np.abs(f).max()
    for e in out[3:]:  # Phi = f cancels the gradient exactly
        assert np.abs(e).max() <= 1e-12 * np.abs(f).max()
