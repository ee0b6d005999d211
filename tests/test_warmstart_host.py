"""Host-side inputs and targets of the learned warm start (no GPU)."""
import numpy as np

from aphid_git_amd import meshgen, warmstart


def _b_of(case):
    nx, ny, nz = case.shape
    b = [np.zeros((nx - 1, ny, nz), complex), np.zeros((nx, ny - 1, nz), complex),
         np.zeros((nx, ny, nz - 1), complex), np.zeros((nx, ny, nz), complex)]
    _, i, j, k = case.sources[0]
    b[0][i - 1, j - 1, k - 1] = 1.0  # +X dipole: the x-edge from node (i,j,k)
    return b


def test_source_mask_marks_the_four_cells_around_an_x_edge():
    case = meshgen.random_block_case((8, 8, 8), 1, 10.0)
    b = _b_of(case)
    m = warmstart.source_mask(b, (8, 8, 8))
    _, i, j, k = case.sources[0]
    assert m.sum() == 4
    assert m[i - 1, j - 2:j, k - 2:k].sum() == 4


def test_features_are_bounded_and_ordered_z_fastest():
    case = meshgen.random_block_case((8, 12, 8), 2, 100.0)
    f = warmstart.features(case, _b_of(case))
    assert f.shape == (8 * 12 * 8, 3) and f.dtype == np.float32
    assert np.all(np.abs(f) <= 1.0 + 1e-6)
    sig = case.cells["sigma"]
    np.testing.assert_allclose(f[:, 0].reshape(8, 12, 8), (np.log10(sig) + 4) / 4, atol=1e-6)


def test_target_and_mapping_round_trip_the_owned_unknowns():
    nx, ny, nz = 5, 6, 7
    rng = np.random.default_rng(3)
    shp = [(nx - 1, ny, nz), (nx, ny - 1, nz), (nx, ny, nz - 1), (nx, ny, nz)]
    x = [rng.standard_normal(s) + 1j * rng.standard_normal(s) for s in shp]
    x[3][0, :, :] = x[3][-1, :, :] = 0
    x[3][:, 0, :] = x[3][:, -1, :] = 0
    x[3][:, :, 0] = x[3][:, :, -1] = 0
    t = warmstart.target(x)
    rms = np.sqrt(sum(np.sum(np.abs(a) ** 2) for a in x) / sum(a.size for a in x))
    np.testing.assert_allclose(t.reshape(nx - 1, ny - 1, nz - 1, 8)[..., 2] * rms,
                               x[1][:-1, :, :-1].real, rtol=1e-6, atol=1e-6)
    u = warmstart.cells_to_unknowns(t.astype(np.float64) * rms, (nx, ny, nz))
    # every unknown with a cell of its own comes back (to float32 rounding)
    np.testing.assert_allclose(u[0][:, :-1, :-1], x[0][:, :-1, :-1], rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(u[2][:-1, :-1, :], x[2][:-1, :-1, :], rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(u[3], x[3], rtol=1e-6, atol=1e-6)
    # edges on the last planes take the nearest cell across
    np.testing.assert_allclose(u[0][:, -1, :-1], x[0][:, -2, :-1], rtol=1e-6, atol=1e-6)
