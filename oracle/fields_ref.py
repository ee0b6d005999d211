"""CPU restatement of the reference's post-processing -- TEST INFRASTRUCTURE
ONLY (imported by tests/, never by the product path).

gen_fields (src/aphid.f90:2173-2252): B = curl A at the centroids of the
primary cell faces, E = -i omega (A - grad Phi) on the primary edges, from a
solution (A_x, A_y, A_z, Phi) in the reference's shapes.  d(i) = node(i+1) -
node(i) is the metric of read_mesh (src/aphid.f90:351-357).
"""
import numpy as np


def gen_fields(x, xn, yn, zn, omega):
    ax, ay, az, ph = (np.asarray(a) for a in x)
    dx = np.diff(xn)[:, None, None]
    dy = np.diff(yn)[None, :, None]
    dz = np.diff(zn)[None, None, :]
    bx = (az[:, 1:, :] - az[:, :-1, :]) / dy - (ay[:, :, 1:] - ay[:, :, :-1]) / dz
    by = (ax[:, :, 1:] - ax[:, :, :-1]) / dz - (az[1:, :, :] - az[:-1, :, :]) / dx
    bz = (ay[1:, :, :] - ay[:-1, :, :]) / dx - (ax[:, 1:, :] - ax[:, :-1, :]) / dy
    ex = -1j * omega * (ax - (ph[1:, :, :] - ph[:-1, :, :]) / dx)
    ey = -1j * omega * (ay - (ph[:, 1:, :] - ph[:, :-1, :]) / dy)
    ez = -1j * omega * (az - (ph[:, :, 1:] - ph[:, :, :-1]) / dz)
    return [bx, by, bz, ex, ey, ez]


def field_shapes(nx, ny, nz):
    return [(nx, ny - 1, nz - 1), (nx - 1, ny, nz - 1), (nx - 1, ny - 1, nz),
            (nx - 1, ny, nz), (nx, ny - 1, nz), (nx, ny, nz - 1)]


def split(flat, shapes):
    out, off = [], 0
    for s in shapes:
        n = int(np.prod(s))
        out.append(np.asarray(flat[off:off + n]).reshape(s, order="F"))
        off += n
    return out
