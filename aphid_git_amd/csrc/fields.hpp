// fields.hpp -- post-processing of a solution: B = curl A on the primary
// cell faces and E = -i omega (A - grad Phi) on the primary edges
// (gen_fields, src/aphid.f90:2173-2252), in the node-grid device layout.
#pragma once
#include "setup.hpp"

// x: the solution (4 blocks of nn: Ax, Ay, Az, Phi).  ob: Bx, By, Bz and
// oe: Ex, Ey, Ez (3 blocks of nn each); entries outside each array's shape
// (Bx (nx,ny-1,nz-1), By (nx-1,ny,nz-1), Bz (nx-1,ny-1,nz); E as A) are 0.
// Dx/Dy/Dz are the metric arrays d(0:n), d(i) = node(i+1) - node(i).  The
// arithmetic follows the reference's expressions term by term.
__global__ void k_fields(const c128* __restrict__ x, Dims d, const double* __restrict__ Dx,
                         const double* __restrict__ Dy, const double* __restrict__ Dz,
                         double omega, c128* __restrict__ ob, c128* __restrict__ oe) {
#pragma clang fp contract(off)
  const long long nn = (long long)d.nx * d.ny * d.nz;
  const c128* Ax = x;
  const c128* Ay = x + nn;
  const c128* Az = x + 2 * nn;
  const c128* Ph = x + 3 * nn;
  const long long sj = d.nz, si = (long long)d.ny * d.nz;
  auto dv = [](c128 a, double s) { return c128{a.re / s, a.im / s}; };
  auto sub = [](c128 a, c128 b) { return c128{a.re - b.re, a.im - b.im}; };
  auto add = [](c128 a, c128 b) { return c128{a.re + b.re, a.im + b.im}; };
  // -II*omega*X = -((0, omega) * X)
  auto miw = [omega](c128 X) {
    const double re = 0.0 * X.re - omega * X.im, im = 0.0 * X.im + omega * X.re;
    return c128{-re, -im};
  };
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < nn;
       q += (long long)gridDim.x * blockDim.x) {
    const int k = (int)(q % d.nz), j = (int)((q / d.nz) % d.ny), i = (int)(q / si);
    const double dx = Dx[i + 1], dy = Dy[j + 1], dz = Dz[k + 1];
    const c128 z = {0.0, 0.0};
    c128 bx = z, by = z, bz = z, ex = z, ey = z, ez = z;
    if (j < d.ny - 1 && k < d.nz - 1)
      bx = add(sub(sub(dv(Az[q + sj], dy), dv(Ay[q + 1], dz)), dv(Az[q], dy)), dv(Ay[q], dz));
    if (i < d.nx - 1 && k < d.nz - 1)
      by = add(sub(sub(dv(Ax[q + 1], dz), dv(Az[q + si], dx)), dv(Ax[q], dz)), dv(Az[q], dx));
    if (i < d.nx - 1 && j < d.ny - 1)
      bz = add(sub(sub(dv(Ay[q + si], dx), dv(Ax[q + sj], dy)), dv(Ay[q], dx)), dv(Ax[q], dy));
    if (i < d.nx - 1) ex = miw(sub(Ax[q], dv(sub(Ph[q + si], Ph[q]), dx)));
    if (j < d.ny - 1) ey = miw(sub(Ay[q], dv(sub(Ph[q + sj], Ph[q]), dy)));
    if (k < d.nz - 1) ez = miw(sub(Az[q], dv(sub(Ph[q + 1], Ph[q]), dz)));
    ob[q] = bx;
    ob[nn + q] = by;
    ob[2 * nn + q] = bz;
    oe[q] = ex;
    oe[nn + q] = ey;
    oe[2 * nn + q] = ez;
  }
}
