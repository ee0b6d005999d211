// setup.hpp -- once-per-frequency device kernels: Fortran<->device layout
// permutations, sigma-hat models, k^2 grids, Jacobi diagonal and its inverse.
//
// These follow the reference's arithmetic order statement by statement
// (FMA contraction off), so the device k^2 and diagonal agree with the
// reference to rounding of the complex divide.
#pragma once
#include "cplx.hpp"

#define APHID_PI 3.141592653589793
#define APHID_MU0 (APHID_PI * 4.0e-7)
#define APHID_EPS0 8.8541878176e-12

struct Dims {
  int nx, ny, nz;
  long long nn;
  __device__ __forceinline__ void ijk(long long idx, int& i, int& j, int& k) const {
    const long long pl = (long long)ny * nz;
    i = (int)(idx / pl);
    const long long r = idx - (long long)i * pl;
    j = (int)(r / nz);
    k = (int)(r - (long long)j * nz);
  }
};

// Fortran-order array of shape (n1,n2,n3) -> device layout on the node grid
// (z fastest), zero outside (n1,n2,n3).
template <typename T>
__global__ void k_f2d(const cplx<T>* __restrict__ src, int n1, int n2, int n3,
                      cplx<T>* __restrict__ dst, Dims d) {
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < d.nn;
       q += (long long)gridDim.x * blockDim.x) {
    int i, j, k;
    d.ijk(q, i, j, k);
    cplx<T> v = {T(0), T(0)};
    if (i < n1 && j < n2 && k < n3) v = src[((long long)k * n2 + j) * n1 + i];
    dst[q] = v;
  }
}

// device layout -> Fortran-order (n1,n2,n3)
template <typename T>
__global__ void k_d2f(const cplx<T>* __restrict__ src, int n1, int n2, int n3,
                      cplx<T>* __restrict__ dst, Dims d) {
  const long long n = (long long)n1 * n2 * n3;
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n;
       q += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(q % n1);
    const long long r = q / n1;
    const int j = (int)(r % n2);
    const int k = (int)(r / n2);
    dst[q] = src[((long long)i * d.ny + j) * d.nz + k];
  }
}

// Cell parameters (FP32, Fortran order) -> sigma-hat (FP64 complex) in the
// device cell layout ((i*(ny-1)+j)*(nz-1)+k), read_mesh src/aphid.f90:379-449.
__global__ void k_sigma(char model, double omega, const float* __restrict__ a,
                        const float* __restrict__ b, const float* __restrict__ c,
                        const float* __restrict__ dd, c128* __restrict__ sig, Dims d) {
#pragma clang fp contract(off)
  const int cx = d.nx - 1, cy = d.ny - 1, cz = d.nz - 1;
  const long long n = (long long)cx * cy * cz;
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n;
       q += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(q / ((long long)cy * cz));
    const long long r = q - (long long)i * cy * cz;
    const int j = (int)(r / cz), k = (int)(r - (long long)j * cz);
    const long long f = ((long long)k * cy + j) * cx + i;
    c128 s = {0.0, 0.0};
    switch (model) {
      case 'R':  // sig_re + II*omega*EPS0
        s = {(double)a[f], omega * APHID_EPS0};
        break;
      case 'C':  // sig_re + II*omega*EPS0*eps_rel
        s = {(double)a[f], (omega * APHID_EPS0) * (double)b[f]};
        break;
      case 'D': {  // sig_re*(1 - eta/(1 + (II*omega*tau)**cee)); (i w tau)^c =
                   // (w tau)^c (cos(c pi/2) + i sin(c pi/2)) for w tau > 0
        const double wt = omega * (double)c[f], ce = (double)dd[f];
        c128 pw = {0.0, 0.0};
        if (wt > 0.0) {
          const double m = pow(wt, ce), ph = ce * (0.5 * APHID_PI);
          pw = {m * cos(ph), m * sin(ph)};
        }
        const c128 den = {1.0 + pw.re, pw.im};
        const c128 fr = cdiv(c128{(double)b[f], 0.0}, den);
        s = (double)a[f] * c128{1.0 - fr.re, -fr.im};
        break;
      }
      case 'I':  // II*omega*EPS0*eps_rel
        s = {0.0, (omega * APHID_EPS0) * (double)b[f]};
        break;
    }
    sig[q] = s;
  }
}

// k^2 grids (gen_k2_{x,y,z,p}grid, src/aphid.f90:1289-1535).  The
// reference fills the interior and then copies boundary faces, edges and
// corners from the nearest interior value before scaling by II*omega*MU0;
// evaluating the interior formula at the clamped index gives exactly those
// copies.  Padding entries of the edge grids are 0.
// d1: dx(0:nx) etc. in Fortran index (length n+1).
__global__ void k_k2(const c128* __restrict__ sig, const double* __restrict__ Dx,
                     const double* __restrict__ Dy, const double* __restrict__ Dz, double omega,
                     c128* __restrict__ k2, Dims d) {
#pragma clang fp contract(off)
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  const int cy = ny - 1, cz = nz - 1;
  auto S = [&](int I, int J, int K) -> c128 {  // 1-based cell (I,J,K)
    return sig[((long long)(I - 1) * cy + (J - 1)) * cz + (K - 1)];
  };
  auto scale = [&](c128 z) -> c128 {  // z*II*omega*MU0 left to right
    c128 t = {-z.im, z.re};
    t = {t.re * omega, t.im * omega};
    return {t.re * APHID_MU0, t.im * APHID_MU0};
  };
  auto clampi = [](int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); };
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < d.nn;
       q += (long long)gridDim.x * blockDim.x) {
    int i0, j0, k0;
    d.ijk(q, i0, j0, k0);
    const int I = i0 + 1, J = j0 + 1, K = k0 + 1;
    // x-edges (:1306-1334)
    c128 vx = {0.0, 0.0};
    if (I <= nx - 1) {
      const int j = clampi(J, 2, ny - 1), k = clampi(K, 2, nz - 1);
      const double dyj = Dy[j], dyjm = Dy[j - 1], dzk = Dz[k], dzkm = Dz[k - 1];
      const double c1 = dyjm * dzkm, c2 = dyj * dzkm, c3 = dyj * dzk, c4 = dyjm * dzk;
      c128 s = c1 * S(I, j - 1, k - 1);
      s += c2 * S(I, j, k - 1);
      s += c3 * S(I, j, k);
      s += c4 * S(I, j - 1, k);
      vx = scale(divr(s, (dyjm + dyj) * (dzkm + dzk)));
    }
    // y-edges (:1359-1387)
    c128 vy = {0.0, 0.0};
    if (J <= ny - 1) {
      const int i = clampi(I, 2, nx - 1), k = clampi(K, 2, nz - 1);
      const double dxi = Dx[i], dxim = Dx[i - 1], dzk = Dz[k], dzkm = Dz[k - 1];
      const double c1 = dxim * dzkm, c2 = dxi * dzkm, c3 = dxi * dzk, c4 = dxim * dzk;
      c128 s = c1 * S(i - 1, J, k - 1);
      s += c2 * S(i, J, k - 1);
      s += c3 * S(i, J, k);
      s += c4 * S(i - 1, J, k);
      vy = scale(divr(s, (dxim + dxi) * (dzkm + dzk)));
    }
    // z-edges (:1412-1440)
    c128 vz = {0.0, 0.0};
    if (K <= nz - 1) {
      const int i = clampi(I, 2, nx - 1), j = clampi(J, 2, ny - 1);
      const double dxi = Dx[i], dxim = Dx[i - 1], dyj = Dy[j], dyjm = Dy[j - 1];
      const double c1 = dxim * dyjm, c2 = dxi * dyjm, c3 = dxi * dyj, c4 = dxim * dyj;
      c128 s = c1 * S(i - 1, j - 1, K);
      s += c2 * S(i, j - 1, K);
      s += c3 * S(i, j, K);
      s += c4 * S(i - 1, j, K);
      vz = scale(divr(s, (dxim + dxi) * (dyjm + dyj)));
    }
    // nodes (:1465-1531)
    c128 vp;
    {
      const int i = clampi(I, 2, nx - 1), j = clampi(J, 2, ny - 1), k = clampi(K, 2, nz - 1);
      const double dxi = Dx[i], dxim = Dx[i - 1], dyj = Dy[j], dyjm = Dy[j - 1];
      const double dzk = Dz[k], dzkm = Dz[k - 1];
      c128 s = (dxim * dyjm * dzk) * S(i - 1, j - 1, k);
      s += (dxi * dyjm * dzk) * S(i, j - 1, k);
      s += (dxi * dyj * dzk) * S(i, j, k);
      s += (dxim * dyj * dzk) * S(i - 1, j, k);
      s += (dxim * dyjm * dzkm) * S(i - 1, j - 1, k - 1);
      s += (dxi * dyjm * dzkm) * S(i, j - 1, k - 1);
      s += (dxi * dyj * dzkm) * S(i, j, k - 1);
      s += (dxim * dyj * dzkm) * S(i - 1, j, k - 1);
      vp = scale(divr(s, (dxim + dxi) * (dyjm + dyj) * (dzkm + dzk)));
    }
    k2[q] = vx;
    k2[q + d.nn] = vy;
    k2[q + 2 * d.nn] = vz;
    k2[q + 3 * d.nn] = vp;
  }
}

// Compact k^2 for the stencil (OpArgs::k2rec / k2re / k2pc): per node the
// pairs {k2x.im, k2p.im} (xp[2q], xp[2q+1]) and {k2y.im, k2z.im}
// (yz[2q], yz[2q+1]); the real parts of the i = 0,
// j = 0 and k = 0 slices of the edge grids as 2-D tables rx(j,k), ry(i,k),
// rz(i,j); and per node one byte c with k2p.re = the double whose bit
// pattern is pbase + c (modulo 2^64).  For a sigma-hat whose imaginary part e is the same
// in every cell (model R), each edge grid's real part is e times metric
// weights of the two transverse axes (gen_k2_{x,y,z}grid,
// src/aphid.f90:1306-1440), so the tables hold every real part; and k2p's
// real part is -omega mu0 e times (sum of the eight cell weights) / (their
// total) (gen_k2_pgrid, :1465-1531), a ratio that is 1 up to rounding, so
// every node's value lies within a few units in the last place of every
// other's.  k_k2c_check verifies both, entry by entry.
__global__ void k_k2c_extract(const c128* __restrict__ k2, double* __restrict__ xp,
                              double* __restrict__ yz, double* __restrict__ rx, double* __restrict__ ry,
                              double* __restrict__ rz, unsigned char* __restrict__ pcode,
                              unsigned long long pbase, Dims d) {
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < d.nn;
       q += (long long)gridDim.x * blockDim.x) {
    int i, j, k;
    d.ijk(q, i, j, k);
    xp[2 * q] = k2[q].im;
    xp[2 * q + 1] = k2[3 * d.nn + q].im;
    yz[2 * q] = k2[d.nn + q].im;
    yz[2 * q + 1] = k2[2 * d.nn + q].im;
    pcode[q] = (unsigned char)((unsigned long long)__double_as_longlong(k2[3 * d.nn + q].re) - pbase);
    if (i == 0) rx[j * d.nz + k] = k2[q].re;
    if (j == 0) ry[i * d.nz + k] = k2[d.nn + q].re;
    if (k == 0) rz[i * d.ny + j] = k2[2 * d.nn + q].re;
  }
}

// Sets *bad = 1 if any real part of an edge grid (padding excluded) differs
// from its table entry in any bit, or any k2p real part's bit pattern lies
// outside pbase + [0, 255] (modulo 2^64).
__global__ void k_k2c_check(const c128* __restrict__ k2, const double* __restrict__ rx,
                            const double* __restrict__ ry, const double* __restrict__ rz,
                            unsigned long long pbase, Dims d, int* bad) {
  bool diff = false;
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < d.nn;
       q += (long long)gridDim.x * blockDim.x) {
    int i, j, k;
    d.ijk(q, i, j, k);
    if (i < d.nx - 1 && __double_as_longlong(k2[q].re) != __double_as_longlong(rx[j * d.nz + k]))
      diff = true;
    if (j < d.ny - 1 &&
        __double_as_longlong(k2[d.nn + q].re) != __double_as_longlong(ry[i * d.nz + k]))
      diff = true;
    if (k < d.nz - 1 &&
        __double_as_longlong(k2[2 * d.nn + q].re) != __double_as_longlong(rz[i * d.ny + j]))
      diff = true;
    const unsigned long long c =
        (unsigned long long)__double_as_longlong(k2[3 * d.nn + q].re) - pbase;
    if (c > 255ull) diff = true;
  }
  if (diff) *bad = 1;  // every writer stores the same value
}

// Jacobi diagonal (gen_diag, src/aphid.f90:1542-1725), including its
// boundary quirks: transverse centred spacings recomputed as
// 0.5*(d(j)+d(j-1)) with d(0)=d(1), d(n)=d(n-1); diag_p = 1 on boundary
// nodes.  Also writes inv = 1/diag (src/aphid_qmr.f90:178-181), with 0 in
// the padding.
__global__ void k_diag(const c128* __restrict__ k2, const double* __restrict__ Dx,
                       const double* __restrict__ Dy, const double* __restrict__ Dz,
                       c128* __restrict__ diag, c128* __restrict__ inv, Dims d) {
#pragma clang fp contract(off)
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  const long long nn = d.nn;
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < nn;
       q += (long long)gridDim.x * blockDim.x) {
    int i0, j0, k0;
    d.ijk(q, i0, j0, k0);
    const int I = i0 + 1, J = j0 + 1, K = k0 + 1;
    const double dxi = Dx[I], dxim = Dx[I - 1], dxci = 0.5 * (dxi + dxim);
    const double dyj = Dy[J], dyjm = Dy[J - 1], dycj = 0.5 * (dyj + dyjm);
    const double dzk = Dz[K], dzkm = Dz[K - 1], dzck = 0.5 * (dzk + dzkm);
    // node-type masked coefficients
    const double ax4 = (double)(I / nx - 1) / dxi / dxci;
    const double ax5 = (double)((nx - I) / (nx - 1) - 1) / dxim / dxci;
    const double ay4 = (double)(J / ny - 1) / dyj / dycj;
    const double ay5 = (double)((ny - J) / (ny - 1) - 1) / dyjm / dycj;
    const double az6 = (double)(K / nz - 1) / dzk / dzck;
    const double az7 = (double)((nz - K) / (nz - 1) - 1) / dzkm / dzck;
    c128 dgx = {0.0, 0.0}, dgy = {0.0, 0.0}, dgz = {0.0, 0.0}, dgp = {1.0, 0.0};
    if (I <= nx - 1) {  // x-edges (:1572-1605)
      const double dxip = Dx[I + 1], dxcip = 0.5 * (dxi + dxip);
      const double c2 = (double)(I / (nx - 1) - 1) / dxcip / dxi;
      const double c3 = (double)((nx - I) / (nx - 1) - 1) / dxci / dxi;
      const double sm = c2 + c3 + ay4 + ay5 + az6 + az7;
      const c128 kk = k2[q];
      dgx = (dxi * dycj * dzck) * c128{kk.re - sm, kk.im};
    }
    if (J <= ny - 1) {  // y-edges (:1610-1644)
      const double dyjp = Dy[J + 1], dycjp = 0.5 * (dyj + dyjp);
      const double c2 = (double)(J / (ny - 1) - 1) / dycjp / dyj;
      const double c3 = (double)((ny - J) / (ny - 1) - 1) / dycj / dyj;
      const double sm = c2 + c3 + ax4 + ax5 + az6 + az7;
      const c128 kk = k2[q + nn];
      dgy = (dxci * dyj * dzck) * c128{kk.re - sm, kk.im};
    }
    if (K <= nz - 1) {  // z-edges (:1650-1682)
      const double dzkp = Dz[K + 1], dzckp = 0.5 * (dzk + dzkp);
      const double c2 = (double)(K / (nz - 1) - 1) / dzckp / dzk;
      const double c3 = (double)((nz - K) / (nz - 1) - 1) / dzck / dzk;
      const double sm = c2 + c3 + ax4 + ax5 + ay4 + ay5;
      const c128 kk = k2[q + 2 * nn];
      dgz = (dxci * dycj * dzk) * c128{kk.re - sm, kk.im};
    }
    if (I > 1 && I < nx && J > 1 && J < ny && K > 1 && K < nz) {  // nodes (:1684-1723)
      const long long pl = (long long)ny * nz;
      const c128 kp = k2[q + 3 * nn];
      const c128 s2 = k2[q], s3 = k2[q - pl];                 // k2x(I), k2x(I-1)
      const c128 s4 = k2[q + nn], s5 = k2[q + nn - nz];       // k2y(J), k2y(J-1)
      const c128 s6 = k2[q + 2 * nn], s7 = k2[q + 2 * nn - 1];  // k2z(K), k2z(K-1)
      auto term = [](double m, c128 s, double a, double b) {
        return divr(divr(m * s, a), b);
      };
      c128 sm = term((double)(I / nx - 1), s2, dxi, dxci);
      sm += term((double)((nx - I) / (nx - 1) - 1), s3, dxim, dxci);
      sm += term((double)(J / ny - 1), s4, dyj, dycj);
      sm += term((double)((ny - J) / (ny - 1) - 1), s5, dyjm, dycj);
      sm += term((double)(K / nz - 1), s6, dzk, dzck);
      sm += term((double)((nz - K) / (nz - 1) - 1), s7, dzkm, dzck);
      dgp = (dxci * dycj * dzck) * (kp * kp - sm);
    }
    diag[q] = dgx;
    diag[q + nn] = dgy;
    diag[q + 2 * nn] = dgz;
    diag[q + 3 * nn] = dgp;
    const c128 one = {1.0, 0.0}, z = {0.0, 0.0};
    inv[q] = (I <= nx - 1) ? cdiv(one, dgx) : z;
    inv[q + nn] = (J <= ny - 1) ? cdiv(one, dgy) : z;
    inv[q + 2 * nn] = (K <= nz - 1) ? cdiv(one, dgz) : z;
    inv[q + 3 * nn] = cdiv(one, dgp);
  }
}

// 1/diag from an uploaded diagonal (padding stays 0)
__global__ void k_invert(const c128* __restrict__ diag, c128* __restrict__ inv, Dims d) {
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < 4 * d.nn;
       q += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(q / d.nn);
    int i, j, k;
    d.ijk(q - (long long)c * d.nn, i, j, k);
    const bool pad = (c == 0 && i == nx - 1) || (c == 1 && j == ny - 1) || (c == 2 && k == nz - 1);
    inv[q] = pad ? c128{0.0, 0.0} : cdiv(c128{1.0, 0.0}, diag[q]);
  }
}

template <typename To, typename From>
__global__ void k_convert(const cplx<From>* __restrict__ a, cplx<To>* __restrict__ b, long long n) {
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n;
       q += (long long)gridDim.x * blockDim.x)
    b[q] = ccast<To>(a[q]);
}
