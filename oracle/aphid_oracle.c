/*
 * aphid_oracle.c -- CPU oracle for the A-Phi Jacobi-QMR solve.
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE: a plain C restatement of the
 * reference algorithm, used only as the checker (tests/, smoke()) and as the
 * "port" CPU baseline in bench.py.  Each function cites the reference lines
 * it restates.  Arithmetic is kept in the reference's order of operations
 * (left-to-right sums, divide-then-divide coefficient formulas, copy-then-
 * scale k^2) so that it agrees with the compiled reference to rounding.
 *
 * Parity pins (tests/test_oracle_*.py): golden vectors written by the
 * reference compiled from its own sources (oracle/Makefile, target `ref`),
 * plus closed forms and invariants that need no reference (symmetry of the
 * assembled operator, uniform-sigma k^2, dense direct solve, the Krylov
 * quasi-minimal-residual characterisation of the QMR iterates).
 */
#include "aphid_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- Fortran-order index helpers (1-based i,j,k as in the reference) ---- */
#define IDX(i, j, k, n1, n2) \
  ((((int64_t)(k)-1) * (int64_t)(n2) + ((int64_t)(j)-1)) * (int64_t)(n1) + ((int64_t)(i)-1))

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* complex / real, component-wise */
static inline ocplx cdivr(ocplx z, double r) {
  return CMPLX(creal(z) / r, cimag(z) / r);
}
/* Fortran-rule complex product (no Annex-G NaN recovery) */
static inline ocplx cmul(ocplx a, ocplx b) {
  return CMPLX(creal(a) * creal(b) - cimag(a) * cimag(b),
               creal(a) * cimag(b) + cimag(a) * creal(b));
}
/* Smith's complex division */
static inline ocplx cdiv(ocplx a, ocplx b) {
  double ar = creal(a), ai = cimag(a), br = creal(b), bi = cimag(b);
  if (fabs(br) >= fabs(bi)) {
    double q = bi / br, den = br + bi * q;
    return CMPLX((ar + ai * q) / den, (ai - ar * q) / den);
  } else {
    double q = br / bi, den = bi + br * q;
    return CMPLX((ar * q + ai) / den, (ai * q - ar) / den);
  }
}
static inline ocplx rmul(double r, ocplx z) { return CMPLX(r * creal(z), r * cimag(z)); }
/* z * II * omega * MU0, evaluated left to right (src/aphid.f90:1336) */
static inline ocplx scale_iwmu(ocplx z, double omega) {
  ocplx t = CMPLX(-cimag(z), creal(z));
  t = CMPLX(creal(t) * omega, cimag(t) * omega);
  return CMPLX(creal(t) * OM_MU0, cimag(t) * OM_MU0);
}

/* ------------------------------------------------------------------------ */
/* metrics: src/aphid.f90:341-361                                            */
void om_metrics(int n, const double *c, double *d, double *dc) {
  for (int i = 1; i <= n - 1; ++i) d[i] = c[i] - c[i - 1]; /* xn(i+1)-xn(i) */
  d[0] = d[1];
  d[n] = d[n - 1];
  for (int i = 2; i <= n - 1; ++i) dc[i] = 0.5 * (d[i - 1] + d[i]);
  dc[1] = dc[2];
  dc[n] = dc[n - 1];
  dc[0] = 0.0;
}

/* sigma-hat: src/aphid.f90:379-449 (inputs are real(4), :288) */
int om_sigma(char model, double omega, int64_t ncell, const float *a,
             const float *b, const float *c, const float *d, ocplx *sig) {
  const ocplx II = CMPLX(0.0, 1.0);
  for (int64_t q = 0; q < ncell; ++q) {
    switch (model) {
      case 'R': /* sig_re + II*omega*EPS0 (:385) */
        sig[q] = CMPLX((double)a[q], omega * OM_EPS0);
        break;
      case 'C': { /* sig_re + ((II*omega)*EPS0)*eps_rel (:404) */
        double im = (omega * OM_EPS0) * (double)b[q];
        sig[q] = CMPLX((double)a[q], im);
        break;
      }
      case 'D': { /* sig_re*(1 - eta/(1 + (II*omega*tau)**cee)) (:432) */
        ocplx iwt = CMPLX(0.0, omega * (double)c[q]);
        ocplx pw = cpow(iwt, CMPLX((double)d[q], 0.0));
        ocplx den = CMPLX(1.0 + creal(pw), cimag(pw));
        ocplx fr = cdiv(CMPLX((double)b[q], 0.0), den);
        ocplx one_m = CMPLX(1.0 - creal(fr), -cimag(fr));
        sig[q] = rmul((double)a[q], one_m);
        break;
      }
      case 'I': /* II*omega*EPS0*eps_rel (:445) */
        sig[q] = CMPLX(0.0, (omega * OM_EPS0) * (double)b[q]);
        break;
      default:
        return -1;
    }
  }
  (void)II;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* k^2 grids: src/aphid.f90:1289-1535                                        */
void om_k2(int nx, int ny, int nz, const double *dx, const double *dy,
           const double *dz, const ocplx *sig, double omega, ocplx *k2x,
           ocplx *k2y, ocplx *k2z, ocplx *k2p) {
  const int cx = nx - 1, cy = ny - 1; /* cell array extents */
#define SIG(i, j, k) sig[IDX(i, j, k, cx, cy)]
  /* x-edges (nx-1,ny,nz): 4-cell dy*dz weighted mean (:1306-1330) */
#define KX(i, j, k) k2x[IDX(i, j, k, nx - 1, ny)]
  for (int i = 1; i <= nx - 1; ++i)
    for (int j = 2; j <= ny - 1; ++j)
      for (int k = 2; k <= nz - 1; ++k) {
        double dyj = dy[j], dyjm = dy[j - 1], dzk = dz[k], dzkm = dz[k - 1];
        double c1 = dyjm * dzkm, c2 = dyj * dzkm, c3 = dyj * dzk, c4 = dyjm * dzk;
        ocplx s = rmul(c1, SIG(i, j - 1, k - 1));
        s += rmul(c2, SIG(i, j, k - 1));
        s += rmul(c3, SIG(i, j, k));
        s += rmul(c4, SIG(i, j - 1, k));
        KX(i, j, k) = cdivr(s, (dyjm + dyj) * (dzkm + dzk));
      }
  for (int i = 1; i <= nx - 1; ++i)
    for (int j = 2; j <= ny - 1; ++j) {
      KX(i, j, 1) = KX(i, j, 2);
      KX(i, j, nz) = KX(i, j, nz - 1);
    }
  for (int k = 1; k <= nz; ++k)
    for (int i = 1; i <= nx - 1; ++i) {
      KX(i, 1, k) = KX(i, 2, k);
      KX(i, ny, k) = KX(i, ny - 1, k);
    }
  /* y-edges (nx,ny-1,nz) (:1359-1387) */
#define KY(i, j, k) k2y[IDX(i, j, k, nx, ny - 1)]
  for (int i = 2; i <= nx - 1; ++i)
    for (int j = 1; j <= ny - 1; ++j)
      for (int k = 2; k <= nz - 1; ++k) {
        double dxi = dx[i], dxim = dx[i - 1], dzk = dz[k], dzkm = dz[k - 1];
        double c1 = dxim * dzkm, c2 = dxi * dzkm, c3 = dxi * dzk, c4 = dxim * dzk;
        ocplx s = rmul(c1, SIG(i - 1, j, k - 1));
        s += rmul(c2, SIG(i, j, k - 1));
        s += rmul(c3, SIG(i, j, k));
        s += rmul(c4, SIG(i - 1, j, k));
        KY(i, j, k) = cdivr(s, (dxim + dxi) * (dzkm + dzk));
      }
  for (int i = 2; i <= nx - 1; ++i)
    for (int j = 1; j <= ny - 1; ++j) {
      KY(i, j, 1) = KY(i, j, 2);
      KY(i, j, nz) = KY(i, j, nz - 1);
    }
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny - 1; ++j) {
      KY(1, j, k) = KY(2, j, k);
      KY(nx, j, k) = KY(nx - 1, j, k);
    }
  /* z-edges (nx,ny,nz-1) (:1412-1440) */
#define KZ(i, j, k) k2z[IDX(i, j, k, nx, ny)]
  for (int i = 2; i <= nx - 1; ++i)
    for (int j = 2; j <= ny - 1; ++j)
      for (int k = 1; k <= nz - 1; ++k) {
        double dxi = dx[i], dxim = dx[i - 1], dyj = dy[j], dyjm = dy[j - 1];
        double c1 = dxim * dyjm, c2 = dxi * dyjm, c3 = dxi * dyj, c4 = dxim * dyj;
        ocplx s = rmul(c1, SIG(i - 1, j - 1, k));
        s += rmul(c2, SIG(i, j - 1, k));
        s += rmul(c3, SIG(i, j, k));
        s += rmul(c4, SIG(i - 1, j, k));
        KZ(i, j, k) = cdivr(s, (dxim + dxi) * (dyjm + dyj));
      }
  for (int k = 1; k <= nz - 1; ++k)
    for (int j = 2; j <= ny - 1; ++j) {
      KZ(1, j, k) = KZ(2, j, k);
      KZ(nx, j, k) = KZ(nx - 1, j, k);
    }
  for (int k = 1; k <= nz - 1; ++k)
    for (int i = 1; i <= nx; ++i) {
      KZ(i, 1, k) = KZ(i, 2, k);
      KZ(i, ny, k) = KZ(i, ny - 1, k);
    }
  /* nodes (nx,ny,nz): 8-cell volume-weighted mean (:1465-1497) */
#define KP(i, j, k) k2p[IDX(i, j, k, nx, ny)]
  for (int i = 2; i <= nx - 1; ++i)
    for (int j = 2; j <= ny - 1; ++j)
      for (int k = 2; k <= nz - 1; ++k) {
        double dxi = dx[i], dxim = dx[i - 1], dyj = dy[j], dyjm = dy[j - 1];
        double dzk = dz[k], dzkm = dz[k - 1];
        double c[8] = {dxim * dyjm * dzk,  dxi * dyjm * dzk,  dxi * dyj * dzk,
                       dxim * dyj * dzk,   dxim * dyjm * dzkm, dxi * dyjm * dzkm,
                       dxi * dyj * dzkm,   dxim * dyj * dzkm};
        ocplx s = rmul(c[0], SIG(i - 1, j - 1, k));
        s += rmul(c[1], SIG(i, j - 1, k));
        s += rmul(c[2], SIG(i, j, k));
        s += rmul(c[3], SIG(i - 1, j, k));
        s += rmul(c[4], SIG(i - 1, j - 1, k - 1));
        s += rmul(c[5], SIG(i, j - 1, k - 1));
        s += rmul(c[6], SIG(i, j, k - 1));
        s += rmul(c[7], SIG(i - 1, j, k - 1));
        KP(i, j, k) = cdivr(s, (dxim + dxi) * (dyjm + dyj) * (dzkm + dzk));
      }
  /* face, edge and corner copies, in the reference's statement order
     (:1499-1531) */
  for (int k = 2; k <= nz - 1; ++k)
    for (int j = 2; j <= ny - 1; ++j) {
      KP(1, j, k) = KP(2, j, k);
      KP(nx, j, k) = KP(nx - 1, j, k);
    }
  for (int k = 2; k <= nz - 1; ++k)
    for (int i = 2; i <= nx - 1; ++i) {
      KP(i, 1, k) = KP(i, 2, k);
      KP(i, ny, k) = KP(i, ny - 1, k);
    }
  for (int j = 2; j <= ny - 1; ++j)
    for (int i = 2; i <= nx - 1; ++i) {
      KP(i, j, 1) = KP(i, j, 2);
      KP(i, j, nz) = KP(i, j, nz - 1);
    }
  for (int k = 2; k <= nz - 1; ++k) {
    KP(1, 1, k) = KP(2, 2, k);
    KP(1, ny, k) = KP(2, ny - 1, k);
    KP(nx, 1, k) = KP(nx - 1, 2, k);
    KP(nx, ny, k) = KP(nx - 1, ny - 1, k);
  }
  for (int i = 2; i <= nx - 1; ++i) {
    KP(i, 1, 1) = KP(i, 2, 2);
    KP(i, ny, 1) = KP(i, ny - 1, 2);
    KP(i, 1, nz) = KP(i, 2, nz - 1);
    KP(i, ny, nz) = KP(i, ny - 1, nz - 1);
  }
  for (int j = 2; j <= ny - 1; ++j) {
    KP(1, j, 1) = KP(2, j, 2);
    KP(nx, j, 1) = KP(nx - 1, j, 2);
    KP(1, j, nz) = KP(2, j, nz - 1);
    KP(nx, j, nz) = KP(nx - 1, j, nz - 1);
  }
  KP(1, 1, 1) = KP(2, 2, 2);
  KP(nx, 1, 1) = KP(nx - 1, 2, 2);
  KP(1, ny, 1) = KP(2, ny - 1, 2);
  KP(nx, ny, 1) = KP(nx - 1, ny - 1, 2);
  KP(1, 1, nz) = KP(2, 2, nz - 1);
  KP(nx, 1, nz) = KP(nx - 1, 2, nz - 1);
  KP(1, ny, nz) = KP(2, ny - 1, nz - 1);
  KP(nx, ny, nz) = KP(nx - 1, ny - 1, nz - 1);
  /* scale by II*omega*MU0 after the copies (:1336,1389,1442,1533) */
  int64_t n;
  n = (int64_t)(nx - 1) * ny * nz;
  for (int64_t q = 0; q < n; ++q) k2x[q] = scale_iwmu(k2x[q], omega);
  n = (int64_t)nx * (ny - 1) * nz;
  for (int64_t q = 0; q < n; ++q) k2y[q] = scale_iwmu(k2y[q], omega);
  n = (int64_t)nx * ny * (nz - 1);
  for (int64_t q = 0; q < n; ++q) k2z[q] = scale_iwmu(k2z[q], omega);
  n = (int64_t)nx * ny * nz;
  for (int64_t q = 0; q < n; ++q) k2p[q] = scale_iwmu(k2p[q], omega);
#undef SIG
}

/* ------------------------------------------------------------------------ */
/* Jacobi diagonal: src/aphid.f90:1542-1725                                  */
void om_diag(int nx, int ny, int nz, const double *dx, const double *dy,
             const double *dz, const ocplx *k2x, const ocplx *k2y,
             const ocplx *k2z, const ocplx *k2p, ocplx *dgx, ocplx *dgy,
             ocplx *dgz, ocplx *dgp) {
  int64_t n;
  /* initialised to unity (:1564-1567) */
  n = (int64_t)(nx - 1) * ny * nz;
  for (int64_t q = 0; q < n; ++q) dgx[q] = 1.0;
  n = (int64_t)nx * (ny - 1) * nz;
  for (int64_t q = 0; q < n; ++q) dgy[q] = 1.0;
  n = (int64_t)nx * ny * (nz - 1);
  for (int64_t q = 0; q < n; ++q) dgz[q] = 1.0;
  n = (int64_t)nx * ny * nz;
  for (int64_t q = 0; q < n; ++q) dgp[q] = 1.0;

  /* x-edges (:1572-1605) */
  for (int k = 1; k <= nz; ++k) {
    double dzk = dz[k], dzkm = dz[k - 1], dzck = 0.5 * (dzk + dzkm);
    double c6 = (double)(k / nz - 1) / dzk / dzck;
    double c7 = (double)((nz - k) / (nz - 1) - 1) / dzkm / dzck;
    for (int j = 1; j <= ny; ++j) {
      double dyj = dy[j], dyjm = dy[j - 1], dycj = 0.5 * (dyj + dyjm);
      double c4 = (double)(j / ny - 1) / dyj / dycj;
      double c5 = (double)((ny - j) / (ny - 1) - 1) / dyjm / dycj;
      for (int i = 1; i <= nx - 1; ++i) {
        double dxip = dx[i + 1], dxi = dx[i], dxim = dx[i - 1];
        double dxcip = 0.5 * (dxi + dxip), dxci = 0.5 * (dxi + dxim);
        double vol = dxi * dycj * dzck;
        ocplx k2 = k2x[IDX(i, j, k, nx - 1, ny)];
        double c2 = (double)(i / (nx - 1) - 1) / dxcip / dxi;
        double c3 = (double)((nx - i) / (nx - 1) - 1) / dxci / dxi;
        double sm = c2 + c3 + c4 + c5 + c6 + c7;
        ocplx dg = CMPLX(-sm + creal(k2), cimag(k2));
        dgx[IDX(i, j, k, nx - 1, ny)] = rmul(vol, dg);
      }
    }
  }
  /* y-edges (:1610-1644) */
  for (int k = 1; k <= nz; ++k) {
    double dzk = dz[k], dzkm = dz[k - 1], dzck = 0.5 * (dzk + dzkm);
    double c6 = (double)(k / nz - 1) / dzk / dzck;
    double c7 = (double)((nz - k) / (nz - 1) - 1) / dzkm / dzck;
    for (int j = 1; j <= ny - 1; ++j) {
      double dyjp = dy[j + 1], dyj = dy[j], dyjm = dy[j - 1];
      double dycjp = 0.5 * (dyj + dyjp), dycj = 0.5 * (dyj + dyjm);
      double c2 = (double)(j / (ny - 1) - 1) / dycjp / dyj;
      double c3 = (double)((ny - j) / (ny - 1) - 1) / dycj / dyj;
      for (int i = 1; i <= nx; ++i) {
        double dxi = dx[i], dxim = dx[i - 1], dxci = 0.5 * (dxi + dxim);
        double vol = dxci * dyj * dzck;
        ocplx k2 = k2y[IDX(i, j, k, nx, ny - 1)];
        double c4 = (double)(i / nx - 1) / dxi / dxci;
        double c5 = (double)((nx - i) / (nx - 1) - 1) / dxim / dxci;
        double sm = c2 + c3 + c4 + c5 + c6 + c7;
        ocplx dg = CMPLX(-sm + creal(k2), cimag(k2));
        dgy[IDX(i, j, k, nx, ny - 1)] = rmul(vol, dg);
      }
    }
  }
  /* z-edges (:1650-1682) */
  for (int k = 1; k <= nz - 1; ++k) {
    double dzkp = dz[k + 1], dzk = dz[k], dzkm = dz[k - 1];
    double dzckp = 0.5 * (dzk + dzkp), dzck = 0.5 * (dzk + dzkm);
    double c2 = (double)(k / (nz - 1) - 1) / dzckp / dzk;
    double c3 = (double)((nz - k) / (nz - 1) - 1) / dzck / dzk;
    for (int j = 1; j <= ny; ++j) {
      double dyj = dy[j], dyjm = dy[j - 1], dycj = 0.5 * (dyj + dyjm);
      double c6 = (double)(j / ny - 1) / dyj / dycj;
      double c7 = (double)((ny - j) / (ny - 1) - 1) / dyjm / dycj;
      for (int i = 1; i <= nx; ++i) {
        double dxi = dx[i], dxim = dx[i - 1], dxci = 0.5 * (dxi + dxim);
        double vol = dxci * dycj * dzk;
        ocplx k2 = k2z[IDX(i, j, k, nx, ny)];
        double c4 = (double)(i / nx - 1) / dxi / dxci;
        double c5 = (double)((nx - i) / (nx - 1) - 1) / dxim / dxci;
        double sm = c2 + c3 + c4 + c5 + c6 + c7;
        ocplx dg = CMPLX(-sm + creal(k2), cimag(k2));
        dgz[IDX(i, j, k, nx, ny)] = rmul(vol, dg);
      }
    }
  }
  /* interior nodes (:1684-1723) */
  for (int k = 2; k <= nz - 1; ++k) {
    double dzk = dz[k], dzkm = dz[k - 1], dzck = 0.5 * (dzk + dzkm);
    for (int j = 2; j <= ny - 1; ++j) {
      double dyj = dy[j], dyjm = dy[j - 1], dycj = 0.5 * (dyj + dyjm);
      for (int i = 2; i <= nx - 1; ++i) {
        double dxi = dx[i], dxim = dx[i - 1], dxci = 0.5 * (dxi + dxim);
        double vol = dxci * dycj * dzck;
        ocplx k2 = k2p[IDX(i, j, k, nx, ny)];
        ocplx s2 = k2x[IDX(i < nx - 1 ? i : nx - 1, j, k, nx - 1, ny)];
        ocplx s3 = k2x[IDX(i - 1 > 1 ? i - 1 : 1, j, k, nx - 1, ny)];
        ocplx s4 = k2y[IDX(i, j < ny - 1 ? j : ny - 1, k, nx, ny - 1)];
        ocplx s5 = k2y[IDX(i, j - 1 > 1 ? j - 1 : 1, k, nx, ny - 1)];
        ocplx s6 = k2z[IDX(i, j, k < nz - 1 ? k : nz - 1, nx, ny)];
        ocplx s7 = k2z[IDX(i, j, k - 1 > 1 ? k - 1 : 1, nx, ny)];
        /* (int)*s/d/dc: integer factor times complex, then two divisions */
        ocplx c2 = cdivr(cdivr(rmul((double)(i / nx - 1), s2), dxi), dxci);
        ocplx c3 = cdivr(cdivr(rmul((double)((nx - i) / (nx - 1) - 1), s3), dxim), dxci);
        ocplx c4 = cdivr(cdivr(rmul((double)(j / ny - 1), s4), dyj), dycj);
        ocplx c5 = cdivr(cdivr(rmul((double)((ny - j) / (ny - 1) - 1), s5), dyjm), dycj);
        ocplx c6 = cdivr(cdivr(rmul((double)(k / nz - 1), s6), dzk), dzck);
        ocplx c7 = cdivr(cdivr(rmul((double)((nz - k) / (nz - 1) - 1), s7), dzkm), dzck);
        ocplx sm = c2 + c3 + c4 + c5 + c6 + c7;
        ocplx dg = -sm + cmul(k2, k2);
        dgp[IDX(i, j, k, nx, ny)] = rmul(vol, dg);
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* prepared operator: coefficient tables of gen_inprod_{x,y,z,p}_coefs       */
struct om_op {
  int nx, ny, nz;
  const double *dx, *dy, *dz, *dxc, *dyc, *dzc; /* borrowed, Fortran index */
  const ocplx *k2x, *k2y, *k2z, *k2p;           /* borrowed */
  /* per component c (0..3) and axis a (0..2): plus/minus neighbour
     coefficient indexed by the 1-based coordinate along that axis */
  double *cp[4][3], *cm[4][3];
};

/* edge-type axis (the component's own direction, length n-1 entries):
   coef_2vec = (i/(n-1)-1)/dc(i+1)/d(i), coef_3vec = ((n-i)/(n-1)-1)/dc(i)/d(i)
   (src/aphid.f90:787-790) */
static void tab_edge(int n, const double *d, const double *dc, double *p, double *m) {
  for (int i = 1; i <= n - 1; ++i) {
    p[i] = (double)(i / (n - 1) - 1) / dc[i + 1] / d[i];
    m[i] = (double)((n - i) / (n - 1) - 1) / dc[i] / d[i];
  }
}
/* node-type axis: coef = (j/n-1)/d(j)/dc(j), ((n-j)/(n-1)-1)/d(j-1)/dc(j)
   (src/aphid.f90:767-779,1161-1164) */
static void tab_node(int n, const double *d, const double *dc, double *p, double *m) {
  for (int j = 1; j <= n; ++j) {
    p[j] = (double)(j / n - 1) / d[j] / dc[j];
    m[j] = (double)((n - j) / (n - 1) - 1) / d[j - 1] / dc[j];
  }
}

om_op *om_op_create(int nx, int ny, int nz, const double *dx, const double *dy,
                    const double *dz, const double *dxc, const double *dyc,
                    const double *dzc, const ocplx *k2x, const ocplx *k2y,
                    const ocplx *k2z, const ocplx *k2p) {
  om_op *op = (om_op *)calloc(1, sizeof(om_op));
  op->nx = nx; op->ny = ny; op->nz = nz;
  op->dx = dx; op->dy = dy; op->dz = dz;
  op->dxc = dxc; op->dyc = dyc; op->dzc = dzc;
  op->k2x = k2x; op->k2y = k2y; op->k2z = k2z; op->k2p = k2p;
  const int n[3] = {nx, ny, nz};
  const double *d[3] = {dx, dy, dz}, *dc[3] = {dxc, dyc, dzc};
  for (int c = 0; c < 4; ++c)
    for (int a = 0; a < 3; ++a) {
      op->cp[c][a] = (double *)calloc((size_t)n[a] + 2, sizeof(double));
      op->cm[c][a] = (double *)calloc((size_t)n[a] + 2, sizeof(double));
      if (c == a)
        tab_edge(n[a], d[a], dc[a], op->cp[c][a], op->cm[c][a]);
      else
        tab_node(n[a], d[a], dc[a], op->cp[c][a], op->cm[c][a]);
    }
  return op;
}

void om_op_destroy(om_op *op) {
  if (!op) return;
  for (int c = 0; c < 4; ++c)
    for (int a = 0; a < 3; ++a) {
      free(op->cp[c][a]);
      free(op->cm[c][a]);
    }
  free(op);
}

static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }

/* inprod_x: src/aphid.f90:803-871 */
static void apply_x(const om_op *op, const ocplx *ix, const ocplx *ip, ocplx *out) {
  const int nx = op->nx, ny = op->ny, nz = op->nz;
  const double *cE = op->cp[0][0], *cW = op->cm[0][0], *cN = op->cp[0][1],
               *cS = op->cm[0][1], *cU = op->cp[0][2], *cD = op->cm[0][2];
  for (int k = 1; k <= nz; ++k) {
    int ka = imin(k + 1, nz), kb = imax(1, k - 1);
    for (int j = 1; j <= ny; ++j) {
      int ja = imin(j + 1, ny), jb = imax(1, j - 1);
      for (int i = 1; i <= nx - 1; ++i) {
        int ia = imin(i + 1, nx - 1), ib = imax(1, i - 1);
        double dxi = op->dx[i];
        double vol = dxi * op->dyc[j] * op->dzc[k];
        ocplx k2 = op->k2x[IDX(i, j, k, nx - 1, ny)];
        ocplx kp1 = op->k2p[IDX(i, j, k, nx, ny)];
        ocplx kp2 = op->k2p[IDX(i + 1, j, k, nx, ny)];
        double sm = cE[i] + cW[i] + cN[j] + cS[j] + cU[k] + cD[k];
        ocplx c1 = CMPLX(-sm + creal(k2), cimag(k2));
        ocplx acc = cmul(c1, ix[IDX(i, j, k, nx - 1, ny)]);
        acc += rmul(cE[i], ix[IDX(ia, j, k, nx - 1, ny)]);
        acc += rmul(cW[i], ix[IDX(ib, j, k, nx - 1, ny)]);
        acc += rmul(cN[j], ix[IDX(i, ja, k, nx - 1, ny)]);
        acc += rmul(cS[j], ix[IDX(i, jb, k, nx - 1, ny)]);
        acc += rmul(cU[k], ix[IDX(i, j, ka, nx - 1, ny)]);
        acc += rmul(cD[k], ix[IDX(i, j, kb, nx - 1, ny)]);
        ocplx cpl = cmul(ip[IDX(i + 1, j, k, nx, ny)], kp2 - k2) -
                    cmul(ip[IDX(i, j, k, nx, ny)], kp1 - k2);
        out[IDX(i, j, k, nx - 1, ny)] = rmul(vol, acc + cdivr(cpl, dxi));
      }
    }
  }
}

/* inprod_y: src/aphid.f90:932-1000 */
static void apply_y(const om_op *op, const ocplx *iy, const ocplx *ip, ocplx *out) {
  const int nx = op->nx, ny = op->ny, nz = op->nz;
  const double *cN = op->cp[1][1], *cS = op->cm[1][1], *cE = op->cp[1][0],
               *cW = op->cm[1][0], *cU = op->cp[1][2], *cD = op->cm[1][2];
  for (int k = 1; k <= nz; ++k) {
    int ka = imin(k + 1, nz), kb = imax(1, k - 1);
    for (int j = 1; j <= ny - 1; ++j) {
      int ja = imin(j + 1, ny - 1), jb = imax(1, j - 1);
      double dyj = op->dy[j];
      for (int i = 1; i <= nx; ++i) {
        int ia = imin(i + 1, nx), ib = imax(1, i - 1);
        double vol = op->dxc[i] * dyj * op->dzc[k];
        ocplx k2 = op->k2y[IDX(i, j, k, nx, ny - 1)];
        ocplx kp1 = op->k2p[IDX(i, j, k, nx, ny)];
        ocplx kp2 = op->k2p[IDX(i, j + 1, k, nx, ny)];
        double sm = cN[j] + cS[j] + cE[i] + cW[i] + cU[k] + cD[k];
        ocplx c1 = CMPLX(-sm + creal(k2), cimag(k2));
        ocplx acc = cmul(c1, iy[IDX(i, j, k, nx, ny - 1)]);
        acc += rmul(cN[j], iy[IDX(i, ja, k, nx, ny - 1)]);
        acc += rmul(cS[j], iy[IDX(i, jb, k, nx, ny - 1)]);
        acc += rmul(cE[i], iy[IDX(ia, j, k, nx, ny - 1)]);
        acc += rmul(cW[i], iy[IDX(ib, j, k, nx, ny - 1)]);
        acc += rmul(cU[k], iy[IDX(i, j, ka, nx, ny - 1)]);
        acc += rmul(cD[k], iy[IDX(i, j, kb, nx, ny - 1)]);
        ocplx cpl = cmul(ip[IDX(i, j + 1, k, nx, ny)], kp2 - k2) -
                    cmul(ip[IDX(i, j, k, nx, ny)], kp1 - k2);
        out[IDX(i, j, k, nx, ny - 1)] = rmul(vol, acc + cdivr(cpl, dyj));
      }
    }
  }
}

/* inprod_z: src/aphid.f90:1061-1128 */
static void apply_z(const om_op *op, const ocplx *iz, const ocplx *ip, ocplx *out) {
  const int nx = op->nx, ny = op->ny, nz = op->nz;
  const double *cU = op->cp[2][2], *cD = op->cm[2][2], *cE = op->cp[2][0],
               *cW = op->cm[2][0], *cN = op->cp[2][1], *cS = op->cm[2][1];
  for (int k = 1; k <= nz - 1; ++k) {
    int ka = imin(k + 1, nz - 1), kb = imax(1, k - 1);
    double dzk = op->dz[k];
    for (int j = 1; j <= ny; ++j) {
      int ja = imin(j + 1, ny), jb = imax(1, j - 1);
      for (int i = 1; i <= nx; ++i) {
        int ia = imin(i + 1, nx), ib = imax(1, i - 1);
        double vol = op->dxc[i] * op->dyc[j] * dzk;
        ocplx k2 = op->k2z[IDX(i, j, k, nx, ny)];
        ocplx kp1 = op->k2p[IDX(i, j, k, nx, ny)];
        ocplx kp2 = op->k2p[IDX(i, j, k + 1, nx, ny)];
        double sm = cU[k] + cD[k] + cE[i] + cW[i] + cN[j] + cS[j];
        ocplx c1 = CMPLX(-sm + creal(k2), cimag(k2));
        ocplx acc = cmul(c1, iz[IDX(i, j, k, nx, ny)]);
        acc += rmul(cU[k], iz[IDX(i, j, ka, nx, ny)]);
        acc += rmul(cD[k], iz[IDX(i, j, kb, nx, ny)]);
        acc += rmul(cE[i], iz[IDX(ia, j, k, nx, ny)]);
        acc += rmul(cW[i], iz[IDX(ib, j, k, nx, ny)]);
        acc += rmul(cN[j], iz[IDX(i, ja, k, nx, ny)]);
        acc += rmul(cS[j], iz[IDX(i, jb, k, nx, ny)]);
        ocplx cpl = cmul(ip[IDX(i, j, k + 1, nx, ny)], kp2 - k2) -
                    cmul(ip[IDX(i, j, k, nx, ny)], kp1 - k2);
        out[IDX(i, j, k, nx, ny)] = rmul(vol, acc + cdivr(cpl, dzk));
      }
    }
  }
}

/* inprod_p: src/aphid.f90:1191-1282 (interior nodes only; others 0) */
static void apply_p(const om_op *op, const ocplx *ix, const ocplx *iy,
                    const ocplx *iz, const ocplx *ip, ocplx *out) {
  const int nx = op->nx, ny = op->ny, nz = op->nz;
  const double *cE = op->cp[3][0], *cW = op->cm[3][0], *cN = op->cp[3][1],
               *cS = op->cm[3][1], *cU = op->cp[3][2], *cD = op->cm[3][2];
  memset(out, 0, sizeof(ocplx) * (size_t)nx * ny * nz);
  for (int k = 2; k <= nz - 1; ++k) {
    int ka = imin(k + 1, nz), kb = imax(1, k - 1), kc = imin(k, nz - 1);
    double dzck = op->dzc[k];
    for (int j = 2; j <= ny - 1; ++j) {
      int ja = imin(j + 1, ny), jb = imax(1, j - 1), jc = imin(j, ny - 1);
      double dycj = op->dyc[j];
      for (int i = 2; i <= nx - 1; ++i) {
        int ia = imin(i + 1, nx), ib = imax(1, i - 1), ic = imin(i, nx - 1);
        double dxci = op->dxc[i];
        double vol = dxci * dycj * dzck;
        ocplx s1 = op->k2p[IDX(i, j, k, nx, ny)];
        ocplx s2 = op->k2x[IDX(ic, j, k, nx - 1, ny)];
        ocplx s3 = op->k2x[IDX(ib, j, k, nx - 1, ny)];
        ocplx s4 = op->k2y[IDX(i, jc, k, nx, ny - 1)];
        ocplx s5 = op->k2y[IDX(i, jb, k, nx, ny - 1)];
        ocplx s6 = op->k2z[IDX(i, j, kc, nx, ny)];
        ocplx s7 = op->k2z[IDX(i, j, kb, nx, ny)];
        ocplx c2 = rmul(cE[i], s2), c3 = rmul(cW[i], s3), c4 = rmul(cN[j], s4);
        ocplx c5 = rmul(cS[j], s5), c6 = rmul(cU[k], s6), c7 = rmul(cD[k], s7);
        ocplx c1 = -(c2 + c3 + c4 + c5 + c6 + c7) + cmul(s1, s1);
        ocplx acc = cmul(c1, ip[IDX(i, j, k, nx, ny)]);
        acc += cmul(c2, ip[IDX(ia, j, k, nx, ny)]);
        acc += cmul(c3, ip[IDX(ib, j, k, nx, ny)]);
        acc += cmul(c4, ip[IDX(i, ja, k, nx, ny)]);
        acc += cmul(c5, ip[IDX(i, jb, k, nx, ny)]);
        acc += cmul(c6, ip[IDX(i, j, ka, nx, ny)]);
        acc += cmul(c7, ip[IDX(i, j, kb, nx, ny)]);
        ocplx o = rmul(vol, acc);
        /* A-coupling terms (:1262-1278) */
        ocplx g2 = cdivr(s2 - s1, dxci), g3 = -cdivr(s3 - s1, dxci);
        ocplx g4 = cdivr(s4 - s1, dycj), g5 = -cdivr(s5 - s1, dycj);
        ocplx g6 = cdivr(s6 - s1, dzck), g7 = -cdivr(s7 - s1, dzck);
        ocplx a2 = cmul(g2, ix[IDX(ic, j, k, nx - 1, ny)]);
        a2 += cmul(g3, ix[IDX(ib, j, k, nx - 1, ny)]);
        a2 += cmul(g4, iy[IDX(i, jc, k, nx, ny - 1)]);
        a2 += cmul(g5, iy[IDX(i, jb, k, nx, ny - 1)]);
        a2 += cmul(g6, iz[IDX(i, j, kc, nx, ny)]);
        a2 += cmul(g7, iz[IDX(i, j, kb, nx, ny)]);
        out[IDX(i, j, k, nx, ny)] = o + rmul(vol, a2);
      }
    }
  }
}

void om_apply_comp(const om_op *op, int comp, const ocplx *ix, const ocplx *iy,
                   const ocplx *iz, const ocplx *ip, ocplx *out) {
  switch (comp) {
    case 0: apply_x(op, ix, ip, out); break;
    case 1: apply_y(op, iy, ip, out); break;
    case 2: apply_z(op, iz, ip, out); break;
    default: apply_p(op, ix, iy, iz, ip, out); break;
  }
}

void om_apply(const om_op *op, const ocplx *ix, const ocplx *iy,
              const ocplx *iz, const ocplx *ip, ocplx *ox, ocplx *oy,
              ocplx *oz, ocplx *opp) {
  apply_x(op, ix, ip, ox);
  apply_y(op, iy, ip, oy);
  apply_z(op, iz, ip, oz);
  apply_p(op, ix, iy, iz, ip, opp);
}

/* ------------------------------------------------------------------------ */
/* right-hand sides: src/aphid.f90:534-743                                   */
int om_rhs(int mode, int nx, int ny, int nz, const double *xn,
           const double *yn, const double *zn, const double *dx,
           const double *dy, const double *dz, int nsrc, const int *dir,
           const int *ip_, const int *jp_, const int *kp_, int *skipped,
           ocplx *bx, ocplx *by, ocplx *bz, ocplx *bp) {
  memset(bx, 0, sizeof(ocplx) * (size_t)(nx - 1) * ny * nz);
  memset(by, 0, sizeof(ocplx) * (size_t)nx * (ny - 1) * nz);
  memset(bz, 0, sizeof(ocplx) * (size_t)nx * ny * (nz - 1));
  memset(bp, 0, sizeof(ocplx) * (size_t)nx * ny * nz);
#define BX(i, j, k) bx[IDX(i, j, k, nx - 1, ny)]
#define BY(i, j, k) by[IDX(i, j, k, nx, ny - 1)]
#define BZ(i, j, k) bz[IDX(i, j, k, nx, ny)]
#define BP(i, j, k) bp[IDX(i, j, k, nx, ny)]
  if (mode == 1) { /* x dipole at the node at the origin (:566-587) */
    const double tol = (double)1.0e-4f; /* real(4) literal 1.e-4 */
    int ip = -1, jp = -1, kp = -1;
    for (int k = 1; k <= nz; ++k) {
      if (fabs(zn[k - 1] - 0.0) < tol) kp = k;
      for (int j = 1; j <= ny; ++j) {
        if (fabs(yn[j - 1] - 0.0) < tol) jp = j;
        for (int i = 1; i <= nx; ++i)
          if (fabs(xn[i - 1] - 0.0) < tol) ip = i;
      }
    }
    if (ip < 1 || jp < 1 || kp < 1 || ip > nx - 1) return -1;
    BX(ip, jp, kp) = 1.0 * OM_MU0;
    BP(ip, jp, kp) = 1.0 * OM_MU0 / dx[ip];
    BP(ip + 1, jp, kp) = -1.0 * OM_MU0 / dx[ip];
  } else if (mode == 2) { /* x-polarised MT sheet at k=2 (:592-604) */
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx - 1; ++i) BX(i, j, 2) = 1.0 * OM_MU0;
    for (int i = 1; i <= nx - 1; ++i)
      for (int j = 1; j <= ny; ++j) BX(i, j, 2) = rmul(dx[i], BX(i, j, 2));
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx - 1; ++i)
        BX(i, j, 2) = rmul(0.5 * (dy[j - 1] + dy[j]), BX(i, j, 2));
  } else if (mode == 3) { /* y-polarised MT sheet at k=2 (:608-620) */
    for (int j = 1; j <= ny - 1; ++j)
      for (int i = 1; i <= nx; ++i) BY(i, j, 2) = 1.0 * OM_MU0;
    for (int j = 1; j <= ny - 1; ++j)
      for (int i = 1; i <= nx; ++i) BY(i, j, 2) = rmul(dy[j], BY(i, j, 2));
    for (int i = 1; i <= nx; ++i)
      for (int j = 1; j <= ny - 1; ++j)
        BY(i, j, 2) = rmul(0.5 * (dx[i - 1] + dx[i]), BY(i, j, 2));
  } else if (mode == 4) { /* fixed 40-cell square loop (:624-631) */
    const int n = 20;
    if (nx < 51 + n + 1 || ny < 51 + n + 1 || nz < 51) return -1;
    for (int i = 51 - n; i <= 51 + n - 1; ++i) {
      BX(i, 51 + n, 51) = 1.0 * OM_MU0 * dx[i];
      BX(i, 51 - n, 51) = -1.0 * OM_MU0 * dx[i];
    }
    for (int j = 51 - n; j <= 51 + n - 1; ++j) {
      BY(51 - n, j, 51) = 1.0 * OM_MU0 * dy[j];
      BY(51 + n, j, 51) = -1.0 * OM_MU0 * dy[j];
    }
  } else if (mode == 5) { /* dipole list (:638-741) */
    for (int s = 0; s < nsrc; ++s) {
      int ip = ip_[s], jp = jp_[s], kp = kp_[s], dr = dir[s], skip = 0;
      switch (dr) {
        case +1:
          if (ip < 1 || ip > nx - 1 || jp < 1 || jp > ny || kp < 1 || kp > nz) { skip = 1; break; }
          BX(ip, jp, kp) += 1.0 * OM_MU0 * dx[ip];
          BP(ip, jp, kp) += 1.0 * OM_MU0;
          BP(ip + 1, jp, kp) -= 1.0 * OM_MU0;
          break;
        case -1:
          if (ip < 2 || ip > nx || jp < 1 || jp > ny || kp < 1 || kp > nz) { skip = 1; break; }
          BX(ip - 1, jp, kp) -= 1.0 * OM_MU0 * dx[ip - 1];
          BP(ip - 1, jp, kp) -= 1.0 * OM_MU0;
          BP(ip, jp, kp) += 1.0 * OM_MU0;
          break;
        case +2:
          if (ip < 1 || ip > nx || jp < 1 || jp > ny - 1 || kp < 1 || kp > nz) { skip = 1; break; }
          BY(ip, jp, kp) += 1.0 * OM_MU0 * dy[jp];
          BP(ip, jp, kp) += 1.0 * OM_MU0;
          BP(ip, jp + 1, kp) -= 1.0 * OM_MU0;
          break;
        case -2:
          if (ip < 1 || ip > nx || jp < 2 || jp > ny || kp < 1 || kp > nz) { skip = 1; break; }
          BY(ip, jp - 1, kp) -= 1.0 * OM_MU0 * dy[jp - 1];
          BP(ip, jp - 1, kp) -= 1.0 * OM_MU0;
          BP(ip, jp, kp) += 1.0 * OM_MU0;
          break;
        case +3:
          if (ip < 1 || ip > nx || jp < 1 || jp > ny || kp < 1 || kp > nz - 1) { skip = 1; break; }
          BZ(ip, jp, kp) += 1.0 * OM_MU0 * dz[kp];
          BP(ip, jp, kp) += 1.0 * OM_MU0;
          BP(ip, jp, kp + 1) -= 1.0 * OM_MU0;
          break;
        case -3:
          if (ip < 1 || ip > nx || jp < 1 || jp > ny || kp < 2 || kp > nz) { skip = 1; break; }
          BZ(ip, jp, kp - 1) -= 1.0 * OM_MU0 * dz[kp - 1];
          BP(ip, jp, kp - 1) -= 1.0 * OM_MU0;
          BP(ip, jp, kp) += 1.0 * OM_MU0;
          break;
        default:
          skip = 1;
      }
      if (skipped) skipped[s] = skip;
    }
  }
#undef BX
#undef BY
#undef BZ
#undef BP
  return 0;
}

/* ------------------------------------------------------------------------ */
/* QMR engine: src/aphid_qmr.f90:5-434 (serial) and :441-972 (4 threads)     */

typedef void (*mv_fn)(const void *ctx, int comp, ocplx *const *in, ocplx *out);

static int qmr_engine(int ncomp, const int64_t *len, mv_fn mv, const void *ctx,
                      ocplx *const *x, ocplx *const *dg, ocplx *const *x0,
                      double tol, int max_it, int nthreads, double *hist,
                      int hist_cap, om_qmr_report *rep) {
  double t1 = now_s();
  ocplx *invd[4], *p[4], *d[4], *vt[4], *v[4], *tmp[4], *r[4], *s[4];
  for (int c = 0; c < ncomp; ++c) {
    size_t nb = sizeof(ocplx) * (size_t)len[c];
    invd[c] = malloc(nb); p[c] = malloc(nb); d[c] = malloc(nb);
    vt[c] = malloc(nb);   v[c] = malloc(nb); tmp[c] = malloc(nb);
    r[c] = malloc(nb);    s[c] = malloc(nb);
  }
  /* ||b|| (:169-172) */
  double rhsnorm = 0.0;
  for (int c = 0; c < ncomp; ++c) {
    double a = 0.0;
    for (int64_t q = 0; q < len[c]; ++q)
      a += creal(x[c][q]) * creal(x[c][q]) + cimag(x[c][q]) * cimag(x[c][q]);
    rhsnorm += a;
  }
  rhsnorm = sqrt(rhsnorm);
  for (int c = 0; c < ncomp; ++c) {
    for (int64_t q = 0; q < len[c]; ++q) invd[c][q] = cdiv(1.0, dg[c][q]); /* :178-181 */
    memset(p[c], 0, sizeof(ocplx) * (size_t)len[c]);
    memset(d[c], 0, sizeof(ocplx) * (size_t)len[c]);
    memcpy(s[c], x[c], sizeof(ocplx) * (size_t)len[c]); /* s := b (:225-228) */
    if (x0 && x0[c]) memcpy(x[c], x0[c], sizeof(ocplx) * (size_t)len[c]);
    else memset(x[c], 0, sizeof(ocplx) * (size_t)len[c]); /* x := 0 (:230-233) */
  }
  /* initial residual (:245-258) */
  for (int c = 0; c < ncomp; ++c) mv(ctx, c, x, tmp[c]);
  double rho = 0.0;
  for (int c = 0; c < ncomp; ++c) {
    double a = 0.0;
    for (int64_t q = 0; q < len[c]; ++q) {
      ocplx e = s[c][q] - tmp[c][q];
      vt[c][q] = e;
      r[c][q] = e;
      a += creal(e) * creal(e) + cimag(e) * cimag(e);
    }
    rho += a;
  }
  rho = sqrt(rho);
  for (int c = 0; c < ncomp; ++c) {
    for (int64_t q = 0; q < len[c]; ++q) v[c][q] = cdivr(vt[c][q], rho);
    memset(s[c], 0, sizeof(ocplx) * (size_t)len[c]);
  }
  double c_n = 1.0, theta_n = 0.0;
  ocplx eps = 1.0, eta = -1.0;
  int iter = 1, status = 1, conv = 0, stop = 0;
  double rnorm = rho;
  const double rho0 = rho;
  ocplx ps[4];
  double pr[4];
  ocplx del = 0, k1 = 0, beta = 0;
  double rho_n = 0, c_nm = 0, theta_nm = 0, k2 = 0;
  if (!(rho > 0.0) || !isfinite(rho)) { stop = 1; status = (rho == 0.0) ? 0 : 2; conv = rho == 0.0; iter = 0; }
  int nt = nthreads > 1 ? nthreads : 1;
  if (nt > ncomp) nt = ncomp;

#pragma omp parallel num_threads(nt) if (nt > 1)
  {
#ifdef _OPENMP
    const int tid = omp_get_thread_num();
    const int ntt = omp_get_num_threads();
#else
    const int tid = 0, ntt = 1;
#endif
    while (!stop) {
      /* step 1: tmp = v*invD, del = v^T tmp (:306-311) */
      for (int c = tid; c < ncomp; c += ntt) {
        ocplx a = 0;
        for (int64_t q = 0; q < len[c]; ++q) {
          ocplx t = cmul(v[c][q], invd[c][q]);
          tmp[c][q] = t;
          a += cmul(v[c][q], t);
        }
        ps[c] = a;
      }
#pragma omp barrier
#pragma omp single
      {
        c_nm = c_n; rho_n = rho; theta_nm = theta_n;
        del = 0;
        for (int c = 0; c < ncomp; ++c) del += ps[c];
        k1 = cdiv(rmul(rho_n, del), eps); /* :313 */
      }
      /* step 2: p = tmp - p*k1 (:317-320) */
      for (int c = tid; c < ncomp; c += ntt)
        for (int64_t q = 0; q < len[c]; ++q) p[c][q] = tmp[c][q] - cmul(p[c][q], k1);
#pragma omp barrier
      /* step 3: tmp = A p, eps = p^T tmp (:324-329) */
      for (int c = tid; c < ncomp; c += ntt) {
        mv(ctx, c, p, tmp[c]);
        ocplx a = 0;
        for (int64_t q = 0; q < len[c]; ++q) a += cmul(p[c][q], tmp[c][q]);
        ps[c] = a;
      }
#pragma omp barrier
#pragma omp single
      {
        eps = 0;
        for (int c = 0; c < ncomp; ++c) eps += ps[c];
        beta = cdiv(eps, del); /* :330 */
      }
      /* vt = tmp - v*beta, rho = ||vt|| (:332-338) */
      for (int c = tid; c < ncomp; c += ntt) {
        double a = 0.0;
        for (int64_t q = 0; q < len[c]; ++q) {
          ocplx e = tmp[c][q] - cmul(v[c][q], beta);
          vt[c][q] = e;
          a += creal(e) * creal(e) + cimag(e) * cimag(e);
        }
        pr[c] = a;
      }
#pragma omp barrier
#pragma omp single
      {
        double a = 0.0;
        for (int c = 0; c < ncomp; ++c) a += pr[c];
        rho = sqrt(a);
        /* step 4 scalars (:342-345) */
        double bb = creal(beta) * creal(beta) + cimag(beta) * cimag(beta);
        theta_n = rho * rho / c_nm / bb;
        c_n = 1.0 / (1.0 + theta_n);
        ocplx num = rmul(c_n, rmul(rho_n, -eta));
        eta = cdiv(num, rmul(c_nm, beta));
        k2 = theta_nm * c_n;
      }
      /* d, x, s, r, v updates (:347-369) and ||r|| (:371-372) */
      for (int c = tid; c < ncomp; c += ntt) {
        double a = 0.0;
        for (int64_t q = 0; q < len[c]; ++q) {
          ocplx dd = cmul(p[c][q], eta) + rmul(k2, d[c][q]);
          d[c][q] = dd;
          x[c][q] = x[c][q] + dd;
          ocplx ss = cmul(tmp[c][q], eta) + rmul(k2, s[c][q]);
          s[c][q] = ss;
          ocplx rr = r[c][q] - ss;
          r[c][q] = rr;
          v[c][q] = cdivr(vt[c][q], rho);
          a += creal(rr) * creal(rr) + cimag(rr) * cimag(rr);
        }
        pr[c] = a;
      }
#pragma omp barrier
#pragma omp single
      {
        double a = 0.0;
        for (int c = 0; c < ncomp; ++c) a += pr[c];
        rnorm = sqrt(a);
        if (hist && iter <= hist_cap) hist[iter - 1] = rnorm;
        int bad = !(isfinite(rnorm) && isfinite(rho)) || del == 0 || eps == 0 || beta == 0;
        /* step 5: stop test and counter (:386-390) */
        if (rnorm < tol * rhsnorm) conv = 1;
        if (bad && !conv) { stop = 1; status = 2; }
        else if (conv) { stop = 1; status = 0; }
        else if (iter <= max_it) { iter = iter + 1; if (iter > max_it) { stop = 1; status = 1; } }
        else { stop = 1; status = 1; }
      }
    }
  }
  double t2 = now_s();
  for (int c = 0; c < ncomp; ++c) {
    free(invd[c]); free(p[c]); free(d[c]); free(vt[c]);
    free(v[c]); free(tmp[c]); free(r[c]); free(s[c]);
  }
  if (rep) {
    rep->iters = iter;
    rep->status = status;
    rep->rnorm = rnorm;
    rep->rhsnorm = rhsnorm;
    rep->rho0 = rho0;
    rep->seconds = t2 - t1;
  }
  return status;
}

static void mv_op(const void *ctx, int comp, ocplx *const *in, ocplx *out) {
  om_apply_comp((const om_op *)ctx, comp, in[0], in[1], in[2], in[3], out);
}

int om_qmr(const om_op *op, ocplx *bx, ocplx *by, ocplx *bz, ocplx *bp,
           const ocplx *dgx, const ocplx *dgy, const ocplx *dgz,
           const ocplx *dgp, const ocplx *x0x, const ocplx *x0y,
           const ocplx *x0z, const ocplx *x0p, double tol, int max_it,
           int nthreads, double *hist, int hist_cap, om_qmr_report *rep) {
  const int nx = op->nx, ny = op->ny, nz = op->nz;
  int64_t len[4] = {(int64_t)(nx - 1) * ny * nz, (int64_t)nx * (ny - 1) * nz,
                    (int64_t)nx * ny * (nz - 1), (int64_t)nx * ny * nz};
  ocplx *x[4] = {bx, by, bz, bp};
  ocplx *dg[4] = {(ocplx *)dgx, (ocplx *)dgy, (ocplx *)dgz, (ocplx *)dgp};
  ocplx *x0[4] = {(ocplx *)x0x, (ocplx *)x0y, (ocplx *)x0z, (ocplx *)x0p};
  int have_x0 = x0x && x0y && x0z && x0p;
  return qmr_engine(4, len, mv_op, op, x, dg, have_x0 ? x0 : NULL, tol, max_it,
                    nthreads, hist, hist_cap, rep);
}

typedef struct {
  int n;
  const ocplx *A;
} dense_ctx;

static void mv_dense(const void *ctx, int comp, ocplx *const *in, ocplx *out) {
  const dense_ctx *dc = (const dense_ctx *)ctx;
  (void)comp;
  for (int i = 0; i < dc->n; ++i) {
    ocplx a = 0;
    for (int j = 0; j < dc->n; ++j) a += cmul(dc->A[(int64_t)i * dc->n + j], in[0][j]);
    out[i] = a;
  }
}

int om_qmr_dense(int n, const ocplx *A, const ocplx *diag, ocplx *b,
                 double tol, int max_it, double *hist, int hist_cap,
                 om_qmr_report *rep) {
  dense_ctx dc = {n, A};
  int64_t len[1] = {n};
  ocplx *x[1] = {b};
  ocplx *dg[1] = {(ocplx *)diag};
  return qmr_engine(1, len, mv_dense, &dc, x, dg, NULL, tol, max_it, 1, hist,
                    hist_cap, rep);
}
