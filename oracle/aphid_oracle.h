/*
 * aphid_oracle.h -- CPU oracle for the A-Phi Jacobi-QMR solve.
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE.  This is a plain C restatement of
 * the reference's algorithm (src/aphid.f90, src/aphid_qmr.f90) used only as
 * the checker in tests/, by __graft_entry__.smoke(), and as the "port"
 * flavour of bench.py's CPU baseline.  The product path (the HIP library)
 * never links, loads or calls it.
 *
 * Conventions (all functions):
 *   - nx, ny, nz are NODE counts, exactly as in the reference's `mesh`
 *     module (src/aphid_mod.f90:39).
 *   - 3-D arrays are in Fortran order (i fastest), with the reference's
 *     shapes: x-edges (nx-1,ny,nz), y-edges (nx,ny-1,nz), z-edges
 *     (nx,ny,nz-1), nodes (nx,ny,nz), cells (nx-1,ny-1,nz-1).
 *   - complex values are C99 `double _Complex` (= Fortran complex(8)).
 *   - metric arrays keep the reference's Fortran index: d[0..n] holds
 *     dx(0:nx) and dc[1..n] holds dxc(1:nx) (dc[0] unused), so both have
 *     n+1 entries.
 */
#ifndef APHID_ORACLE_H
#define APHID_ORACLE_H

#include <complex.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef double _Complex ocplx;

/* Reference constants, src/aphid_mod.f90:9-12. */
#define OM_PI   3.141592653589793
#define OM_MU0  (OM_PI * 4.0e-7)
#define OM_EPS0 8.8541878176e-12

/* dx(0:n), dxc(1:n) from node coordinates c[0..n-1]; src/aphid.f90:341-361. */
void om_metrics(int n, const double *c, double *d, double *dc);

/* Cell sigma-hat from FP32 inputs; src/aphid.f90:379-449.
 * model 'R': a=sigma            -> a + i w eps0
 * model 'C': a=sigma, b=eps_rel -> a + i w eps0 eps_rel
 * model 'D': a=sigma, b=eta, c=tau, d=cee -> a (1 - b / (1 + (i w c)^d))
 * model 'I': b=eps_rel          -> i w eps0 eps_rel
 * Returns 0, or -1 for an unknown model letter. */
int om_sigma(char model, double omega, int64_t ncell, const float *a,
             const float *b, const float *c, const float *d, ocplx *sig);

/* k^2 grids (edge 4-cell and node 8-cell averages, copy-then-scale);
 * src/aphid.f90:1289-1535. */
void om_k2(int nx, int ny, int nz, const double *dx, const double *dy,
           const double *dz, const ocplx *sig, double omega, ocplx *k2x,
           ocplx *k2y, ocplx *k2z, ocplx *k2p);

/* Jacobi diagonal with the reference's boundary quirks;
 * src/aphid.f90:1542-1725. */
void om_diag(int nx, int ny, int nz, const double *dx, const double *dy,
             const double *dz, const ocplx *k2x, const ocplx *k2y,
             const ocplx *k2z, const ocplx *k2p, ocplx *dgx, ocplx *dgy,
             ocplx *dgz, ocplx *dgp);

/* Prepared operator: metrics + coefficient tables (gen_inprod_*_coefs,
 * src/aphid.f90:750-797,878-925,1007-1054,1135-1185) + borrowed k^2. */
typedef struct om_op om_op;
om_op *om_op_create(int nx, int ny, int nz, const double *dx, const double *dy,
                    const double *dz, const double *dxc, const double *dyc,
                    const double *dzc, const ocplx *k2x, const ocplx *k2y,
                    const ocplx *k2z, const ocplx *k2p);
void om_op_destroy(om_op *op);

/* out = K in, all four row blocks (inprod_x/y/z/p, src/aphid.f90:803-1282). */
void om_apply(const om_op *op, const ocplx *ix, const ocplx *iy,
              const ocplx *iz, const ocplx *ip, ocplx *ox, ocplx *oy,
              ocplx *oz, ocplx *opp);

/* One row block at a time (component 0..3 = x,y,z,p). */
void om_apply_comp(const om_op *op, int comp, const ocplx *ix, const ocplx *iy,
                   const ocplx *iz, const ocplx *ip, ocplx *out);

/* Right-hand side, src/aphid.f90:534-743.  Modes 1-4 as the reference; mode 5
 * takes nsrc dipoles: dir[s] in {+1,-1,+2,-2,+3,-3} (= +X,-X,+Y,-Y,+Z,-Z)
 * with 1-based tail node (ip,jp,kp).  skipped[s] (optional) is set to 1 for
 * a dipole the reference skips as out of bounds.  Outputs are zeroed first.
 * Returns 0, or -1 when mode 1 finds no node at the origin. */
int om_rhs(int mode, int nx, int ny, int nz, const double *xn,
           const double *yn, const double *zn, const double *dx,
           const double *dy, const double *dz, int nsrc, const int *dir,
           const int *ip, const int *jp, const int *kp, int *skipped,
           ocplx *bx, ocplx *by, ocplx *bz, ocplx *bp);

/* Jacobi-preconditioned complex-symmetric QMR, src/aphid_qmr.f90:5-434
 * (nthreads=1) or the 4-thread component decomposition of :441-972 with a
 * correct reduction (nthreads=4).  x[c] holds b on entry and the solution on
 * exit (the reference's in-place convention).  x0 (optional, may be NULL)
 * is an initial guess; NULL means x0 = 0 as in the reference.
 * hist (optional) receives rnorm for each iteration (the qmr.iter column).
 * Returns 0 converged, 1 max_it exhausted, 2 breakdown/non-finite. *iters
 * follows the reference's `iter` (max_it+1 on exhaustion). */
typedef struct {
  int iters;
  int status;
  double rnorm;
  double rhsnorm;
  double rho0;
  double seconds;
} om_qmr_report;

int om_qmr(const om_op *op, ocplx *bx, ocplx *by, ocplx *bz, ocplx *bp,
           const ocplx *dgx, const ocplx *dgy, const ocplx *dgz,
           const ocplx *dgp, const ocplx *x0x, const ocplx *x0y,
           const ocplx *x0z, const ocplx *x0p, double tol, int max_it,
           int nthreads, double *hist, int hist_cap, om_qmr_report *rep);

/* Generic dense-matrix QMR with the same recurrence (for the Krylov
 * property tests): A is n x n complex, row-major. */
int om_qmr_dense(int n, const ocplx *A, const ocplx *diag, ocplx *b,
                 double tol, int max_it, double *hist, int hist_cap,
                 om_qmr_report *rep);

#ifdef __cplusplus
}
#endif
#endif
