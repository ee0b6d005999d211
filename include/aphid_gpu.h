/*
 * aphid_gpu.h -- C ABI of the MI355X A-Phi solve (libaphid_gpu.so).
 *
 * Drop-in for the reference's hot path: the matrix-free operator
 * (inprod_x/y/z/p, src/aphid.f90:803-1282), its setup (gen_k2_*grid
 * :1289-1535, gen_diag :1542-1725, gen_inprod_*_coefs :750-1185, the
 * sigma-hat models of read_mesh :379-449, gen_rhs :534-743) and the
 * Jacobi-preconditioned complex-symmetric QMR (qmr / qmr_mthread,
 * src/aphid_qmr.f90:5-972).  In the reference these communicate through
 * module globals (src/aphid_mod.f90:5-75); here they take explicit
 * arguments.  A Fortran caller binds them with iso_c_binding
 * (aphid_git_amd/fortran/aphid_gpu_mod.f90).
 *
 * Conventions
 *  - nx, ny, nz are NODE counts (src/aphid_mod.f90:39).
 *  - Host arrays are caller-owned, in Fortran order (i fastest) with the
 *    reference's shapes: x-edges (nx-1,ny,nz), y-edges (nx,ny-1,nz),
 *    z-edges (nx,ny,nz-1), nodes (nx,ny,nz), cells (nx-1,ny-1,nz-1).
 *  - Complex arrays are interleaved (re,im) doubles, i.e. Fortran
 *    complex(c_double_complex) / C double _Complex.
 *  - Return values: 0 success (for a solve: converged), 1 max_it reached
 *    (the iterate is still returned, as the reference does), 2 breakdown or
 *    a non-finite value; negative = error (APHID_E_*), with a message from
 *    aphid_last_error().  No entry point exits or aborts.
 *  - One context per host thread; calls are synchronous; the library owns
 *    all device memory of a context.  The device layout is private (z
 *    fastest); the Fortran<->device permutation happens inside the calls.
 */
#ifndef APHID_GPU_H
#define APHID_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define APHID_OK 0
#define APHID_MAXIT 1
#define APHID_BREAKDOWN 2
#define APHID_E_ARG (-1)
#define APHID_E_HIP (-2)
#define APHID_E_STATE (-3)
#define APHID_E_NOMEM (-4)

/* Largest mesh, in nodes nx*ny*nz (exclusive): every stencil addresses a
 * vector of four complex128 component blocks through 32-bit byte offsets,
 * so 4*nn*16 bytes must stay below 2^32 (e.g. 406^3 nodes fit, 407^3 do
 * not).  aphid_set_mesh and aphid_batch_add_operator return APHID_E_ARG,
 * before any allocation, for a larger mesh. */
#define APHID_MAX_NODES 67108864LL

/* aphid_qmr_solve flags */
#define APHID_F_X0 1u          /* x0 arrays hold an initial guess (else x0 = 0) */
#define APHID_F_NO_TRUE_RES 2u /* skip the final ||b - A x|| check */
#define APHID_F_MIXED 4u       /* complex64 stencil and Krylov vectors, FP64
                                  inner products and recurrence scalars, inside
                                  an FP64 iterative refinement on the true
                                  residual b - K x; iters counts inner
                                  iterations, outer_iters the refinement sweeps,
                                  hist the true relative residual per sweep */

typedef struct aphid_ctx aphid_ctx;

typedef struct {
  int iters;            /* the reference's `iter` (max_it+1 when exhausted) */
  int status;           /* APHID_OK / APHID_MAXIT / APHID_BREAKDOWN */
  double rnorm;         /* final recurrence residual norm ||r|| */
  double rhsnorm;       /* ||b|| */
  double true_rel_res;  /* ||b - A x|| / ||b|| recomputed in complex128 */
  double seconds_qmr;   /* device time of the iteration loop */
  double seconds_total; /* wall time of the call (H2D, setup, solve, D2H) */
  double *hist;         /* optional, caller-owned: ||r|| per iteration */
  int hist_cap;         /* capacity of hist (0 = none) */
  int hist_len;         /* out: entries written */
  int outer_iters;      /* out: refinement sweeps (APHID_F_MIXED), else 1 */
  int reserved;
  double *hist_time;    /* optional, caller-owned, hist_cap entries: wall
                           seconds since the solve started, as of the latest
                           host read of the device state at or before each
                           iteration (the t2-t1 column of the reference's
                           qmr.iter, src/aphid_qmr.f90:376-381) */
} aphid_qmr_report;

/* Context life cycle.  device < 0 selects the current HIP device. */
int aphid_ctx_create(aphid_ctx **ctx, int device);
int aphid_ctx_destroy(aphid_ctx *ctx);
const char *aphid_last_error(void);

/* Mesh from node coordinates: metrics dx(0:n), dxc(1:n) (read_mesh,
 * src/aphid.f90:341-361) and the operator's 1-D coefficient tables
 * (gen_inprod_{x,y,z,p}_coefs, src/aphid.f90:750-797,878-925,1007-1054,
 * 1135-1185).  Requires nx, ny, nz >= 3 and nx*ny*nz < APHID_MAX_NODES. */
int aphid_set_mesh(aphid_ctx *ctx, int nx, int ny, int nz, const double *xn,
                   const double *yn, const double *zn);

/* Cell parameters -> sigma-hat on the device (read_mesh's models,
 * src/aphid.f90:379-449: 'R' a=sigma; 'C' a=sigma,b=eps_rel;
 * 'D' a=sigma,b=eta,c=tau,d=cee; 'I' b=eps_rel; unused pointers may be
 * NULL), then k^2 grids (gen_k2_*grid) and the Jacobi diagonal (gen_diag)
 * for angular frequency omega. */
int aphid_set_cells(aphid_ctx *ctx, char model, double omega, const float *a,
                    const float *b, const float *c, const float *d);

/* Complex cell sigma-hat (Fortran-order (nx-1,ny-1,nz-1)) -> k^2 grids and
 * the Jacobi diagonal on the device (the reference's `sig` module array). */
int aphid_set_sigma(aphid_ctx *ctx, const double *sig_c, double omega);

/* Upload k^2 grids directly (module k2_*grid); rebuilds the diagonal. */
int aphid_set_k2(aphid_ctx *ctx, const double *k2x_c, const double *k2y_c,
                 const double *k2z_c, const double *k2p_c);

/* Override the Jacobi diagonal (module diag_*). */
int aphid_set_diag(aphid_ctx *ctx, const double *dgx_c, const double *dgy_c,
                   const double *dgz_c, const double *dgp_c);

/* Download the device k^2 grids / diagonal (for checks). */
int aphid_get_k2(aphid_ctx *ctx, double *k2x_c, double *k2y_c, double *k2z_c,
                 double *k2p_c);
int aphid_get_diag(aphid_ctx *ctx, double *dgx_c, double *dgy_c, double *dgz_c,
                   double *dgp_c);

/* out = K in over all four row blocks: inprod_x, inprod_y, inprod_z,
 * inprod_p (src/aphid.f90:803-1282).  Inputs and outputs may not alias. */
int aphid_apply(aphid_ctx *ctx, const double *in_x, const double *in_y,
                const double *in_z, const double *in_p, double *out_x,
                double *out_y, double *out_z, double *out_p);

/* aphid_apply without any device allocation.  No counterpart in the
 * reference: this is for callers that cannot allocate device memory (code
 * under stream/graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, semantics and results as aphid_apply; the
 * device copies of in and out and the layout staging live in one
 * library-wide device array of APHID_NOALLOC_SCRATCH_BYTES per GPU instead
 * of the context's vectors, so the mesh must satisfy
 * 9*nx*ny*nz*16 <= APHID_NOALLOC_SCRATCH_BYTES (e.g. up to 77^3 nodes), else
 * APHID_E_ARG.  The operator (mesh, k^2) must already be set on the
 * context; the call uses the context's tile as it stands (it never runs the
 * autotuner).  All calls on one GPU share that array, so two calls that
 * overlap in time on the same GPU (from different host threads) race. */
#define APHID_NOALLOC_SCRATCH_BYTES (64ull << 20)
int aphid_apply_noalloc(aphid_ctx *ctx, const double *in_x, const double *in_y,
                        const double *in_z, const double *in_p, double *out_x,
                        double *out_y, double *out_z, double *out_p);

/* aphid_apply with the stencil in a chosen precision: 0 = complex128
 * (identical to aphid_apply), 1 = complex64 operator data and arithmetic
 * (the mixed-precision inner operator; inputs and outputs stay
 * complex128 on the host). */
int aphid_apply_prec(aphid_ctx *ctx, int precision, const double *in_x,
                     const double *in_y, const double *in_z, const double *in_p,
                     double *out_x, double *out_y, double *out_z, double *out_p);

/* Right-hand side (gen_rhs, src/aphid.f90:534-743) into caller arrays.
 * Mode 5 takes nsrc dipoles with dir in {+1,-1,+2,-2,+3,-3} = +X..-Z and a
 * 1-based tail node; skipped[s] (optional) is set to 1 for a dipole the
 * reference would skip.  Host-side; needs only aphid_set_mesh. */
int aphid_build_rhs(aphid_ctx *ctx, int mode, int nsrc, const int *dir,
                    const int *ip, const int *jp, const int *kp, int *skipped,
                    double *bx_c, double *by_c, double *bz_c, double *bp_c);

/* Jacobi-QMR solve of K x = b (qmr / qmr_mthread, src/aphid_qmr.f90).
 * x_* hold b on entry and the solution on exit (the reference's in-place
 * rhs_* convention).  With APHID_F_X0, x0_* hold the initial guess
 * (src/aphid_qmr.f90:245-258 already forms r0 = b - A x0); the stop test
 * stays ||r|| < tol*||b|| (:386). */
int aphid_qmr_solve(aphid_ctx *ctx, double *x_x, double *x_y, double *x_z,
                    double *x_p, const double *x0_x, const double *x0_y,
                    const double *x0_z, const double *x0_p, double tol,
                    int max_it, unsigned flags, aphid_qmr_report *rep);

/* Device-resident variant of aphid_qmr_solve for callers that solve many
 * right-hand sides or time the solve with its inputs already in HBM:
 * aphid_load_rhs uploads b (and x0 when given; else x0 = 0),
 * aphid_solve_resident runs the same solve on the resident vectors and
 * leaves x on the device, aphid_fetch_solution downloads it.  Together they
 * are aphid_qmr_solve split at its host transfers. */
int aphid_load_rhs(aphid_ctx *ctx, const double *bx_c, const double *by_c,
                   const double *bz_c, const double *bp_c, const double *x0_x,
                   const double *x0_y, const double *x0_z, const double *x0_p);
int aphid_solve_resident(aphid_ctx *ctx, double tol, int max_it, unsigned flags,
                         aphid_qmr_report *rep);
int aphid_fetch_solution(aphid_ctx *ctx, double *x_x, double *x_y, double *x_z,
                         double *x_p);

/* Stencil launch shape in use (picked by the autotuner at the first apply or
 * solve on a mesh, or fixed by aphid_set_tile): tile along z (tk), along y
 * (tj), and x planes per block (chunk >= 1: the x-marching kernel with
 * LDS-staged planes; chunk = 0: the one-node-per-thread kernel).
 * precision: 0 = complex128 kernel, 1 = complex64 kernel. */
int aphid_get_tile(aphid_ctx *ctx, int precision, int *tk, int *tj, int *chunk);
int aphid_set_tile(aphid_ctx *ctx, int precision, int tk, int tj, int chunk);

/* Timing hooks for benchmarks: run `reps` device-resident operator applies
 * (complex128 if precision == 0, complex64 if 1) on the context's current
 * vectors and return the mean kernel time in seconds. */
int aphid_time_apply(aphid_ctx *ctx, int precision, int reps, double *seconds);

/* Reproducible complex128 solves (no counterpart in the reference, whose
 * reduction order is fixed by its OpenMP thread count, src/aphid_qmr.f90:
 * 441-972).  With on != 0 the stencil's eps = p^T K p partials are kept
 * per (node column, 4-plane x-segment), each summed in ascending x by the
 * thread that owns the column, and a fixed grid sums them in a fixed
 * order; the iterates then no longer depend on the stencil's tile or chunk
 * (every other reduction of the solve already runs on fixed grids), so the
 * same inputs give bit-identical solves on any launch shape the tuner may
 * pick, e.g. on GPUs with other CU counts.  The autotuner (and
 * aphid_set_tile) then uses x-march tiles with chunks that are multiples
 * of 4.  Costs a ny*nz*ceil(nx/4) pair buffer, its write in the stencil and
 * one more kernel per iteration.  The mixed-precision path keeps per-block
 * partials.  Default off. */
int aphid_set_reproducible(aphid_ctx *ctx, int on);

/* Compact k^2 for the complex128 x-march stencil (no reference counterpart:
 * a storage choice inside the operator apply, inprod_{x,y,z,p} of
 * src/aphid.f90:803-1282, that leaves every value it computes unchanged).
 * When every cell's sigma-hat has the same imaginary part (model R:
 * sigma + i omega eps0, src/aphid.f90:379-449), the real parts of the edge
 * grids k2x, k2y, k2z depend on two coordinates only, and the real parts of
 * the node grid k2p lie within a few units in the last place of each
 * other; the stencil then reads 33 bytes of k^2 per node instead of 64:
 * the four imaginary parts as two 16-byte pairs, k2p's real part as a
 * one-byte offset of its bit pattern from a common base, and the edge
 * grids' real parts from 2-D tables.  The library checks every real part
 * against this form, bit for bit, each time k^2 is set, and uses it only
 * if all agree (so models C, D and I, or k^2 given by aphid_set_k2 without
 * that structure, keep the whole grids).  on = 0 switches it off (the
 * stencil then reads the whole grids; the results are the same bits).
 * *active (may be NULL) returns whether the compact form is in use.
 * Default on. */
int aphid_set_compact_k2(aphid_ctx *ctx, int on, int *active);

/* Mean seconds per launch of the stencil's memory traffic alone: the tuned
 * complex128 march tile's loads (own columns of the four fields and four
 * k^2 grids, halo cells) and stores (with its store hint) with no
 * arithmetic, LDS staging or barriers, at the same occupancy, repeated on the context's vectors (the
 * spare vector receives meaningless values).  The floor the stencil's
 * access pattern sets under the HBM spec.  No counterpart in the reference;
 * for benchmarks. */
int aphid_time_traffic(aphid_ctx *ctx, int reps, double *seconds);

/* Mean device time (seconds) of each kernel of one QMR iteration, timed on
 * the context's vectors after a solve, each inside the iteration's
 * back-to-back kernel sequence: secs[0] p-update, [1] stencil with the eps
 * epilogue, [2] ||t - beta v||, [3] fused update (their one-block finalize
 * kernels excluded), [4] the whole iteration (finalize kernels included).
 * Overwrites the resident iterate. */
int aphid_time_iteration(aphid_ctx *ctx, int precision, int reps, double *secs);

/* Mean device time of a plain device-to-device copy kernel over `bytes`
 * (read `bytes`, write `bytes`): the achievable-HBM-bandwidth probe that
 * calibrates the roofline and the hardware traffic counters. */
int aphid_time_copy(aphid_ctx *ctx, double bytes, int reps, double *seconds);

/* ---------------------------------------------------------------------------
 * Batched solve: many independent systems (frequencies x sources x meshes of
 * uneven size) in one launch per phase of the QMR iteration.  This replaces
 * the reference's MPI process farm (src/aphid_parallel.f90:75-201, one
 * (frequency, source) system per rank) on a single GPU.  Systems that share a
 * mesh and frequency share one operator.  Each system stops on its own test
 * (||r|| < tol ||b||, src/aphid_qmr.f90:386); converged systems drop out of
 * every later launch.  Host arrays as in the single-system API. */
typedef struct aphid_batch aphid_batch;

typedef struct {
  int iters;           /* the reference's `iter` for this system */
  int status;          /* APHID_OK / APHID_MAXIT / APHID_BREAKDOWN */
  double rnorm;        /* final recurrence residual norm */
  double rhsnorm;      /* ||b|| */
  double true_rel_res; /* ||b - K x|| / ||b|| in complex128 */
  double r0_rel;       /* ||b - K x0|| / ||b|| of the initial guess: 1 from
                          x0 = 0, the fit's value for a sweep guess, NaN
                          for a caller's x0 (not formed) or a mixed solve */
} aphid_sys_report;

int aphid_batch_create(aphid_batch **batch, int device);
int aphid_batch_destroy(aphid_batch *batch);
/* An operator: mesh + cell parameters at angular frequency omega (as
 * aphid_set_mesh + aphid_set_cells).  Returns its index (>= 0). */
int aphid_batch_add_operator(aphid_batch *batch, int nx, int ny, int nz,
                             const double *xn, const double *yn,
                             const double *zn, char model, double omega,
                             const float *a, const float *b, const float *c,
                             const float *d);
/* A system on operator `op`: right-hand side and an optional initial guess
 * (NULL = 0).  Add every operator before the first system and every system
 * before the first solve.  Returns the system index (>= 0). */
int aphid_batch_add_system(aphid_batch *batch, int op, const double *bx_c,
                           const double *by_c, const double *bz_c,
                           const double *bp_c, const double *x0_x,
                           const double *x0_y, const double *x0_z,
                           const double *x0_p);
/* Solve every system.  reports (optional) has one entry per system; returns
 * the worst status.  seconds (optional): wall time of the solve, from the
 * right-hand sides in place to the last system done (the final true
 * residuals excluded).  With
 * APHID_F_X0 the current solutions are the initial guesses (warm restart,
 * e.g. after aphid_batch_set_guess). */
int aphid_batch_solve(aphid_batch *batch, double tol, int max_it,
                      unsigned flags, aphid_sys_report *reports,
                      double *seconds);
int aphid_batch_get_solution(aphid_batch *batch, int sys, double *x_x,
                             double *x_y, double *x_z, double *x_p);
/* Use system from_sys's current solution as the initial guess of system
 * sys (same mesh), e.g. the previous frequency of a sweep. */
int aphid_batch_set_guess(aphid_batch *batch, int sys, int from_sys);
/* Sweep schedule (no counterpart in the reference, whose farm solves every
 * system from x0 = 0, src/aphid_parallel.f90:201 -> src/aphid_qmr.f90:
 * 230-233): mark system sys as fitted from the nfrom
 * (<= APHID_SWEEP_MAX_FROM) systems `from` -- same mesh, e.g. the
 * neighbouring frequencies with the same source.  aphid_batch_solve solves
 * in waves: first every unmarked system, then each marked one in the wave
 * after its latest source's, with its initial guess fitted to its sources'
 * solutions by the Galerkin condition x_a^T (b - K x0) = 0 on its own
 * operator (least squares on the residual if that is worse than x0 = 0,
 * and x0 = 0 if both are); an x0 given for a marked system is replaced.
 * Sources may be marked themselves (a sweep of several levels), but a mark
 * that would make a system depend on itself is refused.  nfrom = 0
 * unmarks.  A mixed-precision batch solve (APHID_F_MIXED) ignores the
 * marks. */
#define APHID_SWEEP_MAX_FROM 4
int aphid_batch_set_sweep_guess(aphid_batch *batch, int sys, int nfrom,
                                const int *from);
/* Residual histories of the batched solve: every later complex128
 * aphid_batch_solve keeps, per system, ||r|| of its first `cap` iterations
 * and the wall-clock time at which each was read, the figures each process
 * of the reference's farm writes into its own qmr.iter
 * (src/aphid_qmr.f90:376-381 under src/aphid_parallel.f90:201).  cap = 0
 * (the default) keeps none.  A mixed-precision solve (APHID_F_MIXED) keeps
 * none (its inner iterations run in complex64 on a changing right-hand
 * side). */
int aphid_batch_set_history(aphid_batch *batch, int cap);
/* History of system sys from the last solve: *len (<= cap) entries of
 * ||r|| after iteration 1, 2, ... into hist and, when time is not NULL,
 * the seconds since that system's QMR began (its wave's batched launches,
 * or its own single-system solve) at the latest host read of its state at
 * or before that iteration (0 before the first read). */
int aphid_batch_get_history(aphid_batch *batch, int sys, double *hist,
                            double *time, int cap, int *len);

/* Post-processing of the resident solution (gen_fields, src/aphid.f90:
 * 2173-2252): B = curl A on the primary cell faces -- Bx (nx,ny-1,nz-1),
 * By (nx-1,ny,nz-1), Bz (nx-1,ny-1,nz) -- and E = -i omega (A - grad Phi)
 * on the primary edges -- Ex (nx-1,ny,nz), Ey (nx,ny-1,nz), Ez (nx,ny,nz-1)
 * -- complex, Fortran order. */
int aphid_gen_fields(aphid_ctx *ctx, double omega, double *bx, double *by,
                     double *bz, double *ex, double *ey, double *ez);
/* The complex sigma-hat of every cell (nx-1,ny-1,nz-1) as the operator
 * uses it (read_mesh's models, src/aphid.f90:379-449), for dump_sigma. */
int aphid_get_sigma(aphid_ctx *ctx, double *sig);

/* ---------------------------------------------------------------------------
 * Learned warm start.  No counterpart in the reference, which always starts
 * QMR from x0 = 0 (src/aphid_qmr.f90:230-233).
 *
 * A 3-level 3-D U-Net (16/32/64 channels) on a cell grid of nx x ny x nz
 * cells (multiples of 4), bf16 weights and activations with FP32
 * accumulation on the matrix cores, trained on the GPU.  Features per cell
 * (z fastest, FP32, 3 per cell): normalised log10 sigma, skin-depth
 * normalised log frequency, source mask.  Output per cell: 8 FP32 (re, im
 * of Ax, Ay, Az, Phi) as a normalised field shape.  Parameters are one flat
 * FP32 vector: per layer the weights [cout][cin][k][k][k] then the biases.
 */
typedef struct aphid_unet aphid_unet;

aphid_unet *aphid_unet_create(int device, int nx, int ny, int nz,
                              unsigned long long seed);
void aphid_unet_destroy(aphid_unet *net);
long long aphid_unet_param_count(const aphid_unet *net);
/* multiply-add FLOPs (x2) of one forward pass */
double aphid_unet_forward_flops(const aphid_unet *net);
int aphid_unet_get_params(aphid_unet *net, float *params);
int aphid_unet_set_params(aphid_unet *net, const float *params);
/* out (optional): [nx*ny*nz][8] */
int aphid_unet_forward(aphid_unet *net, const float *features, float *out);
/* One forward and backward pass of the mean squared error against
 * target [nx*ny*nz][8]; grad (optional) receives d loss / d params. */
int aphid_unet_loss_grad(aphid_unet *net, const float *features,
                         const float *target, double *loss, float *grad);
/* A stored tensor of the last pass (index into the network's buffer list;
 * grad != 0: its gradient), as FP32. */
int aphid_unet_activation(aphid_unet *net, int which, int grad, float *out);
/* Training set held in device memory (bf16). */
int aphid_unet_reserve(aphid_unet *net, int capacity);
int aphid_unet_add_sample(aphid_unet *net, const float *features,
                          const float *target);
/* `steps` Adam steps over random mini-batches of `batch` samples;
 * losses (optional) gets the mean loss of each step. */
int aphid_unet_train(aphid_unet *net, int steps, int batch, float lr,
                     unsigned long long seed, float *losses, double *seconds);
/* Device time of one forward pass and of one forward+backward pass. */
int aphid_unet_time(aphid_unet *net, int reps, double *fwd_seconds,
                    double *step_seconds);

#define APHID_WS_PREV 1u /* include the context's current solution */
#define APHID_WS_LSQ 2u  /* fit amplitudes by least squares on the residual */
#define APHID_WS_JLSQ 4u /* network only: amplitudes by least squares on the
                            Jacobi-weighted residual ||W (b - K x0)||,
                            W = 1/diag, the fit aphid_warm_train_residual
                            trains with; x0 = 0 if it leaves
                            ||b - K x0|| > ||b|| */
/* Initial guess for aphid_solve_resident(..., APHID_F_X0) from the network
 * (net may be NULL) and/or the context's current solution.  One complex
 * amplitude per candidate (each network component, the previous solution)
 * from the Galerkin condition u^T (b - K x0) = 0, kept only if
 * ||b - K x0|| <= ||b||; otherwise (or with APHID_WS_LSQ) minimising
 * ||b - K x0||, which is never worse than x0 = 0.  Each cell
 * (i,j,k) of the network owns the unknowns at its lower corner node; the
 * mesh must have (nx+1, ny+1, nz+1) nodes for nx x ny x nz cells.
 * r0_rel (optional): ||b - K x0|| / ||b||. */
int aphid_warm_start(aphid_ctx *ctx, aphid_unet *net, const float *features,
                     unsigned flags, double *r0_rel);

/* The residual loss below for one case -- the context's operator and
 * resident right-hand side, the network's features -- and its gradient
 * with respect to the parameters (grad, optional, FP32 [param count]); the
 * network's output gradient stays readable with aphid_unet_activation. */
int aphid_warm_residual_loss_grad(aphid_ctx *ctx, aphid_unet *net, const float *features,
                                  unsigned flags, double *loss, float *grad);

/* Train the network on the residual of its own guess (no counterpart in
 * the reference).  Each step draws `batch` of the network's stored samples
 * (aphid_unet_add_sample, in the order of sigma and omega), rebuilds that
 * sample's operator on the device from its cells (model R: sigma
 * [nsamples][(nx-1)(ny-1)(nz-1)], Fortran order per sample; angular
 * frequency omega[sample]), fits the four component amplitudes of the
 * mapped output by least squares on the residual, and back-propagates
 *   L = ||W (b - K x0)||^2 / ||W b||^2   (b the context's resident
 * right-hand side, shared by all samples; flags APHID_WS_JLSQ: W = 1/diag,
 * the fit of aphid_warm_start's APHID_WS_JLSQ; else W = 1, the fit of its
 * APHID_WS_LSQ for a network alone), whose gradient needs one
 * more operator apply (K is complex symmetric: K^H y = conj(K conj y)); then
 * one Adam step.  losses (optional): mean L per step.  Overwrites the
 * context's operator and resident vectors (reload b before solving). */
int aphid_warm_train_residual(aphid_ctx *ctx, aphid_unet *net, const float *sigma,
                              const double *omega, unsigned flags, int steps, int batch,
                              float lr, unsigned long long seed, float *losses,
                              double *seconds);

#ifdef __cplusplus
}
#endif
#endif
