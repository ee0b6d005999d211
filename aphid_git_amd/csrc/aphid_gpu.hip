// aphid_gpu.hip -- host side of libaphid_gpu.so: the C ABI declared in
// include/aphid_gpu.h, the per-context device state, the stencil
// autotuner and the graph-replayed QMR loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "aphid_gpu.h"
#include "qmr.hpp"
#include "setup.hpp"
#include "stencil.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define CK(call)                                                                     \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
      return fail(APHID_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));   \
  } while (0)

constexpr int kBlock = 256;
constexpr int kRedBlocks = 2048;        // grid of the reduction kernels
constexpr int kMaxPartBlocks = 1 << 16; // capacity of the partial-sum buffer

// ---------------------------------------------------------------------------
// stencil launch shapes (compile-time tiles, runtime chunk)
struct Tile {
  int tk, tj, chunk;  // chunk = x-planes per block (march); 0 = direct variant
};
const Tile kTiles[] = {{16, 16, 0}, {32, 8, 0}, {64, 4, 0}, {16, 8, 0}, {32, 4, 0}};

template <typename R>
hipError_t launch_apply(const Tile& t, const OpArgs<R>& op, const cplx<R>* in, cplx<R>* out,
                        ApplyEpilogue ep, hipStream_t s) {
  const int ntk = (op.nz + t.tk - 1) / t.tk, ntj = (op.ny + t.tj - 1) / t.tj;
  dim3 block(t.tk * t.tj);
  if (t.chunk <= 0) {
    dim3 grid((unsigned)ntk * ntj * op.nx);
#define APHID_TILE(TK, TJ)                                                                \
    if (t.tk == TK && t.tj == TJ) {                                                       \
      hipLaunchKernelGGL((k_apply_direct<R, TK, TJ>), grid, block, 0, s, op, in, out, ep); \
      return hipGetLastError();                                                           \
    }
    APHID_TILE(16, 16)
    APHID_TILE(32, 8)
    APHID_TILE(64, 4)
    APHID_TILE(16, 8)
    APHID_TILE(32, 4)
#undef APHID_TILE
    return hipErrorInvalidValue;
  }
  const int chunk = t.chunk;
  dim3 grid((unsigned)ntk * ntj * ((op.nx + chunk - 1) / chunk));
#define APHID_TILE(TK, TJ)                                                           \
  if (t.tk == TK && t.tj == TJ) {                                                    \
    hipLaunchKernelGGL((k_apply<R, TK, TJ>), grid, block, 0, s, op, in, out, chunk, ep); \
    return hipGetLastError();                                                        \
  }
  APHID_TILE(16, 16)
  APHID_TILE(32, 8)
  APHID_TILE(64, 4)
  APHID_TILE(16, 8)
  APHID_TILE(32, 4)
#undef APHID_TILE
  return hipErrorInvalidValue;
}

unsigned apply_blocks(const Tile& t, int nx, int ny, int nz) {
  const unsigned ntk = (nz + t.tk - 1) / t.tk, ntj = (ny + t.tj - 1) / t.tj;
  if (t.chunk <= 0) return ntk * ntj * (unsigned)nx;
  return ntk * ntj * (unsigned)((nx + t.chunk - 1) / t.chunk);
}

// 1-D tables of one axis (see OpArgs in stencil.hpp), from node coordinates
struct AxisTables {
  std::vector<double> xn;     // node coordinates
  std::vector<double> d, dc;  // Fortran-indexed d(0:n), dc(1:n) (dc[0] unused)
  std::vector<double> eP, eM, dE, rdE, nP, nM, dN, rdN;
};

AxisTables make_axis(const double* c, int n) {
  AxisTables a;
  a.xn.assign(c, c + n);
  a.d.assign(n + 1, 0.0);
  a.dc.assign(n + 1, 0.0);
  for (int i = 1; i <= n - 1; ++i) a.d[i] = c[i] - c[i - 1];  // read_mesh :351-357
  a.d[0] = a.d[1];
  a.d[n] = a.d[n - 1];
  for (int i = 2; i <= n - 1; ++i) a.dc[i] = 0.5 * (a.d[i - 1] + a.d[i]);  // :359-361
  a.dc[1] = a.dc[2];
  a.dc[n] = a.dc[n - 1];
  for (int e = 0; e < n - 1; ++e) {  // edge-type (own axis), gen_inprod_*_coefs
    const int E = e + 1;
    a.eP.push_back((double)(E / (n - 1) - 1) / a.dc[E + 1] / a.d[E]);
    a.eM.push_back((double)((n - E) / (n - 1) - 1) / a.dc[E] / a.d[E]);
    a.dE.push_back(a.d[E]);
    a.rdE.push_back(1.0 / a.d[E]);
  }
  a.eP.push_back(0.0);  // pad to length n (the padding entry is never a row)
  a.eM.push_back(0.0);
  a.dE.push_back(1.0);
  a.rdE.push_back(1.0);
  for (int q = 0; q < n; ++q) {  // node-type (transverse axes and Phi rows)
    const int Q = q + 1;
    a.nP.push_back((double)(Q / n - 1) / a.d[Q] / a.dc[Q]);
    a.nM.push_back((double)((n - Q) / (n - 1) - 1) / a.d[Q - 1] / a.dc[Q]);
    a.dN.push_back(a.dc[Q]);
    a.rdN.push_back(1.0 / a.dc[Q]);
  }
  return a;
}

}  // namespace

// ---------------------------------------------------------------------------
struct aphid_ctx {
  int dev = 0;
  hipStream_t s = nullptr;
  int nx = 0, ny = 0, nz = 0;
  long long nn = 0;
  AxisTables ax[3];
  // device tables: 8 arrays per axis, each padded to length n, double + float
  double* tab64 = nullptr;
  float* tab32 = nullptr;
  double* dmet = nullptr;  // d(0:n) per axis for the setup kernels
  OpArgs<double> op64{};
  OpArgs<float> op32{};
  // operator data
  c128* sig = nullptr;
  c128* k2 = nullptr;
  c64* k2f = nullptr;
  c128* diag = nullptr;
  c128* invd = nullptr;
  c64* invdf = nullptr;
  bool have_k2 = false, have_diag = false;
  // Krylov vectors, complex128 and complex64 sets
  c128* V[9] = {};  // x b v p t d s r spare
  c64* F[8] = {};   // x b v p t d s r (complex64)
  c128* stage = nullptr;
  float* stagef = nullptr;  // staging for FP32 cell inputs (4 blocks)
  // QMR state
  QmrState* st = nullptr;
  QmrState* hst = nullptr;  // pinned mirror
  double* part = nullptr;
  double* hist = nullptr;
  int hist_cap = 0;
  double* resbuf = nullptr;  // [||b-Ax||, ||b||]
  // graph of `giters` iterations per precision
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  int giters[2] = {0, 0};
  // stencil tiles per precision
  Tile tile[2] = {{16, 16, 8}, {16, 16, 8}};
  bool tuned[2] = {false, false};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace {

void free_mesh(aphid_ctx* c) {
  for (int p = 0; p < 2; ++p) {
    if (c->gexec[p]) hipGraphExecDestroy(c->gexec[p]);
    c->gexec[p] = nullptr;
    c->giters[p] = 0;
  }
  void* ptrs[] = {c->tab64, c->tab32, c->dmet, c->sig, c->k2, c->k2f, c->diag, c->invd, c->invdf,
                  c->stage, c->stagef, c->hist};
  for (void* p : ptrs)
    if (p) hipFree(p);
  for (auto& v : c->V)
    if (v) hipFree(v), v = nullptr;
  for (auto& v : c->F)
    if (v) hipFree(v), v = nullptr;
  c->tab64 = nullptr; c->tab32 = nullptr; c->dmet = nullptr; c->sig = nullptr; c->k2 = nullptr;
  c->k2f = nullptr; c->diag = nullptr; c->invd = nullptr; c->invdf = nullptr; c->stage = nullptr;
  c->stagef = nullptr; c->hist = nullptr; c->hist_cap = 0;
  c->have_k2 = c->have_diag = false;
  c->tuned[0] = c->tuned[1] = false;
}

unsigned grid_for(long long n) {
  long long g = (n + kBlock - 1) / kBlock;
  return (unsigned)std::max(1LL, std::min<long long>(g, 8192));
}

Dims dims(const aphid_ctx* c) { return Dims{c->nx, c->ny, c->nz, c->nn}; }

int comp_shape(const aphid_ctx* c, int comp, int& n1, int& n2, int& n3) {
  n1 = c->nx - (comp == 0);
  n2 = c->ny - (comp == 1);
  n3 = c->nz - (comp == 2);
  return 0;
}

// host Fortran-order component arrays -> device vector (4 blocks of nn)
int upload4(aphid_ctx* c, const double* const h[4], c128* dst) {
  for (int comp = 0; comp < 4; ++comp) {
    int n1, n2, n3;
    comp_shape(c, comp, n1, n2, n3);
    const size_t bytes = sizeof(c128) * (size_t)n1 * n2 * n3;
    CK(hipMemcpyAsync(c->stage, h[comp], bytes, hipMemcpyHostToDevice, c->s));
    hipLaunchKernelGGL(k_f2d<double>, dim3(grid_for(c->nn)), dim3(kBlock), 0, c->s, c->stage, n1,
                       n2, n3, dst + comp * c->nn, dims(c));
    CK(hipGetLastError());
  }
  return 0;
}

int download4(aphid_ctx* c, const c128* src, double* const h[4]) {
  for (int comp = 0; comp < 4; ++comp) {
    int n1, n2, n3;
    comp_shape(c, comp, n1, n2, n3);
    const long long n = (long long)n1 * n2 * n3;
    hipLaunchKernelGGL(k_d2f<double>, dim3(grid_for(n)), dim3(kBlock), 0, c->s,
                       src + comp * c->nn, n1, n2, n3, c->stage, dims(c));
    CK(hipGetLastError());
    CK(hipMemcpyAsync(h[comp], c->stage, sizeof(c128) * (size_t)n, hipMemcpyDeviceToHost, c->s));
  }
  CK(hipStreamSynchronize(c->s));
  return 0;
}

int ensure_vectors(aphid_ctx* c, bool f32) {
  const size_t n4 = 4 * (size_t)c->nn;
  for (auto& v : c->V)
    if (!v) CK(hipMalloc(&v, sizeof(c128) * n4));
  if (f32)
    for (auto& v : c->F)
      if (!v) CK(hipMalloc(&v, sizeof(c64) * n4));
  return 0;
}

// derived complex64 operator data (k2, 1/diag)
int refresh_f32(aphid_ctx* c) {
  const long long n4 = 4 * c->nn;
  hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_for(n4)), dim3(kBlock), 0, c->s, c->k2,
                     c->k2f, n4);
  hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_for(n4)), dim3(kBlock), 0, c->s,
                     c->invd, c->invdf, n4);
  CK(hipGetLastError());
  return 0;
}

int build_diag(aphid_ctx* c) {
  const double* Dx = c->dmet;
  const double* Dy = Dx + (c->nx + 1);
  const double* Dz = Dy + (c->ny + 1);
  hipLaunchKernelGGL(k_diag, dim3(grid_for(c->nn)), dim3(kBlock), 0, c->s, c->k2, Dx, Dy, Dz,
                     c->diag, c->invd, dims(c));
  CK(hipGetLastError());
  c->have_diag = true;
  return refresh_f32(c);
}

int build_k2(aphid_ctx* c, double omega) {
  const double* Dx = c->dmet;
  const double* Dy = Dx + (c->nx + 1);
  const double* Dz = Dy + (c->ny + 1);
  hipLaunchKernelGGL(k_k2, dim3(grid_for(c->nn)), dim3(kBlock), 0, c->s, c->sig, Dx, Dy, Dz, omega,
                     c->k2, dims(c));
  CK(hipGetLastError());
  c->have_k2 = true;
  return build_diag(c);
}

// time one stencil configuration (mean of reps launches)
template <typename R>
int time_tile(aphid_ctx* c, const Tile& t, const OpArgs<R>& op, const cplx<R>* in, cplx<R>* out,
              int reps, float* ms) {
  ApplyEpilogue ep{nullptr, nullptr};
  CK(launch_apply<R>(t, op, in, out, ep, c->s));  // warm
  CK(hipEventRecord(c->ev0, c->s));
  for (int r = 0; r < reps; ++r) CK(launch_apply<R>(t, op, in, out, ep, c->s));
  CK(hipEventRecord(c->ev1, c->s));
  CK(hipEventSynchronize(c->ev1));
  CK(hipEventElapsedTime(ms, c->ev0, c->ev1));
  *ms /= reps;
  return 0;
}

// Autotuner: time every (tile, chunk) candidate for this mesh and keep the
// fastest.  Results are cached per mesh shape and precision in the file
// named by APHID_TUNE_CACHE (if set).
int autotune(aphid_ctx* c, int prec) {
  if (c->tuned[prec]) return 0;
  char key[96];
  snprintf(key, sizeof key, "%d %d %d %d", c->nx, c->ny, c->nz, prec);
  const char* cache = getenv("APHID_TUNE_CACHE");
  if (cache) {
    if (FILE* f = fopen(cache, "r")) {
      char line[160];
      while (fgets(line, sizeof line, f)) {
        int a, b, cc, p, tk, tj, ch;
        if (sscanf(line, "%d %d %d %d %d %d %d", &a, &b, &cc, &p, &tk, &tj, &ch) == 7) {
          char k2[96];
          snprintf(k2, sizeof k2, "%d %d %d %d", a, b, cc, p);
          if (!strcmp(k2, key)) {
            c->tile[prec] = {tk, tj, ch};
            c->tuned[prec] = true;
          }
        }
      }
      fclose(f);
      if (c->tuned[prec]) return 0;
    }
  }
  int rc = ensure_vectors(c, prec == 1);
  if (rc) return rc;
  const int chunks[] = {0, 2, 4, 6, 8, 12, 16};  // 0 = direct variant
  float best = 1e30f;
  Tile bt = c->tile[prec];
  const int reps = c->nn > 2000000 ? 5 : 20;
  for (const Tile& base : kTiles) {
    for (int ch : chunks) {
      Tile t = {base.tk, base.tj, ch};
      if (apply_blocks(t, c->nx, c->ny, c->nz) > (unsigned)kMaxPartBlocks) continue;
      float ms = 0.f;
      if (prec == 0)
        rc = time_tile<double>(c, t, c->op64, c->V[3], c->V[4], reps, &ms);
      else
        rc = time_tile<float>(c, t, c->op32, c->F[3], c->F[4], reps, &ms);
      if (rc) return rc;
      if (ms < best) {
        best = ms;
        bt = t;
      }
    }
  }
  c->tile[prec] = bt;
  c->tuned[prec] = true;
  if (cache)
    if (FILE* f = fopen(cache, "a")) {
      fprintf(f, "%s %d %d %d\n", key, bt.tk, bt.tj, bt.chunk);
      fclose(f);
    }
  return 0;
}

// ---------------------------------------------------------------------------
// QMR on the device for precision V (double: plain complex128 path, float:
// complex64 vectors).  Vectors: X x, B b, Vv v, P p, T t, D d, S s, Rr r.
template <typename V>
struct Vecs {
  cplx<V>*x, *b, *v, *p, *t, *d, *s, *r;
};

template <typename V>
Vecs<V> vecs(aphid_ctx* c);
template <>
Vecs<double> vecs<double>(aphid_ctx* c) {
  return {c->V[0], c->V[1], c->V[2], c->V[3], c->V[4], c->V[5], c->V[6], c->V[7]};
}
template <>
Vecs<float> vecs<float>(aphid_ctx* c) {
  return {c->F[0], c->F[1], c->F[2], c->F[3], c->F[4], c->F[5], c->F[6], c->F[7]};
}

template <typename V>
const OpArgs<V>& opargs(aphid_ctx* c);
template <>
const OpArgs<double>& opargs<double>(aphid_ctx* c) { return c->op64; }
template <>
const OpArgs<float>& opargs<float>(aphid_ctx* c) { return c->op32; }

template <typename V>
const cplx<V>* invd_of(aphid_ctx* c);
template <>
const cplx<double>* invd_of<double>(aphid_ctx* c) { return c->invd; }
template <>
const cplx<float>* invd_of<float>(aphid_ctx* c) { return c->invdf; }

template <typename V>
int enqueue_iteration(aphid_ctx* c) {
  constexpr int prec = std::is_same<V, float>::value ? 1 : 0;
  const Vecs<V> w = vecs<V>(c);
  const long long n4 = 4 * c->nn;
  const cplx<V>* invd = invd_of<V>(c);
  hipLaunchKernelGGL((k_pupdate<V, 4>), dim3(kRedBlocks), dim3(kBlock), 0, c->s, w.v, invd, w.p, n4,
                     c->st);
  ApplyEpilogue ep{c->st, c->part};
  const Tile& tl = c->tile[prec];
  CK(launch_apply<V>(tl, opargs<V>(c), w.p, w.t, ep, c->s));
  hipLaunchKernelGGL(k_fin_eps, dim3(1), dim3(kBlock), 0, c->s, c->part,
                     (int)apply_blocks(tl, c->nx, c->ny, c->nz), c->st);
  hipLaunchKernelGGL(k_vtnorm<V>, dim3(kRedBlocks), dim3(kBlock), 0, c->s, w.t, w.v, n4, c->st,
                     c->part);
  hipLaunchKernelGGL(k_fin_vtnorm, dim3(1), dim3(kBlock), 0, c->s, c->part, kRedBlocks, c->st);
  hipLaunchKernelGGL(k_update<V>, dim3(kRedBlocks), dim3(kBlock), 0, c->s, w.p, w.d, w.x, w.t, w.s,
                     w.r, w.v, invd, n4, c->st, c->part);
  hipLaunchKernelGGL(k_fin_update, dim3(1), dim3(kBlock), 0, c->s, c->part, kRedBlocks, c->st,
                     c->hist);
  CK(hipGetLastError());
  return 0;
}

// iterations per captured graph: about 3 ms of work per replay
int graph_iters(aphid_ctx* c, int prec) {
  const double bytes = 22.0 * 4.0 * (double)c->nn * (prec ? 8.0 : 16.0);
  const double t_it = std::max(bytes / 4.0e12, 25e-6);
  return std::max(4, std::min(128, (int)std::ceil(3e-3 / t_it)));
}

// APHID_NO_GRAPH=1 enqueues the iterations directly on the stream instead
// of replaying a captured graph (same kernels and launch shapes); profilers
// that intercept hipGraphLaunch need it.
bool use_graphs() {
  const char* e = getenv("APHID_NO_GRAPH");
  return !(e && e[0] == '1');
}

template <typename V>
int run_loop(aphid_ctx* c) {
  constexpr int prec = std::is_same<V, float>::value ? 1 : 0;
  if (!use_graphs()) {
    const int k = graph_iters(c, prec);
    const long long max_launch = (long long)c->hst->max_it / k + 3;
    int last_iter = c->hst->iter;
    for (long long l = 0;; ++l) {
      for (int q = 0; q < k; ++q) {
        const int rc = enqueue_iteration<V>(c);
        if (rc) return rc;
      }
      CK(hipMemcpyAsync(c->hst, c->st, sizeof(QmrState), hipMemcpyDeviceToHost, c->s));
      CK(hipStreamSynchronize(c->s));
      if (c->hst->done) return 0;
      if (c->hst->iter <= last_iter || l > max_launch)
        return fail(APHID_E_STATE, "QMR state did not advance on the device");
      last_iter = c->hst->iter;
    }
  }
  if (!c->gexec[prec]) {
    const int k = graph_iters(c, prec);
    hipGraph_t g;
    CK(hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    for (int q = 0; q < k && !rc; ++q) rc = enqueue_iteration<V>(c);
    hipError_t e = hipStreamEndCapture(c->s, &g);
    if (rc) return rc;
    if (e != hipSuccess) return fail(APHID_E_HIP, std::string("capture: ") + hipGetErrorString(e));
    CK(hipGraphInstantiate(&c->gexec[prec], g, nullptr, nullptr, 0));
    CK(hipGraphDestroy(g));
    c->giters[prec] = k;
  }
  // every replay advances the iteration counter by giters unless the solve
  // stops; bound the replays so a state that fails to advance is an error,
  // never an endless host loop
  const long long max_launch = (long long)c->hst->max_it / c->giters[prec] + 3;
  int last_iter = c->hst->iter;
  for (long long l = 0;; ++l) {
    CK(hipGraphLaunch(c->gexec[prec], c->s));
    CK(hipMemcpyAsync(c->hst, c->st, sizeof(QmrState), hipMemcpyDeviceToHost, c->s));
    CK(hipStreamSynchronize(c->s));
    if (c->hst->done) break;
    if (c->hst->iter <= last_iter || l > max_launch)
      return fail(APHID_E_STATE, "QMR state did not advance on the device");
    last_iter = c->hst->iter;
  }
  return 0;
}

// One QMR solve on the device vectors of precision V: b in w.b, x0 in w.x
// (zero unless have_x0).  Leaves the solution in w.x and the state in
// c->hst.
template <typename V>
int qmr_device(aphid_ctx* c, double tol, int max_it, bool have_x0, double* secs) {
  constexpr int prec = std::is_same<V, float>::value ? 1 : 0;
  const Vecs<V> w = vecs<V>(c);
  const long long n4 = 4 * c->nn;
  QmrState init{};
  init.max_it = max_it;
  init.hist_cap = c->hist_cap;
  init.done = 0;
  CK(hipMemcpyAsync(c->st, &init, sizeof init, hipMemcpyHostToDevice, c->s));
  CK(hipEventRecord(c->ev0, c->s));
  hipLaunchKernelGGL(k_norm_b<V>, dim3(kRedBlocks), dim3(kBlock), 0, c->s, w.b, n4, c->part);
  hipLaunchKernelGGL(k_fin_norm_b, dim3(1), dim3(kBlock), 0, c->s, c->part, kRedBlocks, c->st, tol);
  if (have_x0) {
    ApplyEpilogue ep{nullptr, nullptr};
    CK(launch_apply<V>(c->tile[prec], opargs<V>(c), w.x, w.t, ep, c->s));
  }
  hipLaunchKernelGGL(k_init_r<V>, dim3(kRedBlocks), dim3(kBlock), 0, c->s, w.b,
                     have_x0 ? (const cplx<V>*)w.t : nullptr, w.r, n4, c->part);
  hipLaunchKernelGGL(k_fin_init_r, dim3(1), dim3(kBlock), 0, c->s, c->part, kRedBlocks, c->st);
  hipLaunchKernelGGL(k_init_v<V>, dim3(kRedBlocks), dim3(kBlock), 0, c->s, w.r, invd_of<V>(c), w.v,
                     w.p, w.d, w.s, n4, c->st, c->part);
  hipLaunchKernelGGL(k_fin_init_v, dim3(1), dim3(kBlock), 0, c->s, c->part, kRedBlocks, c->st);
  CK(hipGetLastError());
  CK(hipMemcpyAsync(c->hst, c->st, sizeof(QmrState), hipMemcpyDeviceToHost, c->s));
  CK(hipStreamSynchronize(c->s));
  int rc = 0;
  if (!c->hst->done) rc = run_loop<V>(c);
  if (rc) return rc;
  CK(hipEventRecord(c->ev1, c->s));
  CK(hipEventSynchronize(c->ev1));
  float ms = 0.f;
  CK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  *secs = 1e-3 * ms;
  return 0;
}

// ||b - K x|| / ||b|| in complex128 (x, b: c->V[0], c->V[1]; uses V[8])
int true_residual(aphid_ctx* c, double* rel, double* rnorm, double* bnorm) {
  ApplyEpilogue ep{nullptr, nullptr};
  CK(launch_apply<double>(c->tile[0], c->op64, c->V[0], c->V[8], ep, c->s));
  hipLaunchKernelGGL(k_resnorm<double>, dim3(kRedBlocks), dim3(kBlock), 0, c->s, c->V[1], c->V[8],
                     4 * c->nn, c->part);
  hipLaunchKernelGGL(k_fin_resnorm, dim3(1), dim3(kBlock), 0, c->s, c->part, kRedBlocks,
                     c->resbuf);
  CK(hipGetLastError());
  double h[2];
  CK(hipMemcpyAsync(h, c->resbuf, sizeof h, hipMemcpyDeviceToHost, c->s));
  CK(hipStreamSynchronize(c->s));
  *rnorm = h[0];
  *bnorm = h[1];
  *rel = h[1] > 0 ? h[0] / h[1] : h[0];
  return 0;
}

// Mixed precision (APHID_F_MIXED): FP64 iterative refinement on the true
// residual around complex64 QMR solves.  x (V[0]) and b (V[1]) stay
// complex128; each sweep solves K e = (b - K x)/||b - K x|| with the
// complex64 stencil and vectors (FP64 inner products and scalars) to a
// relative residual inner_tol, then x += ||b - K x|| e in FP64.
double mixed_inner_floor() {
  const char* e = getenv("APHID_MIXED_INNER_TOL");
  const double v = e ? atof(e) : 0.0;
  return v > 0.0 ? v : 1e-5;
}

int solve_mixed(aphid_ctx* c, double tol, int max_it, bool have_x0, QmrState* out,
                int* sweeps_out, std::vector<double>* sweep_res, double* secs) {
  const auto t0 = std::chrono::steady_clock::now();
  int rc = ensure_vectors(c, true);
  if (rc) return rc;
  rc = autotune(c, 1);
  if (rc) return rc;
  const long long n4 = 4 * c->nn;
  if (!have_x0) CK(hipMemsetAsync(c->V[0], 0, sizeof(c128) * n4, c->s));
  const double floor_tol = mixed_inner_floor();
  int total = 0, sweeps = 0, status = ST_MAXIT;
  double rel = NAN, rn = NAN, bn = NAN;
  QmrState last{};
  for (;;) {
    rc = true_residual(c, &rel, &rn, &bn);  // t = K x in V[8]; resbuf = [||b-Kx||, ||b||]
    if (rc) return rc;
    sweep_res->push_back(rel);
    if (rn < tol * bn || bn == 0.0) {
      status = ST_CONV;
      break;
    }
    if (!std::isfinite(rn)) {
      status = ST_BREAK;
      break;
    }
    if (total >= std::max(max_it, 1) || sweeps >= 64) break;
    hipLaunchKernelGGL(k_inner_rhs, dim3(grid_for(n4)), dim3(kBlock), 0, c->s, c->V[1], c->V[8],
                       c->resbuf, c->F[1], n4);
    CK(hipGetLastError());
    CK(hipMemsetAsync(c->F[0], 0, sizeof(c64) * n4, c->s));
    const double inner_tol = std::max(0.5 * tol * bn / rn, floor_tol);
    double s_in = 0.0;
    rc = qmr_device<float>(c, inner_tol, std::max(max_it - total, 1), false, &s_in);
    if (rc) return rc;
    last = *c->hst;
    const int executed = last.status == ST_MAXIT ? std::max(max_it - total, 1) : last.iter;
    total += std::max(executed, 0);
    if (last.status == ST_BREAK && executed <= 1) {
      status = ST_BREAK;
      break;
    }
    hipLaunchKernelGGL(k_refine, dim3(grid_for(n4)), dim3(kBlock), 0, c->s, c->V[0], c->F[0],
                       c->resbuf, n4);
    CK(hipGetLastError());
    ++sweeps;
  }
  CK(hipStreamSynchronize(c->s));
  *secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  *out = last;
  out->iter = status == ST_MAXIT ? std::max(max_it, 1) + 1 : total;
  out->status = status;
  out->rnorm = rn;
  out->rhsnorm = bn;
  *sweeps_out = sweeps;
  return 0;
}

bool need_mesh(aphid_ctx* c) { return c->nn == 0; }

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
extern "C" {

const char* aphid_last_error(void) { return g_err.c_str(); }

int aphid_ctx_create(aphid_ctx** out, int device) {
  if (!out) return fail(APHID_E_ARG, "ctx pointer is NULL");
  *out = nullptr;
  int ndev = 0;
  CK(hipGetDeviceCount(&ndev));
  if (ndev <= 0) return fail(APHID_E_HIP, "no HIP device");
  if (device < 0) CK(hipGetDevice(&device));
  if (device >= ndev) return fail(APHID_E_ARG, "device index out of range");
  CK(hipSetDevice(device));
  aphid_ctx* c = new aphid_ctx();
  c->dev = device;
  CK(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
  CK(hipMalloc(&c->st, sizeof(QmrState)));
  CK(hipHostMalloc(&c->hst, sizeof(QmrState), hipHostMallocDefault));
  CK(hipMalloc(&c->part, sizeof(double) * 4 * kMaxPartBlocks));
  CK(hipMalloc(&c->resbuf, sizeof(double) * 4));
  CK(hipMemset(c->st, 0, sizeof(QmrState)));
  CK(hipEventCreate(&c->ev0));
  CK(hipEventCreate(&c->ev1));
  *out = c;
  return 0;
}

int aphid_ctx_destroy(aphid_ctx* c) {
  if (!c) return 0;
  hipSetDevice(c->dev);
  if (c->s) hipStreamSynchronize(c->s);
  free_mesh(c);
  if (c->st) hipFree(c->st);
  if (c->hst) hipHostFree(c->hst);
  if (c->part) hipFree(c->part);
  if (c->resbuf) hipFree(c->resbuf);
  if (c->ev0) hipEventDestroy(c->ev0);
  if (c->ev1) hipEventDestroy(c->ev1);
  if (c->s) hipStreamDestroy(c->s);
  delete c;
  return 0;
}

int aphid_set_mesh(aphid_ctx* c, int nx, int ny, int nz, const double* xn, const double* yn,
                   const double* zn) {
  if (!c || !xn || !yn || !zn) return fail(APHID_E_ARG, "NULL argument");
  if (nx < 3 || ny < 3 || nz < 3) return fail(APHID_E_ARG, "need at least 3 nodes per axis");
  for (int a = 0; a < 3; ++a) {
    const double* v = a == 0 ? xn : (a == 1 ? yn : zn);
    const int n = a == 0 ? nx : (a == 1 ? ny : nz);
    for (int i = 1; i < n; ++i)
      if (!(v[i] > v[i - 1])) return fail(APHID_E_ARG, "node coordinates must increase");
  }
  CK(hipSetDevice(c->dev));
  CK(hipStreamSynchronize(c->s));
  free_mesh(c);
  c->nx = nx;
  c->ny = ny;
  c->nz = nz;
  c->nn = (long long)nx * ny * nz;
  const int n[3] = {nx, ny, nz};
  c->ax[0] = make_axis(xn, nx);
  c->ax[1] = make_axis(yn, ny);
  c->ax[2] = make_axis(zn, nz);
  // pack 8 tables per axis, each of length n_a
  std::vector<double> h64;
  size_t off[3][8];
  for (int a = 0; a < 3; ++a) {
    const AxisTables& t = c->ax[a];
    const std::vector<double>* arrs[8] = {&t.eP, &t.eM, &t.nP, &t.nM, &t.dE, &t.rdE, &t.dN, &t.rdN};
    for (int q = 0; q < 8; ++q) {
      off[a][q] = h64.size();
      h64.insert(h64.end(), arrs[q]->begin(), arrs[q]->end());
      (void)n;
    }
  }
  std::vector<float> h32(h64.begin(), h64.end());
  CK(hipMalloc(&c->tab64, sizeof(double) * h64.size()));
  CK(hipMalloc(&c->tab32, sizeof(float) * h32.size()));
  CK(hipMemcpy(c->tab64, h64.data(), sizeof(double) * h64.size(), hipMemcpyHostToDevice));
  CK(hipMemcpy(c->tab32, h32.data(), sizeof(float) * h32.size(), hipMemcpyHostToDevice));
  std::vector<double> dm;
  for (int a = 0; a < 3; ++a) dm.insert(dm.end(), c->ax[a].d.begin(), c->ax[a].d.end());
  CK(hipMalloc(&c->dmet, sizeof(double) * dm.size()));
  CK(hipMemcpy(c->dmet, dm.data(), sizeof(double) * dm.size(), hipMemcpyHostToDevice));
  const size_t n4 = 4 * (size_t)c->nn;
  CK(hipMalloc(&c->k2, sizeof(c128) * n4));
  CK(hipMalloc(&c->k2f, sizeof(c64) * n4));
  CK(hipMalloc(&c->diag, sizeof(c128) * n4));
  CK(hipMalloc(&c->invd, sizeof(c128) * n4));
  CK(hipMalloc(&c->invdf, sizeof(c64) * n4));
  CK(hipMalloc(&c->stage, sizeof(c128) * (size_t)c->nn));
  CK(hipMemset(c->k2, 0, sizeof(c128) * n4));
  CK(hipMemset(c->k2f, 0, sizeof(c64) * n4));
  for (int a = 0; a < 3; ++a) {
    c->op64.eP[a] = c->tab64 + off[a][0];
    c->op64.eM[a] = c->tab64 + off[a][1];
    c->op64.nP[a] = c->tab64 + off[a][2];
    c->op64.nM[a] = c->tab64 + off[a][3];
    c->op64.dE[a] = c->tab64 + off[a][4];
    c->op64.rdE[a] = c->tab64 + off[a][5];
    c->op64.dN[a] = c->tab64 + off[a][6];
    c->op64.rdN[a] = c->tab64 + off[a][7];
    c->op32.eP[a] = c->tab32 + off[a][0];
    c->op32.eM[a] = c->tab32 + off[a][1];
    c->op32.nP[a] = c->tab32 + off[a][2];
    c->op32.nM[a] = c->tab32 + off[a][3];
    c->op32.dE[a] = c->tab32 + off[a][4];
    c->op32.rdE[a] = c->tab32 + off[a][5];
    c->op32.dN[a] = c->tab32 + off[a][6];
    c->op32.rdN[a] = c->tab32 + off[a][7];
  }
  c->op64.nx = c->op32.nx = nx;
  c->op64.ny = c->op32.ny = ny;
  c->op64.nz = c->op32.nz = nz;
  c->op64.nn = c->op32.nn = c->nn;
  c->op64.k2 = c->k2;
  c->op32.k2 = c->k2f;
  return 0;
}

int aphid_set_sigma(aphid_ctx* c, const double* sig_c, double omega) {
  if (!c || !sig_c) return fail(APHID_E_ARG, "NULL argument");
  if (need_mesh(c)) return fail(APHID_E_STATE, "aphid_set_mesh first");
  CK(hipSetDevice(c->dev));
  const long long nc = (long long)(c->nx - 1) * (c->ny - 1) * (c->nz - 1);
  if (!c->sig) CK(hipMalloc(&c->sig, sizeof(c128) * nc));
  // Fortran-order cells -> device cell layout
  CK(hipMemcpyAsync(c->stage, sig_c, sizeof(c128) * nc, hipMemcpyHostToDevice, c->s));
  Dims dc{c->nx - 1, c->ny - 1, c->nz - 1, nc};
  hipLaunchKernelGGL(k_f2d<double>, dim3(grid_for(nc)), dim3(kBlock), 0, c->s, c->stage, c->nx - 1,
                     c->ny - 1, c->nz - 1, c->sig, dc);
  CK(hipGetLastError());
  int rc = build_k2(c, omega);
  if (rc) return rc;
  CK(hipStreamSynchronize(c->s));
  return 0;
}

int aphid_set_cells(aphid_ctx* c, char model, double omega, const float* a, const float* b,
                    const float* cc, const float* d) {
  if (!c) return fail(APHID_E_ARG, "NULL context");
  if (need_mesh(c)) return fail(APHID_E_STATE, "aphid_set_mesh first");
  const bool need[4] = {model != 'I', model == 'C' || model == 'D' || model == 'I', model == 'D',
                        model == 'D'};
  if (model != 'R' && model != 'C' && model != 'D' && model != 'I')
    return fail(APHID_E_ARG, "model must be R, C, D or I");
  const float* in[4] = {a, b, cc, d};
  for (int q = 0; q < 4; ++q)
    if (need[q] && !in[q]) return fail(APHID_E_ARG, "missing cell parameter block");
  CK(hipSetDevice(c->dev));
  const long long nc = (long long)(c->nx - 1) * (c->ny - 1) * (c->nz - 1);
  if (!c->sig) CK(hipMalloc(&c->sig, sizeof(c128) * nc));
  if (!c->stagef) CK(hipMalloc(&c->stagef, sizeof(float) * 4 * nc));
  const float* dptr[4] = {nullptr, nullptr, nullptr, nullptr};
  for (int q = 0; q < 4; ++q)
    if (need[q]) {
      CK(hipMemcpyAsync(c->stagef + q * nc, in[q], sizeof(float) * nc, hipMemcpyHostToDevice, c->s));
      dptr[q] = c->stagef + q * nc;
    }
  hipLaunchKernelGGL(k_sigma, dim3(grid_for(nc)), dim3(kBlock), 0, c->s, model, omega, dptr[0],
                     dptr[1], dptr[2], dptr[3], c->sig, dims(c));
  CK(hipGetLastError());
  int rc = build_k2(c, omega);
  if (rc) return rc;
  CK(hipStreamSynchronize(c->s));
  return 0;
}

int aphid_set_k2(aphid_ctx* c, const double* kx, const double* ky, const double* kz,
                 const double* kp) {
  if (!c || !kx || !ky || !kz || !kp) return fail(APHID_E_ARG, "NULL argument");
  if (need_mesh(c)) return fail(APHID_E_STATE, "aphid_set_mesh first");
  CK(hipSetDevice(c->dev));
  const double* h[4] = {kx, ky, kz, kp};
  int rc = upload4(c, h, c->k2);
  if (rc) return rc;
  c->have_k2 = true;
  rc = build_diag(c);
  if (rc) return rc;
  CK(hipStreamSynchronize(c->s));
  return 0;
}

int aphid_set_diag(aphid_ctx* c, const double* dx, const double* dy, const double* dz,
                   const double* dp) {
  if (!c || !dx || !dy || !dz || !dp) return fail(APHID_E_ARG, "NULL argument");
  if (need_mesh(c)) return fail(APHID_E_STATE, "aphid_set_mesh first");
  CK(hipSetDevice(c->dev));
  const double* h[4] = {dx, dy, dz, dp};
  int rc = upload4(c, h, c->diag);
  if (rc) return rc;
  hipLaunchKernelGGL(k_invert, dim3(grid_for(4 * c->nn)), dim3(kBlock), 0, c->s, c->diag, c->invd,
                     dims(c));
  CK(hipGetLastError());
  c->have_diag = true;
  rc = refresh_f32(c);
  if (rc) return rc;
  CK(hipStreamSynchronize(c->s));
  return 0;
}

int aphid_get_k2(aphid_ctx* c, double* kx, double* ky, double* kz, double* kp) {
  if (!c || !kx || !ky || !kz || !kp) return fail(APHID_E_ARG, "NULL argument");
  if (!c->have_k2) return fail(APHID_E_STATE, "no k2 on the device");
  CK(hipSetDevice(c->dev));
  double* h[4] = {kx, ky, kz, kp};
  return download4(c, c->k2, h);
}

int aphid_get_diag(aphid_ctx* c, double* dx, double* dy, double* dz, double* dp) {
  if (!c || !dx || !dy || !dz || !dp) return fail(APHID_E_ARG, "NULL argument");
  if (!c->have_diag) return fail(APHID_E_STATE, "no diagonal on the device");
  CK(hipSetDevice(c->dev));
  double* h[4] = {dx, dy, dz, dp};
  return download4(c, c->diag, h);
}

int aphid_apply(aphid_ctx* c, const double* ix, const double* iy, const double* iz,
                const double* ip, double* ox, double* oy, double* oz, double* op) {
  if (!c || !ix || !iy || !iz || !ip || !ox || !oy || !oz || !op)
    return fail(APHID_E_ARG, "NULL argument");
  if (!c->have_k2) return fail(APHID_E_STATE, "set k2 (sigma, cells or k2) first");
  CK(hipSetDevice(c->dev));
  int rc = ensure_vectors(c, false);
  if (rc) return rc;
  rc = autotune(c, 0);
  if (rc) return rc;
  const double* hin[4] = {ix, iy, iz, ip};
  rc = upload4(c, hin, c->V[3]);
  if (rc) return rc;
  ApplyEpilogue ep{nullptr, nullptr};
  CK(launch_apply<double>(c->tile[0], c->op64, c->V[3], c->V[4], ep, c->s));
  double* hout[4] = {ox, oy, oz, op};
  return download4(c, c->V[4], hout);
}

int aphid_apply_prec(aphid_ctx* c, int precision, const double* ix, const double* iy,
                     const double* iz, const double* ip, double* ox, double* oy, double* oz,
                     double* op) {
  if (precision == 0) return aphid_apply(c, ix, iy, iz, ip, ox, oy, oz, op);
  if (precision != 1) return fail(APHID_E_ARG, "precision must be 0 or 1");
  if (!c || !ix || !iy || !iz || !ip || !ox || !oy || !oz || !op)
    return fail(APHID_E_ARG, "NULL argument");
  if (!c->have_k2) return fail(APHID_E_STATE, "set k2 (sigma, cells or k2) first");
  CK(hipSetDevice(c->dev));
  int rc = ensure_vectors(c, true);
  if (rc) return rc;
  rc = autotune(c, 1);
  if (rc) return rc;
  const double* hin[4] = {ix, iy, iz, ip};
  rc = upload4(c, hin, c->V[3]);
  if (rc) return rc;
  const long long n4 = 4 * c->nn;
  hipLaunchKernelGGL((k_cast<float, double>), dim3(grid_for(n4)), dim3(kBlock), 0, c->s, c->V[3],
                     c->F[3], n4);
  ApplyEpilogue ep{nullptr, nullptr};
  CK(launch_apply<float>(c->tile[1], c->op32, c->F[3], c->F[4], ep, c->s));
  hipLaunchKernelGGL((k_cast<double, float>), dim3(grid_for(n4)), dim3(kBlock), 0, c->s, c->F[4],
                     c->V[4], n4);
  CK(hipGetLastError());
  double* hout[4] = {ox, oy, oz, op};
  return download4(c, c->V[4], hout);
}

int aphid_build_rhs(aphid_ctx* c, int mode, int nsrc, const int* dir, const int* ip_,
                    const int* jp_, const int* kp_, int* skipped, double* bx, double* by,
                    double* bz, double* bp) {
  if (!c || !bx || !by || !bz || !bp) return fail(APHID_E_ARG, "NULL argument");
  if (need_mesh(c)) return fail(APHID_E_STATE, "aphid_set_mesh first");
  if (mode == 5 && nsrc > 0 && (!dir || !ip_ || !jp_ || !kp_))
    return fail(APHID_E_ARG, "mode 5 needs the dipole arrays");
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  const double MU0 = APHID_MU0;
  const std::vector<double>&dx = c->ax[0].d, &dy = c->ax[1].d, &dz = c->ax[2].d;
  const size_t nxs = 2 * (size_t)(nx - 1) * ny * nz, nys = 2 * (size_t)nx * (ny - 1) * nz;
  const size_t nzs = 2 * (size_t)nx * ny * (nz - 1), nps = 2 * (size_t)nx * ny * nz;
  std::fill(bx, bx + nxs, 0.0);
  std::fill(by, by + nys, 0.0);
  std::fill(bz, bz + nzs, 0.0);
  std::fill(bp, bp + nps, 0.0);
  // 1-based Fortran element (re part) of each component
  auto X = [&](int i, int j, int k) -> double& { return bx[2 * (((size_t)(k - 1) * ny + (j - 1)) * (nx - 1) + (i - 1))]; };
  auto Y = [&](int i, int j, int k) -> double& { return by[2 * (((size_t)(k - 1) * (ny - 1) + (j - 1)) * nx + (i - 1))]; };
  auto Z = [&](int i, int j, int k) -> double& { return bz[2 * (((size_t)(k - 1) * ny + (j - 1)) * nx + (i - 1))]; };
  auto P = [&](int i, int j, int k) -> double& { return bp[2 * (((size_t)(k - 1) * ny + (j - 1)) * nx + (i - 1))]; };
  if (mode == 1) {  // x dipole at the origin node (src/aphid.f90:566-587)
    const double tol = (double)1.0e-4f;  // the real(4) literal 1.e-4
    int ip = -1, jp = -1, kp = -1;  // last match wins, as in the loop nest
    for (int k = 1; k <= nz; ++k) {
      if (std::fabs(c->ax[2].xn[k - 1]) < tol) kp = k;
      for (int j = 1; j <= ny; ++j) {
        if (std::fabs(c->ax[1].xn[j - 1]) < tol) jp = j;
        for (int i = 1; i <= nx; ++i)
          if (std::fabs(c->ax[0].xn[i - 1]) < tol) ip = i;
      }
    }
    if (ip < 1 || jp < 1 || kp < 1 || ip > nx - 1)
      return fail(APHID_E_ARG, "mode 1: no node at the origin");
    X(ip, jp, kp) = 1.0 * MU0;
    P(ip, jp, kp) = 1.0 * MU0 / dx[ip];
    P(ip + 1, jp, kp) = -1.0 * MU0 / dx[ip];
  } else if (mode == 2) {  // x-polarised MT sheet (:592-604)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx - 1; ++i) X(i, j, 2) = 1.0 * MU0;
    for (int i = 1; i <= nx - 1; ++i)
      for (int j = 1; j <= ny; ++j) X(i, j, 2) = X(i, j, 2) * dx[i];
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx - 1; ++i) X(i, j, 2) = X(i, j, 2) * (0.5 * (dy[j - 1] + dy[j]));
  } else if (mode == 3) {  // y-polarised MT sheet (:608-620)
    for (int j = 1; j <= ny - 1; ++j)
      for (int i = 1; i <= nx; ++i) Y(i, j, 2) = 1.0 * MU0;
    for (int j = 1; j <= ny - 1; ++j)
      for (int i = 1; i <= nx; ++i) Y(i, j, 2) = Y(i, j, 2) * dy[j];
    for (int i = 1; i <= nx; ++i)
      for (int j = 1; j <= ny - 1; ++j) Y(i, j, 2) = Y(i, j, 2) * (0.5 * (dx[i - 1] + dx[i]));
  } else if (mode == 4) {  // fixed square loop (:624-631)
    const int n = 20;
    if (nx < 72 || ny < 72 || nz < 51) return fail(APHID_E_ARG, "mode 4 needs >= 72x72x51 nodes");
    for (int i = 51 - n; i <= 51 + n - 1; ++i) {
      X(i, 51 + n, 51) = 1.0 * MU0 * dx[i];
      X(i, 51 - n, 51) = -1.0 * MU0 * dx[i];
    }
    for (int j = 51 - n; j <= 51 + n - 1; ++j) {
      Y(51 - n, j, 51) = 1.0 * MU0 * dy[j];
      Y(51 + n, j, 51) = -1.0 * MU0 * dy[j];
    }
  } else if (mode == 5) {  // dipole list (:638-741)
    for (int s = 0; s < nsrc; ++s) {
      const int ip = ip_[s], jp = jp_[s], kp = kp_[s];
      bool skip = false;
      switch (dir[s]) {
        case +1:
          if (ip < 1 || ip > nx - 1 || jp < 1 || jp > ny || kp < 1 || kp > nz) { skip = true; break; }
          X(ip, jp, kp) += 1.0 * MU0 * dx[ip];
          P(ip, jp, kp) += 1.0 * MU0;
          P(ip + 1, jp, kp) -= 1.0 * MU0;
          break;
        case -1:
          if (ip < 2 || ip > nx || jp < 1 || jp > ny || kp < 1 || kp > nz) { skip = true; break; }
          X(ip - 1, jp, kp) -= 1.0 * MU0 * dx[ip - 1];
          P(ip - 1, jp, kp) -= 1.0 * MU0;
          P(ip, jp, kp) += 1.0 * MU0;
          break;
        case +2:
          if (ip < 1 || ip > nx || jp < 1 || jp > ny - 1 || kp < 1 || kp > nz) { skip = true; break; }
          Y(ip, jp, kp) += 1.0 * MU0 * dy[jp];
          P(ip, jp, kp) += 1.0 * MU0;
          P(ip, jp + 1, kp) -= 1.0 * MU0;
          break;
        case -2:
          if (ip < 1 || ip > nx || jp < 2 || jp > ny || kp < 1 || kp > nz) { skip = true; break; }
          Y(ip, jp - 1, kp) -= 1.0 * MU0 * dy[jp - 1];
          P(ip, jp - 1, kp) -= 1.0 * MU0;
          P(ip, jp, kp) += 1.0 * MU0;
          break;
        case +3:
          if (ip < 1 || ip > nx || jp < 1 || jp > ny || kp < 1 || kp > nz - 1) { skip = true; break; }
          Z(ip, jp, kp) += 1.0 * MU0 * dz[kp];
          P(ip, jp, kp) += 1.0 * MU0;
          P(ip, jp, kp + 1) -= 1.0 * MU0;
          break;
        case -3:
          if (ip < 1 || ip > nx || jp < 1 || jp > ny || kp < 2 || kp > nz) { skip = true; break; }
          Z(ip, jp, kp - 1) -= 1.0 * MU0 * dz[kp - 1];
          P(ip, jp, kp - 1) -= 1.0 * MU0;
          P(ip, jp, kp) += 1.0 * MU0;
          break;
        default:
          skip = true;
      }
      if (skipped) skipped[s] = skip ? 1 : 0;
    }
  } else {
    return fail(APHID_E_ARG, "mode must be 1..5");
  }
  return 0;
}

int aphid_load_rhs(aphid_ctx* c, const double* bx, const double* by, const double* bz,
                   const double* bp, const double* x0_x, const double* x0_y, const double* x0_z,
                   const double* x0_p) {
  if (!c || !bx || !by || !bz || !bp) return fail(APHID_E_ARG, "NULL argument");
  if (need_mesh(c)) return fail(APHID_E_STATE, "aphid_set_mesh first");
  const bool have_x0 = x0_x || x0_y || x0_z || x0_p;
  if (have_x0 && !(x0_x && x0_y && x0_z && x0_p)) return fail(APHID_E_ARG, "x0 arrays incomplete");
  CK(hipSetDevice(c->dev));
  int rc = ensure_vectors(c, false);
  if (rc) return rc;
  const double* hb[4] = {bx, by, bz, bp};
  rc = upload4(c, hb, c->V[1]);
  if (rc) return rc;
  if (have_x0) {
    const double* h0[4] = {x0_x, x0_y, x0_z, x0_p};
    rc = upload4(c, h0, c->V[0]);
  } else {
    CK(hipMemsetAsync(c->V[0], 0, sizeof(c128) * 4 * c->nn, c->s));
  }
  if (rc) return rc;
  CK(hipStreamSynchronize(c->s));
  return 0;
}

int aphid_solve_resident(aphid_ctx* c, double tol, int max_it, unsigned flags,
                         aphid_qmr_report* rep) {
  const auto t0 = std::chrono::steady_clock::now();
  if (!c) return fail(APHID_E_ARG, "NULL context");
  if (!c->have_k2 || !c->have_diag) return fail(APHID_E_STATE, "operator not set up");
  if (!c->V[1]) return fail(APHID_E_STATE, "aphid_load_rhs first");
  if (!(tol >= 0.0) || max_it < 0) return fail(APHID_E_ARG, "tol and max_it must be >= 0");
  const bool have_x0 = flags & APHID_F_X0;
  CK(hipSetDevice(c->dev));
  int rc = autotune(c, 0);
  if (rc) return rc;
  if (flags & APHID_F_MIXED) {
    QmrState hs{};
    int sweeps = 0;
    std::vector<double> res;
    double secs = 0.0;
    rc = solve_mixed(c, tol, max_it, have_x0, &hs, &sweeps, &res, &secs);
    if (rc) return rc;
    if (rep) {
      rep->iters = hs.iter;
      rep->status = hs.status;
      rep->rnorm = hs.rnorm;
      rep->rhsnorm = hs.rhsnorm;
      rep->true_rel_res = res.empty() ? NAN : res.back();
      rep->seconds_qmr = secs;
      rep->outer_iters = sweeps;
      const int cap = (rep->hist && rep->hist_cap > 0) ? rep->hist_cap : 0;
      rep->hist_len = std::min<int>((int)res.size(), cap);
      for (int q = 0; q < rep->hist_len; ++q) rep->hist[q] = res[q];
      rep->seconds_total =
          std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return hs.status;
  }
  const int cap = (rep && rep->hist && rep->hist_cap > 0) ? rep->hist_cap : 0;
  if (cap > c->hist_cap) {
    if (c->hist) CK(hipFree(c->hist));
    CK(hipMalloc(&c->hist, sizeof(double) * cap));
    c->hist_cap = cap;
    for (auto& g : c->gexec)
      if (g) hipGraphExecDestroy(g), g = nullptr;  // graphs captured the old pointer
  }
  if (!have_x0) CK(hipMemsetAsync(c->V[0], 0, sizeof(c128) * 4 * c->nn, c->s));
  double secs = 0.0;
  rc = qmr_device<double>(c, tol, max_it, have_x0, &secs);
  if (rc) return rc;
  const QmrState hs = *c->hst;
  double rel = NAN, rn = NAN, bn = NAN;
  if (!(flags & APHID_F_NO_TRUE_RES)) {
    rc = true_residual(c, &rel, &rn, &bn);
    if (rc) return rc;
  }
  if (rep) {
    rep->iters = hs.iter;
    rep->status = hs.status;
    rep->rnorm = hs.rnorm;
    rep->rhsnorm = hs.rhsnorm;
    rep->true_rel_res = rel;
    rep->seconds_qmr = secs;
    rep->outer_iters = 1;
    const int executed = hs.status == ST_MAXIT ? std::max(max_it, 1) : hs.iter;
    rep->hist_len = std::min(std::max(executed, 0), cap);
    if (cap && rep->hist_len > 0)
      CK(hipMemcpy(rep->hist, c->hist, sizeof(double) * rep->hist_len, hipMemcpyDeviceToHost));
    rep->seconds_total =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  return hs.status;
}

int aphid_fetch_solution(aphid_ctx* c, double* x_x, double* x_y, double* x_z, double* x_p) {
  if (!c || !x_x || !x_y || !x_z || !x_p) return fail(APHID_E_ARG, "NULL argument");
  if (!c->V[0]) return fail(APHID_E_STATE, "no resident solution");
  CK(hipSetDevice(c->dev));
  double* h[4] = {x_x, x_y, x_z, x_p};
  return download4(c, c->V[0], h);
}

int aphid_qmr_solve(aphid_ctx* c, double* x_x, double* x_y, double* x_z, double* x_p,
                    const double* x0_x, const double* x0_y, const double* x0_z,
                    const double* x0_p, double tol, int max_it, unsigned flags,
                    aphid_qmr_report* rep) {
  const auto t0 = std::chrono::steady_clock::now();
  if (!c || !x_x || !x_y || !x_z || !x_p) return fail(APHID_E_ARG, "NULL argument");
  if (!c->have_k2 || !c->have_diag) return fail(APHID_E_STATE, "operator not set up");
  const bool have_x0 = flags & APHID_F_X0;
  if (have_x0 && (!x0_x || !x0_y || !x0_z || !x0_p)) return fail(APHID_E_ARG, "x0 arrays missing");
  int rc = aphid_load_rhs(c, x_x, x_y, x_z, x_p, have_x0 ? x0_x : nullptr,
                          have_x0 ? x0_y : nullptr, have_x0 ? x0_z : nullptr,
                          have_x0 ? x0_p : nullptr);
  if (rc) return rc;
  const int st = aphid_solve_resident(c, tol, max_it, flags, rep);
  if (st < 0) return st;
  rc = aphid_fetch_solution(c, x_x, x_y, x_z, x_p);
  if (rc) return rc;
  if (rep)
    rep->seconds_total =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return st;
}

int aphid_get_tile(aphid_ctx* c, int precision, int* tk, int* tj, int* chunk) {
  if (!c || precision < 0 || precision > 1 || !tk || !tj || !chunk)
    return fail(APHID_E_ARG, "bad argument");
  *tk = c->tile[precision].tk;
  *tj = c->tile[precision].tj;
  *chunk = c->tile[precision].chunk;
  return 0;
}

int aphid_set_tile(aphid_ctx* c, int precision, int tk, int tj, int chunk) {
  if (!c || precision < 0 || precision > 1 || chunk < 0) return fail(APHID_E_ARG, "bad argument");
  bool ok = false;
  for (const Tile& t : kTiles) ok |= (t.tk == tk && t.tj == tj);
  if (!ok) return fail(APHID_E_ARG, "unsupported tile");
  if (c->nn && apply_blocks(Tile{tk, tj, chunk}, c->nx, c->ny, c->nz) > (unsigned)kMaxPartBlocks)
    return fail(APHID_E_ARG, "too many blocks for this tile");
  c->tile[precision] = {tk, tj, chunk};
  c->tuned[precision] = true;
  if (c->gexec[precision]) hipGraphExecDestroy(c->gexec[precision]), c->gexec[precision] = nullptr;
  return 0;
}

int aphid_time_apply(aphid_ctx* c, int precision, int reps, double* seconds) {
  if (!c || !seconds || reps < 1 || precision < 0 || precision > 1)
    return fail(APHID_E_ARG, "bad argument");
  if (!c->have_k2) return fail(APHID_E_STATE, "operator not set up");
  CK(hipSetDevice(c->dev));
  int rc = ensure_vectors(c, precision == 1);
  if (rc) return rc;
  rc = autotune(c, precision);
  if (rc) return rc;
  float ms = 0.f;
  if (precision == 0)
    rc = time_tile<double>(c, c->tile[0], c->op64, c->V[3], c->V[4], reps, &ms);
  else
    rc = time_tile<float>(c, c->tile[1], c->op32, c->F[3], c->F[4], reps, &ms);
  if (rc) return rc;
  *seconds = 1e-3 * ms;
  return 0;
}

int aphid_time_iteration(aphid_ctx* c, int precision, int reps, double* secs) {
  if (!c || !secs || reps < 1 || precision != 0) return fail(APHID_E_ARG, "bad argument");
  if (!c->have_k2 || !c->V[0]) return fail(APHID_E_STATE, "solve once first");
  CK(hipSetDevice(c->dev));
  // continue the Krylov process of the last solve with the stop test off
  QmrState hs = *c->hst;
  hs.done = 0;
  hs.status = ST_MAXIT;
  hs.iter = 1;
  hs.max_it = 1 << 30;
  hs.target = 0.0;
  if (!(hs.rho_np > 0.0) || !std::isfinite(hs.rho_np) || is_zero(hs.del))
    return fail(APHID_E_STATE, "the last solve left no usable Krylov state");
  CK(hipMemcpyAsync(c->st, &hs, sizeof hs, hipMemcpyHostToDevice, c->s));
  const Vecs<double> w = vecs<double>(c);
  const long long n4 = 4 * c->nn;
  auto K0 = [&]() {
    hipLaunchKernelGGL((k_pupdate<double, 4>), dim3(kRedBlocks), dim3(kBlock), 0, c->s, w.v,
                       c->invd, w.p, n4, c->st);
  };
  auto K1 = [&]() {
    ApplyEpilogue ep{c->st, c->part};
    launch_apply<double>(c->tile[0], c->op64, w.p, w.t, ep, c->s);
    hipLaunchKernelGGL(k_fin_eps, dim3(1), dim3(kBlock), 0, c->s, c->part,
                       (int)apply_blocks(c->tile[0], c->nx, c->ny, c->nz), c->st);
  };
  auto K2 = [&]() {
    hipLaunchKernelGGL(k_vtnorm<double>, dim3(kRedBlocks), dim3(kBlock), 0, c->s, w.t, w.v, n4,
                       c->st, c->part);
    hipLaunchKernelGGL(k_fin_vtnorm, dim3(1), dim3(kBlock), 0, c->s, c->part, kRedBlocks, c->st);
  };
  auto K3 = [&]() {
    hipLaunchKernelGGL(k_update<double>, dim3(kRedBlocks), dim3(kBlock), 0, c->s, w.p, w.d, w.x,
                       w.t, w.s, w.r, w.v, c->invd, n4, c->st, c->part);
    hipLaunchKernelGGL(k_fin_update, dim3(1), dim3(kBlock), 0, c->s, c->part, kRedBlocks, c->st,
                       nullptr);
  };
  float ms[5] = {0, 0, 0, 0, 0};
  for (int rep = 0; rep < reps + 1; ++rep) {  // first sweep warms up
    for (int q = 0; q < 4; ++q) {
      CK(hipEventRecord(c->ev0, c->s));
      if (q == 0) K0();
      if (q == 1) K1();
      if (q == 2) K2();
      if (q == 3) K3();
      CK(hipEventRecord(c->ev1, c->s));
      CK(hipEventSynchronize(c->ev1));
      float t = 0.f;
      CK(hipEventElapsedTime(&t, c->ev0, c->ev1));
      if (rep) ms[q] += t;
    }
  }
  CK(hipGetLastError());
  CK(hipEventRecord(c->ev0, c->s));
  for (int rep = 0; rep < reps; ++rep) {
    K0();
    K1();
    K2();
    K3();
  }
  CK(hipEventRecord(c->ev1, c->s));
  CK(hipEventSynchronize(c->ev1));
  CK(hipEventElapsedTime(&ms[4], c->ev0, c->ev1));
  for (int q = 0; q < 5; ++q) secs[q] = 1e-3 * ms[q] / reps;
  CK(hipMemcpy(c->hst, c->st, sizeof(QmrState), hipMemcpyDeviceToHost));
  if (c->hst->done) return fail(APHID_E_STATE, "Krylov process broke down while timing");
  return 0;
}

int aphid_time_copy(aphid_ctx* c, double bytes, int reps, double* seconds) {
  if (!c || !seconds || reps < 1 || !(bytes >= 16.0) || bytes > 64.0 * (1ull << 30))
    return fail(APHID_E_ARG, "bad argument");
  CK(hipSetDevice(c->dev));
  const long long n = (long long)(bytes / sizeof(c128));
  c128 *a = nullptr, *b = nullptr;
  CK(hipMalloc(&a, sizeof(c128) * n));
  hipError_t e = hipMalloc(&b, sizeof(c128) * n);
  if (e != hipSuccess) {
    hipFree(a);
    return fail(APHID_E_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  int rc = 0;
  float ms = 0.f;
  do {
    if ((e = hipMemsetAsync(a, 0, sizeof(c128) * n, c->s)) != hipSuccess) break;
    hipLaunchKernelGGL(k_copy, dim3(16384), dim3(kBlock), 0, c->s, a, b, n);  // warm
    if ((e = hipEventRecord(c->ev0, c->s)) != hipSuccess) break;
    for (int r = 0; r < reps; ++r)
      hipLaunchKernelGGL(k_copy, dim3(16384), dim3(kBlock), 0, c->s, a, b, n);
    if ((e = hipEventRecord(c->ev1, c->s)) != hipSuccess) break;
    if ((e = hipEventSynchronize(c->ev1)) != hipSuccess) break;
    e = hipEventElapsedTime(&ms, c->ev0, c->ev1);
  } while (0);
  hipFree(a);
  hipFree(b);
  if (e != hipSuccess) rc = fail(APHID_E_HIP, std::string("copy probe: ") + hipGetErrorString(e));
  *seconds = 1e-3 * ms / reps;
  return rc;
}

}  // extern "C"
