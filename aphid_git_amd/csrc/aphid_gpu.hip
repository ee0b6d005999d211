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
#include <map>
#include <mutex>
#include <random>
#include <string>
#include <tuple>
#include <vector>

#include "aphid_gpu.h"
#include "lib_internal.hpp"
#include "qmr.hpp"
#include "setup.hpp"
#include "stencil.hpp"
#include "warm.hpp"
#include "fields.hpp"

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
  int nt = 0;         // march: non-temporal stores of the output (the streams
                      // tuner decides this inside the iteration's sequence)
};
// Launch shape of a streaming iteration kernel: elements unrolled per thread
// and step, grid size (= number of reduction partials) and the
// non-temporal hint.
struct Stream {
  int u, grid, nt;
};
enum { SK_PUPDATE = 0, SK_VTNORM = 1, SK_UPDATE = 2, SK_COUNT = 3 };
constexpr Stream kStreamDefault[SK_COUNT] = {{4, 2048, 0}, {4, 2048, 0}, {2, 2048, 0}};
const Tile kTiles[] = {{16, 16, 0}, {32, 8, 0}, {64, 4, 0}, {16, 8, 0}, {32, 4, 0},
                       {32, 16, 0}, {64, 8, 0}, {16, 32, 0}};

// the x-march stencil of one tile: store hint (nt), reproducible eps
// partials (ep.seg) and, for complex128, the compact k^2 (op.k2rec; not with
// the segment partials, whose registers it would spill)
template <typename R, int TK, int TJ>
hipError_t launch_march(dim3 grid, dim3 block, hipStream_t s, const OpArgs<R>& op,
                        const cplx<R>* in, cplx<R>* out, int chunk, ApplyEpilogue ep, bool nt) {
#define APHID_MARCH(NT, SEG, CKV)                                                            \
  hipLaunchKernelGGL((k_apply<R, TK, TJ, NT, SEG, CKV>), grid, block, 0, s, op, in, out, chunk, ep)
  if constexpr (std::is_same<R, double>::value) {
    const bool ck = op.k2rec != nullptr;
    if (ep.seg) {
      if (nt) APHID_MARCH(true, true, false); else APHID_MARCH(false, true, false);
    } else if (ck) {
      if (nt) APHID_MARCH(true, false, true); else APHID_MARCH(false, false, true);
    } else {
      if (nt) APHID_MARCH(true, false, false); else APHID_MARCH(false, false, false);
    }
  } else {
    if (nt) APHID_MARCH(true, false, false); else APHID_MARCH(false, false, false);
  }
#undef APHID_MARCH
  return hipGetLastError();
}

template <typename R>
hipError_t launch_apply(const Tile& t, const OpArgs<R>& op, const cplx<R>* in, cplx<R>* out,
                        ApplyEpilogue ep, hipStream_t s) {
  const int ntk = (op.nz + t.tk - 1) / t.tk, ntj = (op.ny + t.tj - 1) / t.tj;
  dim3 block(t.tk * t.tj);
  // reproducible eps partials: complex128 x-march with chunks that are
  // multiples of 4 only
  if (ep.seg && (!std::is_same<R, double>::value || t.chunk <= 0 || t.chunk % 4 != 0))
    return hipErrorInvalidValue;
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
    APHID_TILE(32, 16)
    APHID_TILE(64, 8)
    APHID_TILE(16, 32)
#undef APHID_TILE
    return hipErrorInvalidValue;
  }
  const int chunk = t.chunk;
  dim3 grid((unsigned)ntk * ntj * ((op.nx + chunk - 1) / chunk));
#define APHID_TILE(TK, TJ)                                                                 \
  if (t.tk == TK && t.tj == TJ)                                                            \
    return launch_march<R, TK, TJ>(grid, block, s, op, in, out, chunk, ep, t.nt != 0);
  APHID_TILE(16, 16)
  APHID_TILE(32, 8)
  APHID_TILE(64, 4)
  APHID_TILE(16, 8)
  APHID_TILE(32, 4)
  APHID_TILE(32, 16)
  APHID_TILE(64, 8)
  APHID_TILE(16, 32)
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

int aphid_fail(int code, const std::string& msg) { return fail(code, msg); }

// ---------------------------------------------------------------------------
// Scratch of aphid_apply_noalloc: one file-scope device array per GPU (the
// device copies of the input and the output vector, 4*nn complex128 each,
// and the nn-value staging of the layout permutation: 9*nn*16 bytes), held
// from the code object's load until the process ends.  Host code never
// passes `scratch` itself to a kernel (in host code the name is only a
// host-side stand-in): aphid_ctx_create looks up its device address once
// per device with hipGetSymbolAddress, and every launch and copy uses that
// address.
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_scratch_mu;
std::map<int, void*> g_scratch_ptr;  // device -> address of `scratch` there

// the device address of `scratch` on the current device `dev` (once per
// device; later calls return the stored address)
int scratch_address(int dev, void** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  const auto it = g_scratch_ptr.find(dev);
  if (it != g_scratch_ptr.end()) {
    *out = it->second;
    return 0;
  }
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(APHID_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[dev] = p;
  *out = p;
  return 0;
}
}  // namespace

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
  // compact k^2 of the complex128 stencil (OpArgs::k2rec / k2re / k2pc):
  // 2*nn pairs of doubles, the tables rx (ny*nz), ry (nx*nz), rz (nx*ny),
  // then nn k2p codes (bytes) from the bit pattern k2c_base; used when
  // k2c_ok (checked against the grids) and k2c_on (the caller's switch,
  // aphid_set_compact_k2)
  double* k2c = nullptr;
  int* k2c_flag = nullptr;
  bool k2c_ok = false, k2c_on = true;
  unsigned long long k2c_base = 0;
  // Krylov vectors, complex128 and complex64 sets
  c128* V[9] = {};  // x b v p t d s r spare
  bool vborrowed = false;  // V[] are slices of a batch's flat vectors (not owned)
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
  const void* gvec[2] = {nullptr, nullptr};  // x vector the graph was captured with
  // stencil tiles per precision
  Tile tile[2] = {{16, 16, 8}, {16, 16, 8}};
  bool tuned[2] = {false, false};
  // launch shapes of the streaming iteration kernels per precision
  Stream stream[2][SK_COUNT] = {{kStreamDefault[0], kStreamDefault[1], kStreamDefault[2]},
                                {kStreamDefault[0], kStreamDefault[1], kStreamDefault[2]}};
  bool stuned[2] = {false, false};
  QmrState* tst = nullptr;  // scratch QMR state for timing the streaming kernels
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // stagnation stop of a solve (used for the complex64 inner solves of the
  // mixed-precision refinement): stop when the residual has not fallen below
  // stall_factor times its best value for stall_window iterations
  int stall_window = 0;
  double stall_factor = 0.9;
  bool stalled = false;
  // wall-clock marks of the running solve: (iterations done, seconds since
  // the solve started) after every host read of the device state
  std::chrono::steady_clock::time_point t_solve;
  std::vector<std::pair<int, double>> marks;
  void* scratch = nullptr;  // device address of `scratch` on this context's device
  // reproducible complex128 solves (aphid_set_reproducible): eps from the
  // stencil's (column, 4-plane segment) partials, seg_n pairs
  bool repro = false;
  double* seg = nullptr;
  long long seg_n = 0;
};

namespace {

void free_mesh(aphid_ctx* c) {
  for (int p = 0; p < 2; ++p) {
    if (c->gexec[p]) (void)hipGraphExecDestroy(c->gexec[p]);
    c->gexec[p] = nullptr;
    c->giters[p] = 0;
  }
  void* ptrs[] = {c->tab64, c->tab32, c->dmet, c->sig, c->k2, c->k2f, c->diag, c->invd, c->invdf,
                  c->stage, c->stagef, c->hist, c->k2c, c->k2c_flag};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (auto& v : c->V) {
    if (v && !c->vborrowed) (void)hipFree(v);
    v = nullptr;
  }
  c->vborrowed = false;
  for (auto& v : c->F)
    if (v) (void)hipFree(v), v = nullptr;
  c->tab64 = nullptr; c->tab32 = nullptr; c->dmet = nullptr; c->sig = nullptr; c->k2 = nullptr;
  c->k2f = nullptr; c->diag = nullptr; c->invd = nullptr; c->invdf = nullptr; c->stage = nullptr;
  c->stagef = nullptr; c->hist = nullptr; c->hist_cap = 0;
  c->k2c = nullptr; c->k2c_flag = nullptr; c->k2c_ok = false;
  c->op64.k2rec = nullptr;
  for (auto& t : c->op64.k2re) t = nullptr;
  c->op64.k2pc = nullptr;
  c->op64.k2p_base = 0;
  if (c->seg) (void)hipFree(c->seg);
  c->seg = nullptr;
  c->seg_n = 0;
  c->have_k2 = c->have_diag = false;
  c->tuned[0] = c->tuned[1] = false;
  c->stuned[0] = c->stuned[1] = false;
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

// host Fortran-order component arrays -> device vector (4 blocks of nn),
// through a device staging buffer of nn complex values (default c->stage)
int upload4(aphid_ctx* c, const double* const h[4], c128* dst, c128* stage = nullptr) {
  if (!stage) stage = c->stage;
  for (int comp = 0; comp < 4; ++comp) {
    int n1, n2, n3;
    comp_shape(c, comp, n1, n2, n3);
    const size_t bytes = sizeof(c128) * (size_t)n1 * n2 * n3;
    CK(hipMemcpyAsync(stage, h[comp], bytes, hipMemcpyHostToDevice, c->s));
    hipLaunchKernelGGL(k_f2d<double>, dim3(grid_for(c->nn)), dim3(kBlock), 0, c->s, stage, n1,
                       n2, n3, dst + comp * c->nn, dims(c));
    CK(hipGetLastError());
  }
  return 0;
}

int download4(aphid_ctx* c, const c128* src, double* const h[4], c128* stage = nullptr) {
  if (!stage) stage = c->stage;
  for (int comp = 0; comp < 4; ++comp) {
    int n1, n2, n3;
    comp_shape(c, comp, n1, n2, n3);
    const long long n = (long long)n1 * n2 * n3;
    hipLaunchKernelGGL(k_d2f<double>, dim3(grid_for(n)), dim3(kBlock), 0, c->s,
                       src + comp * c->nn, n1, n2, n3, stage, dims(c));
    CK(hipGetLastError());
    CK(hipMemcpyAsync(h[comp], stage, sizeof(c128) * (size_t)n, hipMemcpyDeviceToHost, c->s));
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

// Points the complex128 operator at the compact k^2 when it is valid and
// switched on; a captured iteration graph holds the old kernel arguments,
// so it is dropped when the choice changes.
void select_k2c(aphid_ctx* c) {
  const bool use = c->k2c_ok && c->k2c_on;
  const bool same = use ? c->op64.k2rec != nullptr && c->op64.k2p_base == c->k2c_base
                        : c->op64.k2rec == nullptr;
  if (same) return;
  if (c->gexec[0]) (void)hipGraphExecDestroy(c->gexec[0]), c->gexec[0] = nullptr;
  if (use) {
    c->op64.k2rec = c->k2c;
    c->op64.k2yz_off = (unsigned)(2 * c->nn * sizeof(double));
    c->op64.k2re[0] = c->k2c + 4 * c->nn;
    c->op64.k2re[1] = c->op64.k2re[0] + (size_t)c->ny * c->nz;
    c->op64.k2re[2] = c->op64.k2re[1] + (size_t)c->nx * c->nz;
    c->op64.k2pc = reinterpret_cast<const unsigned char*>(c->op64.k2re[2] + (size_t)c->nx * c->ny);
    c->op64.k2p_base = c->k2c_base;
  } else {
    c->op64.k2rec = nullptr;
    for (auto& t : c->op64.k2re) t = nullptr;
    c->op64.k2pc = nullptr;
    c->op64.k2p_base = 0;
  }
}

// the compact k^2 from the current grids, and whether it holds them exactly
// (the codes' base: 128 below the bit pattern of k2p's real part at the
// first node)
int build_k2c(aphid_ctx* c) {
  const size_t nd = 4 * (size_t)c->nn + (size_t)c->ny * c->nz + (size_t)c->nx * c->nz +
                    (size_t)c->nx * c->ny;
  if (!c->k2c) CK(hipMalloc(&c->k2c, sizeof(double) * nd + (size_t)c->nn));
  if (!c->k2c_flag) CK(hipMalloc(&c->k2c_flag, sizeof(int)));
  double* rx = c->k2c + 4 * c->nn;
  double* ry = rx + (size_t)c->ny * c->nz;
  double* rz = ry + (size_t)c->nx * c->nz;
  unsigned char* pc = reinterpret_cast<unsigned char*>(c->k2c + nd);
  double p0 = 0.0;
  CK(hipMemcpyAsync(&p0, &c->k2[3 * c->nn].re, sizeof(double), hipMemcpyDeviceToHost, c->s));
  CK(hipStreamSynchronize(c->s));
  unsigned long long b0;
  std::memcpy(&b0, &p0, sizeof(b0));
  c->k2c_base = b0 - 128ull;
  CK(hipMemsetAsync(c->k2c_flag, 0, sizeof(int), c->s));
  hipLaunchKernelGGL(k_k2c_extract, dim3(grid_for(c->nn)), dim3(kBlock), 0, c->s, c->k2, c->k2c,
                     c->k2c + 2 * c->nn, rx, ry, rz, pc, c->k2c_base, dims(c));
  hipLaunchKernelGGL(k_k2c_check, dim3(grid_for(c->nn)), dim3(kBlock), 0, c->s, c->k2, rx, ry, rz,
                     c->k2c_base, dims(c), c->k2c_flag);
  CK(hipGetLastError());
  int bad = 1;
  CK(hipMemcpyAsync(&bad, c->k2c_flag, sizeof(int), hipMemcpyDeviceToHost, c->s));
  CK(hipStreamSynchronize(c->s));
  c->k2c_ok = bad == 0;
  select_k2c(c);
  return 0;
}

// derived operator data: complex64 copies (k2, 1/diag) and the compact
// complex128 k^2
int refresh_f32(aphid_ctx* c) {
  const long long n4 = 4 * c->nn;
  hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_for(n4)), dim3(kBlock), 0, c->s, c->k2,
                     c->k2f, n4);
  hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_for(n4)), dim3(kBlock), 0, c->s,
                     c->invd, c->invdf, n4);
  CK(hipGetLastError());
  return build_k2c(c);
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

// Process-wide memo of tuned launch parameters per (device, mesh shape,
// precision): contexts and batch operators on a shape already tuned in
// this process skip the timing.
struct TuneKey {
  int dev, nx, ny, nz, prec;
  bool operator<(const TuneKey& o) const {
    return std::tie(dev, nx, ny, nz, prec) < std::tie(o.dev, o.nx, o.ny, o.nz, o.prec);
  }
};
std::mutex g_tune_mu;
std::map<TuneKey, Tile> g_tiles;
struct StreamSet {
  Stream sh[SK_COUNT];
  int snt = 0;  // the stencil's non-temporal output stores (Tile::nt)
};
std::map<TuneKey, StreamSet> g_streams;

// A stencil launch shape the kernels support on this context's mesh: one of
// kTiles, chunk >= 0, and at most kMaxPartBlocks blocks (the capacity of
// the eps partial-sum buffer).
bool tile_ok(const aphid_ctx* c, int tk, int tj, int chunk) {
  if (chunk < 0) return false;
  bool ok = false;
  for (const Tile& t : kTiles) ok |= (t.tk == tk && t.tj == tj);
  return ok && (!c->nn || apply_blocks(Tile{tk, tj, chunk}, c->nx, c->ny, c->nz) <=
                              (unsigned)kMaxPartBlocks);
}

// Autotuner: time every (tile, chunk) candidate for this mesh and keep the
// fastest.  Results are cached per mesh shape and precision in the file
// named by APHID_TUNE_CACHE (if set), and in memory for the process.
int autotune(aphid_ctx* c, int prec) {
  if (c->tuned[prec]) return 0;
  // a reproducible complex128 context picks among x-march tiles with chunks
  // that are multiples of 4 (its eps partials are per 4-plane segment), and
  // keeps its choice apart from the plain one
  const bool seg = prec == 0 && c->repro;
  const int pk = prec + (seg ? 2 : 0);
  const TuneKey tk0{c->dev, c->nx, c->ny, c->nz, pk};
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    const auto it = g_tiles.find(tk0);
    if (it != g_tiles.end()) {
      c->tile[prec] = it->second;
      c->tuned[prec] = true;
      return 0;
    }
  }
  char key[96];
  snprintf(key, sizeof key, "%d %d %d %d", c->nx, c->ny, c->nz, pk);
  const char* cache = getenv("APHID_TUNE_CACHE");
  if (cache) {
    if (FILE* f = fopen(cache, "r")) {
      char line[160];
      while (fgets(line, sizeof line, f)) {
        int a, b, cc, p, tk, tj, ch;
        if (sscanf(line, "%d %d %d %d %d %d %d", &a, &b, &cc, &p, &tk, &tj, &ch) == 7) {
          char k2[96];
          snprintf(k2, sizeof k2, "%d %d %d %d", a, b, cc, p);
          // the checks of aphid_set_tile: a line that fails them (stale or
          // edited) is ignored, and the tile is timed instead
          if (!strcmp(k2, key) && tile_ok(c, tk, tj, ch) && (!seg || (ch > 0 && ch % 4 == 0))) {
            c->tile[prec] = {tk, tj, ch};
            c->tuned[prec] = true;
          }
        }
      }
      fclose(f);
      if (c->tuned[prec]) {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tiles[tk0] = c->tile[prec];
        return 0;
      }
    }
  }
  int rc = ensure_vectors(c, prec == 1);
  if (rc) return rc;
  const int chunks[] = {0, 2, 4, 6, 8, 12, 16, 24, 32};  // 0 = direct variant
  std::vector<std::pair<Tile, float>> timed;
  float best = 1e30f;
  const int reps = c->nn > 2000000 ? 5 : 20;
  for (const Tile& base : kTiles) {
    for (int ch : chunks) {
      Tile t = {base.tk, base.tj, ch};
      if (apply_blocks(t, c->nx, c->ny, c->nz) > (unsigned)kMaxPartBlocks) continue;
      if (seg && (ch == 0 || ch % 4 != 0)) continue;
      float ms = 0.f;
      if (prec == 0)
        rc = time_tile<double>(c, t, c->op64, c->V[3], c->V[4], reps, &ms);
      else
        rc = time_tile<float>(c, t, c->op32, c->F[3], c->F[4], reps, &ms);
      if (rc) return rc;
      timed.push_back({t, ms});
      best = std::min(best, ms);
    }
  }
  // The tile fixes the order of the eps partial sums, and with it the last
  // bits of the solve.  So that timing noise does not flip a near-tie (and
  // with it the iteration count) from run to run, the choice among the
  // tiles within 3% of the fastest is the one with the fewest blocks (the
  // fewest halo and chunk-boundary re-reads), then the faster one.
  Tile bt = c->tile[prec];
  unsigned bt_blocks = ~0u;
  float bt_ms = 1e30f;
  for (const auto& tm : timed) {
    if (tm.second > 1.03f * best) continue;
    const unsigned nb = apply_blocks(tm.first, c->nx, c->ny, c->nz);
    if (nb < bt_blocks || (nb == bt_blocks && tm.second < bt_ms)) {
      bt = tm.first;
      bt_blocks = nb;
      bt_ms = tm.second;
    }
  }
  c->tile[prec] = bt;
  c->tuned[prec] = true;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tiles[tk0] = bt;
  }
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

// reproducible solves: the eps partials of every (column, 4-plane segment)
int ensure_seg(aphid_ctx* c) {
  if (!c->repro || c->seg) return 0;
  c->seg_n = (long long)((c->nx + 3) / 4) * c->ny * c->nz;
  CK(hipMalloc(&c->seg, sizeof(double) * 2 * c->seg_n));
  return 0;
}

// The stencil with the eps = p^T K p epilogue and its finalize on state st
// (phase B of the QMR iteration): per-block partials summed by one block,
// or, for a reproducible complex128 solve, the stencil's per-segment
// partials summed by a fixed grid (k_seg_reduce) and then by one block.
// `mid` (optional): an event recorded right after the stencil.
constexpr int kSegBlocks = 512;
template <typename V>
hipError_t launch_stencil_eps(aphid_ctx* c, const Tile& tl, const cplx<V>* p, cplx<V>* t,
                              QmrState* st, hipEvent_t mid = nullptr) {
  if (std::is_same<V, double>::value && c->repro) {
    ApplyEpilogue ep{st, c->part, c->seg};
    const hipError_t e = launch_apply<V>(tl, opargs<V>(c), p, t, ep, c->s);
    if (e != hipSuccess) return e;
    if (mid) (void)hipEventRecord(mid, c->s);
    const long long per = (c->seg_n + kSegBlocks - 1) / kSegBlocks;
    hipLaunchKernelGGL(k_seg_reduce, dim3(kSegBlocks), dim3(kBlock), 0, c->s, c->seg, c->seg_n,
                       per, st, c->part);
    hipLaunchKernelGGL(k_fin_eps, dim3(1), dim3(kBlock), 0, c->s, c->part, kSegBlocks, st);
    return hipGetLastError();
  }
  ApplyEpilogue ep{st, c->part};
  const hipError_t e = launch_apply<V>(tl, opargs<V>(c), p, t, ep, c->s);
  if (mid) (void)hipEventRecord(mid, c->s);
  hipLaunchKernelGGL(k_fin_eps, dim3(1), dim3(kBlock), 0, c->s, c->part,
                     (int)apply_blocks(tl, c->nx, c->ny, c->nz), st);
  return e;
}


template <typename V>
const cplx<V>* invd_of(aphid_ctx* c);
template <>
const cplx<double>* invd_of<double>(aphid_ctx* c) { return c->invd; }
template <>
const cplx<float>* invd_of<float>(aphid_ctx* c) { return c->invdf; }

// The three streaming kernels of an iteration, each with its finalize, in
// the launch shape `sh` (the grid is also the number of partials).
#define APHID_STREAM_SHAPES(X) X(2, false) X(2, true) X(4, false) X(4, true)

// the p-update with the d and x updates of the previous iteration
template <typename V>
hipError_t launch_pupdate(const Stream& sh, const Vecs<V>& w, const cplx<V>* invd, long long n,
                          QmrState* st, hipStream_t s) {
#define X(U, NT)                                                                              \
  if (sh.u == U && (sh.nt != 0) == NT) {                                                      \
    hipLaunchKernelGGL((k_pupdate<V, U, NT>), dim3(sh.grid), dim3(kBlock), 0, s, w.v, invd, w.p, \
                       w.d, w.x, n, st);                                                      \
    return hipGetLastError();                                                                 \
  }
  APHID_STREAM_SHAPES(X)
#undef X
  return hipErrorInvalidValue;
}

// `mid` (optional): an event recorded between the kernel and its finalize
template <typename V>
hipError_t launch_vtnorm(const Stream& sh, const cplx<V>* t, const cplx<V>* v, long long n,
                         QmrState* st, double* part, hipStream_t s, hipEvent_t mid = nullptr) {
#define X(U, NT)                                                                              \
  if (sh.u == U && (sh.nt != 0) == NT) {                                                      \
    hipLaunchKernelGGL((k_vtnorm<V, U, NT>), dim3(sh.grid), dim3(kBlock), 0, s, t, v, n, st,    \
                       part);                                                                 \
    if (mid) (void)hipEventRecord(mid, s);                                                    \
    hipLaunchKernelGGL(k_fin_vtnorm, dim3(1), dim3(kBlock), 0, s, part, sh.grid, st);          \
    return hipGetLastError();                                                                 \
  }
  APHID_STREAM_SHAPES(X)
#undef X
  return hipErrorInvalidValue;
}

template <typename V>
hipError_t launch_update(const Stream& sh, const Vecs<V>& w, const cplx<V>* invd, long long n,
                         QmrState* st, double* part, double* hist, hipStream_t s,
                         hipEvent_t mid = nullptr) {
#define X(U, NT)                                                                              \
  if (sh.u == U && (sh.nt != 0) == NT) {                                                      \
    hipLaunchKernelGGL((k_update<V, U, NT>), dim3(sh.grid), dim3(kBlock), 0, s, w.t, w.s, w.r,  \
                       w.v, invd, n, st, part);                                               \
    if (mid) (void)hipEventRecord(mid, s);                                                    \
    hipLaunchKernelGGL(k_fin_update, dim3(1), dim3(kBlock), 0, s, part, sh.grid, st, hist);    \
    return hipGetLastError();                                                                 \
  }
  APHID_STREAM_SHAPES(X)
#undef X
  return hipErrorInvalidValue;
}

// APHID_TUNE_VERBOSE=1 prints every timed stream-shape candidate to stderr
bool tune_verbose() {
  static const bool v = [] {
    const char* e = getenv("APHID_TUNE_VERBOSE");
    return e && e[0] == '1';
  }();
  return v;
}

// Stream-shape tuner.  The streaming kernels interact through the caches:
// a kernel that reads t and v with plain loads leaves them in the L2 and
// the Infinity Cache for the next kernel, and the hint on the p-update's
// stores changes how fast the stencil then reads p.  So the tuner times the
// iteration's own kernel sequence (p-update, stencil with eps, ||t - beta
// v||, fused update), each kernel with its finalize, and picks the shapes
// by one coordinate-descent pass on the sequence total (fused update, then
// p-update, then ||t - beta v||).  The tuned set replaces the
// defaults only when the total is at least 3% lower, so that near-ties do
// not flip the launch shapes (and with them the order of the reductions)
// from run to run.  Timing runs on vectors that the solve overwrites
// before it reads them (v p t d s r and the spare, which stands in for x),
// with a scratch state reloaded before every launch.  Results are cached
// with the stencil tiles in APHID_TUNE_CACHE as lines
// "S nx ny nz prec u grid nt u grid nt u grid nt" (p-update, ||t - beta v||,
// fused update).
template <typename V>
int tune_streams(aphid_ctx* c) {
  constexpr int prec = std::is_same<V, float>::value ? 1 : 0;
  if (c->stuned[prec]) return 0;
  if (int rc = ensure_seg(c)) return rc;
  const TuneKey tk0{c->dev, c->nx, c->ny, c->nz, prec};
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    const auto it = g_streams.find(tk0);
    if (it != g_streams.end()) {
      for (int q = 0; q < SK_COUNT; ++q) c->stream[prec][q] = it->second.sh[q];
      c->tile[prec].nt = c->tile[prec].chunk > 0 ? it->second.snt : 0;
      c->stuned[prec] = true;
      return 0;
    }
  }
  int rc = autotune(c, prec);  // the tile, before its store hint is set below
  if (rc) return rc;
  char key[96];
  snprintf(key, sizeof key, "%d %d %d %d", c->nx, c->ny, c->nz, prec);
  const char* cache = getenv("APHID_TUNE_CACHE");
  if (cache) {
    if (FILE* f = fopen(cache, "r")) {
      char line[200];
      while (fgets(line, sizeof line, f)) {
        int a, b, cc, p, v[10];
        v[9] = 0;  // lines without the stencil's store hint: plain stores
        const int got = sscanf(line, "S %d %d %d %d %d %d %d %d %d %d %d %d %d %d", &a, &b, &cc,
                               &p, &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6], &v[7], &v[8],
                               &v[9]);
        if (got != 13 && got != 14) continue;
        char k2[96];
        snprintf(k2, sizeof k2, "%d %d %d %d", a, b, cc, p);
        if (strcmp(k2, key)) continue;
        bool ok = true;
        Stream sh[SK_COUNT];
        for (int q = 0; q < SK_COUNT; ++q) {
          sh[q] = {v[3 * q], v[3 * q + 1], v[3 * q + 2]};
          ok = ok && (sh[q].u == 2 || sh[q].u == 4) && sh[q].grid >= 1 &&
               sh[q].grid <= kMaxPartBlocks && (sh[q].nt == 0 || sh[q].nt == 1);
        }
        if (ok && (v[9] == 0 || v[9] == 1)) {
          for (int q = 0; q < SK_COUNT; ++q) c->stream[prec][q] = sh[q];
          c->tile[prec].nt = c->tile[prec].chunk > 0 ? v[9] : 0;
          c->stuned[prec] = true;
        }
      }
      fclose(f);
      if (c->stuned[prec]) {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        for (int q = 0; q < SK_COUNT; ++q) g_streams[tk0].sh[q] = c->stream[prec][q];
        g_streams[tk0].snt = c->tile[prec].nt;
        return 0;
      }
    }
  }
  rc = ensure_vectors(c, prec == 1);
  if (rc) return rc;
  const long long n4 = 4 * c->nn;
  Vecs<V> w = vecs<V>(c);
  w.x = reinterpret_cast<cplx<V>*>(c->V[8]);
  const cplx<V>* invd = invd_of<V>(c);
  const Tile& tl = c->tile[prec];
  QmrState ts{};
  ts.k1 = {0.5, 0.25};
  ts.beta = {1.0, 0.5};
  ts.eta = {0.5, 0.0};
  ts.eps = {1.0, 0.0};
  ts.del = {1.0, 0.0};
  ts.k2s = 0.5;
  ts.rho_np = ts.rho_n = 1.0;
  ts.c_nm = 1.0;
  ts.max_it = 1 << 30;
  ts.iter = 1;
  // an even number of sweeps: the p-update and the fused update alternate
  // between iterations that store x and r and iterations that do not
  const int reps = c->nn > 2000000 ? 4 : 8;
  auto launch = [&](int k, const Stream (&sh)[SK_COUNT]) -> hipError_t {
    switch (k) {
      case 0:
        return launch_pupdate<V>(sh[SK_PUPDATE], w, invd, n4, c->tst, c->s);
      case 1:
        return launch_stencil_eps<V>(c, tl, w.p, w.t, c->tst);
      case 2:
        return launch_vtnorm<V>(sh[SK_VTNORM], w.t, w.v, n4, c->tst, c->part, c->s);
      default:
        return launch_update<V>(sh[SK_UPDATE], w, invd, n4, c->tst, c->part, nullptr, c->s);
    }
  };
  // mean over reps of the summed per-kernel times of one iteration
  auto seq_time = [&](const Stream (&sh)[SK_COUNT], float* ms) -> int {
    float tot = 0.f;
    for (int r = 0; r <= reps; ++r)  // the first sweep warms up
      for (int k = 0; k < 4; ++k) {
        ts.iter = 2 + (r & 1);
        CK(hipMemcpyAsync(c->tst, &ts, sizeof ts, hipMemcpyHostToDevice, c->s));
        CK(hipEventRecord(c->ev0, c->s));
        CK(launch(k, sh));
        CK(hipEventRecord(c->ev1, c->s));
        CK(hipEventSynchronize(c->ev1));
        float m = 0.f;
        CK(hipEventElapsedTime(&m, c->ev0, c->ev1));
        if (r) tot += m;
      }
    *ms = tot / reps;
    return 0;
  };
  Stream cur[SK_COUNT] = {kStreamDefault[0], kStreamDefault[1], kStreamDefault[2]};
  float cur_ms = 0.f;
  rc = seq_time(cur, &cur_ms);
  if (rc) return rc;
  const float def_ms = cur_ms;
  // The grids of the two reducing kernels stay at the default: the grid
  // fixes the order of their partial sums (unroll and hint do not: every
  // thread adds its elements in index order), so the results do not depend
  // on what the timing picks.  The p-update has no reduction.
  const int grids[] = {2048, 4096, 8192, 16384};
  const long long max_grid = std::max<long long>(2048, (n4 + kBlock - 1) / kBlock);
  for (int q : {SK_UPDATE, SK_PUPDATE, SK_VTNORM}) {
    for (int u : {2, 4})
      for (int g : grids)
        for (int nt : {0, 1}) {
          if (g > max_grid) continue;
          if (q != SK_PUPDATE && g != kStreamDefault[q].grid) continue;
          Stream trial[SK_COUNT] = {cur[0], cur[1], cur[2]};
          trial[q] = {u, g, nt};
          float ms = 0.f;
          rc = seq_time(trial, &ms);
          if (rc) return rc;
          if (tune_verbose())
            fprintf(stderr,
                    "[aphid] stream tune prec %d kernel %d: u %d grid %d nt %d: %.1f us "
                    "(current %.1f, default %.1f)\n",
                    prec, q, u, g, nt, 1e3 * ms, 1e3 * cur_ms, 1e3 * def_ms);
          if (ms < 0.99f * cur_ms) {
            cur_ms = ms;
            cur[q] = trial[q];
          }
        }
  }
  if (cur_ms < 0.97f * def_ms)
    for (int q = 0; q < SK_COUNT; ++q) c->stream[prec][q] = cur[q];
  else
    for (int q = 0; q < SK_COUNT; ++q) c->stream[prec][q] = kStreamDefault[q];
  // The march stencil's output with or without the non-temporal hint,
  // timed in the same sequence: the hint keeps the stencil's re-read planes
  // and halos in L2, but the next kernels read the output.  The hint does
  // not change any value.  Kept when the sequence is 2% faster.
  if (c->tile[prec].chunk > 0) {
    Stream fin[SK_COUNT] = {c->stream[prec][0], c->stream[prec][1], c->stream[prec][2]};
    float t0 = 0.f, t1 = 0.f;
    c->tile[prec].nt = 0;
    rc = seq_time(fin, &t0);
    if (rc) return rc;
    c->tile[prec].nt = 1;
    rc = seq_time(fin, &t1);
    if (rc) return rc;
    c->tile[prec].nt = t1 < 0.98f * t0 ? 1 : 0;
    if (const char* e = getenv("APHID_STENCIL_NT"))  // forced choice (measurements)
      c->tile[prec].nt = e[0] == '1' ? 1 : 0;
    if (tune_verbose())
      fprintf(stderr, "[aphid] stencil stores prec %d: plain %.1f us, non-temporal %.1f us\n", prec,
              1e3 * t0, 1e3 * t1);
  }
  c->stuned[prec] = true;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (int q = 0; q < SK_COUNT; ++q) g_streams[tk0].sh[q] = c->stream[prec][q];
    g_streams[tk0].snt = c->tile[prec].nt;
  }
  if (tune_verbose()) {
    fprintf(stderr, "[aphid] stream shapes prec %d (%.1f us, default %.1f):", prec, 1e3 * cur_ms,
            1e3 * def_ms);
    for (int q = 0; q < SK_COUNT; ++q)
      fprintf(stderr, " %d %d %d", c->stream[prec][q].u, c->stream[prec][q].grid,
              c->stream[prec][q].nt);
    fprintf(stderr, "\n");
  }
  if (cache)
    if (FILE* f = fopen(cache, "a")) {
      fprintf(f, "S %s", key);
      for (int q = 0; q < SK_COUNT; ++q)
        fprintf(f, " %d %d %d", c->stream[prec][q].u, c->stream[prec][q].grid,
                c->stream[prec][q].nt);
      fprintf(f, " %d\n", c->tile[prec].nt);
      fclose(f);
    }
  return 0;
}

template <typename V>
int enqueue_iteration(aphid_ctx* c) {
  constexpr int prec = std::is_same<V, float>::value ? 1 : 0;
  const Vecs<V> w = vecs<V>(c);
  const long long n4 = 4 * c->nn;
  const cplx<V>* invd = invd_of<V>(c);
  const Stream* sh = c->stream[prec];
  CK(launch_pupdate<V>(sh[SK_PUPDATE], w, invd, n4, c->st, c->s));
  CK(launch_stencil_eps<V>(c, c->tile[prec], w.p, w.t, c->st));
  CK(launch_vtnorm<V>(sh[SK_VTNORM], w.t, w.v, n4, c->st, c->part, c->s));
  CK(launch_update<V>(sh[SK_UPDATE], w, invd, n4, c->st, c->part, c->hist, c->s));
  CK(hipGetLastError());
  return 0;
}

// iterations per captured graph: about 10 ms of work per replay, so that
// the host's state read between replays (an idle gap of a few tens of
// microseconds) costs well under 1%; iterations past the stop return at once
int graph_iters(aphid_ctx* c, int prec) {
  const double bytes = 22.0 * 4.0 * (double)c->nn * (prec ? 8.0 : 16.0);
  const double t_it = std::max(bytes / 4.0e12, 25e-6);
  return std::max(4, std::min(128, (int)std::ceil(1e-2 / t_it)));
}

// APHID_NO_GRAPH=1 enqueues the iterations directly on the stream instead
// of replaying a captured graph (same kernels and launch shapes); profilers
// that intercept hipGraphLaunch need it.
bool use_graphs() {
  const char* e = getenv("APHID_NO_GRAPH");
  return !(e && e[0] == '1');
}

// host-side stagnation test after each batch of iterations (see
// aphid_ctx::stall_window); true when the solve should stop
// APHID_PROGRESS=<n>: the solve loop reports its iteration count and
// residual to stderr each time it passes a multiple of n iterations
int progress_every() {
  static const int n = [] {
    const char* e = getenv("APHID_PROGRESS");
    return e ? std::max(0, atoi(e)) : 0;
  }();
  return n;
}

void mark_time(aphid_ctx* c) {
  c->marks.push_back(
      {c->hst->iter,
       std::chrono::duration<double>(std::chrono::steady_clock::now() - c->t_solve).count()});
}

void report_progress(const aphid_ctx* c, int& next) {
  const int n = progress_every();
  if (n <= 0) return;
  if (next <= 0) next = n;
  if (c->hst->iter >= next) {
    fprintf(stderr, "[aphid] iteration %d residual %.3e (target %.3e)\n", c->hst->iter,
            c->hst->rnorm, c->hst->target);
    fflush(stderr);
    while (next <= c->hst->iter) next += n;
  }
}

struct StallTracker {
  double best = INFINITY;
  int best_it = 0;
  bool step(const aphid_ctx* c) {
    if (c->stall_window <= 0) return false;
    const double rn = c->hst->rnorm;
    const int it = c->hst->iter;
    if (rn < c->stall_factor * best) {
      best = rn;
      best_it = it;
    }
    return it - best_it >= c->stall_window;
  }
};

template <typename V>
int run_loop(aphid_ctx* c) {
  constexpr int prec = std::is_same<V, float>::value ? 1 : 0;
  c->stalled = false;
  StallTracker stall;
  int next_report = 0;
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
      mark_time(c);
      if (c->hst->done) return 0;
      report_progress(c, next_report);
      if (stall.step(c)) {
        c->stalled = true;
        return 0;
      }
      if (c->hst->iter <= last_iter || l > max_launch)
        return fail(APHID_E_STATE, "QMR state did not advance on the device");
      last_iter = c->hst->iter;
    }
  }
  const void* xv = vecs<V>(c).x;
  if (c->gexec[prec] && c->gvec[prec] != xv) {  // vectors moved since the capture
    CK(hipGraphExecDestroy(c->gexec[prec]));
    c->gexec[prec] = nullptr;
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
    c->gvec[prec] = xv;
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
    mark_time(c);
    if (c->hst->done) break;
    report_progress(c, next_report);
    if (stall.step(c)) {
      c->stalled = true;
      break;
    }
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
  int trc = ensure_seg(c);
  if (!trc) trc = tune_streams<V>(c);
  if (trc) return trc;
  const Vecs<V> w = vecs<V>(c);
  const long long n4 = 4 * c->nn;
  QmrState init{};
  init.max_it = max_it;
  init.hist_cap = c->hist_cap;
  init.done = 0;
  c->marks.clear();
  c->t_solve = std::chrono::steady_clock::now();
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
  // the loop leaves c->hst current: the last iteration's d and x updates
  // (due in the p-update that did not run) and an odd last iteration's x
  // and r (left one step behind)
  if (c->hst->upd_it > 0) {
    hipLaunchKernelGGL(k_finish<V>, dim3(kRedBlocks), dim3(kBlock), 0, c->s, w.x, w.d, w.p, w.r,
                       w.s, n4, c->st);
    CK(hipGetLastError());
  }
  CK(hipEventRecord(c->ev1, c->s));
  CK(hipEventSynchronize(c->ev1));
  float ms = 0.f;
  CK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  *secs = 1e-3 * ms;
  return 0;
}

// Wall-clock seconds of iteration q+1 for q < n: the time of the latest
// host read of the device state at or before that iteration (0 before the
// first), like the reference's t2 - t1 column, which it refreshes every 10
// iterations (src/aphid_qmr.f90:376-381).
void fill_hist_time(const std::vector<std::pair<int, double>>& marks, double* t, int n) {
  size_t m = 0;
  double last = 0.0;
  for (int q = 0; q < n; ++q) {
    while (m < marks.size() && marks[m].first <= q + 1) last = marks[m++].second;
    t[q] = last;
  }
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
  const char* sw = getenv("APHID_MIXED_STALL");
  c->stall_window = sw ? std::max(0, atoi(sw)) : 0;
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
    if (rc) {
      c->stall_window = 0;
      return rc;
    }
    last = *c->hst;
    const int executed = c->stalled ? std::max(last.iter - 1, 1)
                         : last.status == ST_MAXIT ? std::max(max_it - total, 1) : last.iter;
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
  c->stall_window = 0;
  c->stalled = false;
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
  void* sp = nullptr;
  const int src = scratch_address(device, &sp);
  if (src) return src;
  aphid_ctx* c = new aphid_ctx();
  c->dev = device;
  c->scratch = sp;
  CK(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
  CK(hipMalloc(&c->st, sizeof(QmrState)));
  CK(hipMalloc(&c->tst, sizeof(QmrState)));
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
  (void)hipSetDevice(c->dev);
  if (c->s) (void)hipStreamSynchronize(c->s);
  free_mesh(c);
  if (c->st) (void)hipFree(c->st);
  if (c->tst) (void)hipFree(c->tst);
  if (c->hst) (void)hipHostFree(c->hst);
  if (c->part) (void)hipFree(c->part);
  if (c->resbuf) (void)hipFree(c->resbuf);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->s) (void)hipStreamDestroy(c->s);
  delete c;
  return 0;
}

// The stencils address a vector of four component blocks through 32-bit
// byte offsets from its base (component * nn + node) * sizeof(complex128),
// so 4 * nn * 16 bytes must stay below 2^32: nn < APHID_MAX_NODES.
bool nodes_fit(int nx, int ny, int nz) {
  return (long long)nx * ny * nz < (long long)APHID_MAX_NODES;
}

int aphid_set_mesh(aphid_ctx* c, int nx, int ny, int nz, const double* xn, const double* yn,
                   const double* zn) {
  if (!c || !xn || !yn || !zn) return fail(APHID_E_ARG, "NULL argument");
  if (nx < 3 || ny < 3 || nz < 3) return fail(APHID_E_ARG, "need at least 3 nodes per axis");
  if (!nodes_fit(nx, ny, nz))
    return fail(APHID_E_ARG, "mesh too large: 4*nx*ny*nz*16 bytes must stay below 2^32");
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

int aphid_apply_noalloc(aphid_ctx* c, const double* ix, const double* iy, const double* iz,
                        const double* ip, double* ox, double* oy, double* oz, double* op) {
  if (!c || !ix || !iy || !iz || !ip || !ox || !oy || !oz || !op)
    return fail(APHID_E_ARG, "NULL argument");
  if (!c->have_k2) return fail(APHID_E_STATE, "set k2 (sigma, cells or k2) first");
  if (9 * (size_t)c->nn * sizeof(c128) > kScratchBytes)
    return fail(APHID_E_ARG, "mesh too large for aphid_apply_noalloc: 9*nx*ny*nz*16 bytes must "
                             "not exceed APHID_NOALLOC_SCRATCH_BYTES");
  if (!c->scratch) return fail(APHID_E_STATE, "no scratch address for this device");
  CK(hipSetDevice(c->dev));
  // every entry of the three regions is written before it is read: the
  // permutation kernel fills all 4*nn input entries (zero padding
  // included), the stencil all 4*nn outputs, and each copy its staging
  c128* din = static_cast<c128*>(c->scratch);
  c128* dout = din + 4 * c->nn;
  c128* stage = dout + 4 * c->nn;
  const double* hin[4] = {ix, iy, iz, ip};
  int rc = upload4(c, hin, din, stage);
  if (rc) return rc;
  // the tuned tile if the context has one, else its default: the apply's
  // outputs do not depend on the launch shape's eps grouping
  ApplyEpilogue ep{nullptr, nullptr};
  CK(launch_apply<double>(c->tile[0], c->op64, din, dout, ep, c->s));
  double* hout[4] = {ox, oy, oz, op};
  return download4(c, dout, hout, stage);
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
      if (g) (void)hipGraphExecDestroy(g), g = nullptr;  // graphs captured the old pointer
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
    if (rep->hist_time) fill_hist_time(c->marks, rep->hist_time, rep->hist_len);
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

namespace {

// Amplitudes of the candidates in g (Galerkin or least squares), solved on
// the host by Gaussian elimination with partial pivoting and a tiny
// relative ridge.  Returns false when the fit is singular or non-finite.
int ws_fit(aphid_ctx* c, const GramArgs& g, bool lsq, std::vector<c128>& al, bool* ok) {
  *ok = false;
  if (lsq)
    hipLaunchKernelGGL(k_gram<false>, dim3(kRedBlocks), dim3(kBlock), 0, c->s, g, c->part);
  else
    hipLaunchKernelGGL(k_gram<true>, dim3(kRedBlocks), dim3(kBlock), 0, c->s, g, c->part);
  CK(hipGetLastError());
  std::vector<double> hp((size_t)kRedBlocks * 2 * kWsPairs);
  CK(hipMemcpyAsync(hp.data(), c->part, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, c->s));
  CK(hipStreamSynchronize(c->s));
  std::vector<double> tot(2 * kWsPairs, 0.0);
  for (int b = 0; b < kRedBlocks; ++b)
    for (int q = 0; q < 2 * kWsPairs; ++q) tot[q] += hp[(size_t)b * 2 * kWsPairs + q];
  const int nb = g.nb;
  std::vector<c128> G(nb * nb), h(nb);
  al.assign(nb, c128{0.0, 0.0});
  int p = 0;
  for (int a = 0; a < kWsMax; ++a)
    for (int b = a; b < kWsMax; ++b, ++p)
      if (a < nb && b < nb) {
        G[a * nb + b] = {tot[2 * p], tot[2 * p + 1]};
        G[b * nb + a] = {tot[2 * p], lsq ? -tot[2 * p + 1] : tot[2 * p + 1]};
      }
  for (int a = 0; a < kWsMax; ++a, ++p)
    if (a < nb) h[a] = {tot[2 * p], tot[2 * p + 1]};
  auto mag = [](c128 z) { return std::hypot(z.re, z.im); };
  double dmax = 0.0;
  for (int a = 0; a < nb; ++a) dmax = std::max(dmax, mag(G[a * nb + a]));
  if (!(dmax > 0.0) || !std::isfinite(dmax)) return 0;
  for (int a = 0; a < nb; ++a) G[a * nb + a].re += 1e-13 * dmax;
  for (int k = 0; k < nb; ++k) {
    int piv = k;
    for (int i = k + 1; i < nb; ++i)
      if (mag(G[i * nb + k]) > mag(G[piv * nb + k])) piv = i;
    if (!(mag(G[piv * nb + k]) > 0.0)) return 0;
    if (piv != k) {
      for (int j = 0; j < nb; ++j) std::swap(G[k * nb + j], G[piv * nb + j]);
      std::swap(h[k], h[piv]);
    }
    for (int i = k + 1; i < nb; ++i) {
      const c128 f = cdiv(G[i * nb + k], G[k * nb + k]);
      for (int j = k; j < nb; ++j) {
        const c128 t = f * G[k * nb + j];
        G[i * nb + j] = {G[i * nb + j].re - t.re, G[i * nb + j].im - t.im};
      }
      const c128 t = f * h[k];
      h[i] = {h[i].re - t.re, h[i].im - t.im};
    }
  }
  for (int i = nb - 1; i >= 0; --i) {
    c128 sacc = h[i];
    for (int j = i + 1; j < nb; ++j) {
      const c128 t = G[i * nb + j] * al[j];
      sacc = {sacc.re - t.re, sacc.im - t.im};
    }
    al[i] = cdiv(sacc, G[i * nb + i]);
    if (!finite_c(al[i])) return 0;
  }
  *ok = true;
  return 0;
}

// The network's output `out` (device, [cells][8]) mapped onto the unknowns
// (u, in V[7]) and its four components u_c = P_c u fitted by least squares
// on the residual, weighted (jacobi: W = 1/diag, the M^-1 of QMR) or not
// (W = 1):  al = argmin || W (b - sum_c al_c K P_c u) ||.
// Leaves w_c = W K P_c u in V[2..5] and W b in V[6]; uses V[8].
int net_wfit(aphid_ctx* c, const float* out, bool jacobi, std::vector<c128>& al, bool* ok) {
  const long long nn = c->nn, n4 = 4 * nn;
  const ApplyEpilogue ep{nullptr, nullptr};
  hipLaunchKernelGGL(k_cells_to_unknowns, dim3(grid_for(n4)), dim3(kBlock), 0, c->s, c->nx, c->ny,
                     c->nz, out, c->V[7]);
  CK(hipGetLastError());
  GramArgs g{};
  for (int q = 0; q < 4; ++q) {
    CK(hipMemsetAsync(c->V[8], 0, sizeof(c128) * n4, c->s));
    CK(hipMemcpyAsync(c->V[8] + q * nn, c->V[7] + q * nn, sizeof(c128) * nn,
                      hipMemcpyDeviceToDevice, c->s));
    CK(launch_apply<double>(c->tile[0], c->op64, c->V[8], c->V[2 + q], ep, c->s));
    if (jacobi)
      hipLaunchKernelGGL(k_scale_diag, dim3(grid_for(n4)), dim3(kBlock), 0, c->s, n4, c->invd,
                         c->V[2 + q]);
    CK(hipGetLastError());
    g.w[q] = c->V[2 + q];
    g.u[q] = c->V[7];
    g.ublk[q] = q;
  }
  CK(hipMemcpyAsync(c->V[6], c->V[1], sizeof(c128) * n4, hipMemcpyDeviceToDevice, c->s));
  if (jacobi)
    hipLaunchKernelGGL(k_scale_diag, dim3(grid_for(n4)), dim3(kBlock), 0, c->s, n4, c->invd,
                       c->V[6]);
  CK(hipGetLastError());
  g.nb = 4;
  g.b = c->V[6];
  g.n = n4;
  g.nn = nn;
  return ws_fit(c, g, true, al, ok);
}

// The residual loss of the network's output `out` on the context's operator
// and right-hand side, L = ||W (b - K x0)||^2 / ||W b||^2 (W = 1/diag if
// jacobi, else 1) with x0 the least-squares fit of its mapped components in
// the same norm (net_wfit), and
// scale * dL/dout into dout ([cells][8], FP32, overwritten).  *ok is false
// (and dout untouched) when the fit is singular or non-finite.
int residual_out_grad(aphid_ctx* c, const float* out, bool jacobi, float* dout, double scale,
                      double* loss, bool* ok) {
  const long long n4 = 4 * c->nn;
  const long long nc = (long long)(c->nx - 1) * (c->ny - 1) * (c->nz - 1);
  std::vector<c128> al;
  int rc = net_wfit(c, out, jacobi, al, ok);
  if (rc || !*ok) return rc;
  CombineArgs cb{};
  for (int q = 0; q < 4; ++q) {
    cb.u[q] = c->V[2 + q];
    cb.al[q] = al[q];
  }
  cb.nb = 4;
  hipLaunchKernelGGL(k_wres, dim3(kRedBlocks), dim3(kBlock), 0, c->s, n4, cb,
                     (const c128*)c->V[6], jacobi ? (const c128*)c->invd : nullptr, c->V[8],
                     c->part);
  CK(hipGetLastError());
  std::vector<double> hp(2 * (size_t)kRedBlocks);
  CK(hipMemcpyAsync(hp.data(), c->part, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, c->s));
  CK(hipStreamSynchronize(c->s));
  double rr = 0.0, bb = 0.0;
  for (int b = 0; b < kRedBlocks; ++b) rr += hp[2 * b], bb += hp[2 * b + 1];
  if (!(bb > 0.0) || !std::isfinite(rr)) {
    *ok = false;
    return 0;
  }
  const ApplyEpilogue ep{nullptr, nullptr};
  CK(launch_apply<double>(c->tile[0], c->op64, c->V[8], c->V[0], ep, c->s));  // z = K W conj(r)
  hipLaunchKernelGGL(k_adj_cells, dim3(grid_for(nc)), dim3(kBlock), 0, c->s, c->nx, c->ny, c->nz,
                     (const c128*)c->V[0], cb, 2.0 * scale / bb, dout);
  CK(hipGetLastError());
  CK(hipStreamSynchronize(c->s));
  *loss = rr / bb;
  return 0;
}

}  // namespace

// Initial guess for aphid_solve_resident(..., APHID_F_X0).  The candidate
// vectors are the network's four components (each a block of its mapped
// prediction) and/or the current solution; their amplitudes come from
//  - the Galerkin condition u_a^T (b - K x0) = 0 for every candidate u_a
//    (default: exact for an exact guess, and insensitive to small-scale
//    error that the residual norm amplifies), kept only when the true
//    residual ||b - K x0|| does not exceed ||b||, otherwise
//  - least squares on ||b - K x0|| (APHID_WS_LSQ forces it), which never
//    exceeds ||b|| (x0 = 0 is among its candidates).
// A singular or non-finite fit gives x0 = 0.
int aphid_warm_start(aphid_ctx* c, aphid_unet* net, const float* feat, unsigned flags,
                     double* r0_rel) {
  if (!c) return fail(APHID_E_ARG, "NULL context");
  if (!c->have_k2 || !c->have_diag) return fail(APHID_E_STATE, "operator not set up");
  if (!c->V[1]) return fail(APHID_E_STATE, "aphid_load_rhs first");
  if (net && !feat) return fail(APHID_E_ARG, "aphid_warm_start: features required with a network");
  const bool use_prev = flags & APHID_WS_PREV;
  if ((flags & APHID_WS_JLSQ) && (!net || use_prev))
    return fail(APHID_E_ARG, "aphid_warm_start: APHID_WS_JLSQ takes a network and no other candidate");
  CK(hipSetDevice(c->dev));
  int rc = autotune(c, 0);
  if (rc) return rc;
  const long long nn = c->nn, n4 = 4 * nn;
  if (flags & APHID_WS_JLSQ) {
    const float* out = nullptr;
    int cx = 0, cy = 0, cz = 0;
    rc = unet_internal::forward_host(net, feat, &out, &cx, &cy, &cz);
    if (rc) return rc;
    if (cx != c->nx - 1 || cy != c->ny - 1 || cz != c->nz - 1)
      return fail(APHID_E_ARG, "aphid_warm_start: network grid differs from the mesh's cells");
    std::vector<c128> al;
    bool ok = false;
    rc = net_wfit(c, out, true, al, &ok);
    if (rc) return rc;
    double rel = 1.0, rn = 0.0, bn = 0.0;
    if (ok) {
      WsCoef cf{};
      cf.use_net = 1;
      for (int k = 0; k < 4; ++k) cf.a[k] = al[k];
      hipLaunchKernelGGL(k_ws_combine, dim3(grid_for(n4)), dim3(kBlock), 0, c->s, nn,
                         (const c128*)c->V[7], (const c128*)c->V[6], c->V[0], cf);
      CK(hipGetLastError());
      rc = true_residual(c, &rel, &rn, &bn);
      if (rc) return rc;
    }
    if (!ok || !std::isfinite(rel) || rel > 1.0) {  // the judge: never worse than x0 = 0
      CK(hipMemsetAsync(c->V[0], 0, sizeof(c128) * n4, c->s));
      CK(hipStreamSynchronize(c->s));
      rel = 1.0;
    }
    if (r0_rel) *r0_rel = rel;
    return 0;
  }
  const ApplyEpilogue ep{nullptr, nullptr};
  GramArgs g{};
  g.b = c->V[1];
  g.n = n4;
  g.nn = nn;
  if (net) {
    const float* out = nullptr;
    int cx = 0, cy = 0, cz = 0;
    rc = unet_internal::forward_host(net, feat, &out, &cx, &cy, &cz);
    if (rc) return rc;
    if (cx != c->nx - 1 || cy != c->ny - 1 || cz != c->nz - 1)
      return fail(APHID_E_ARG, "aphid_warm_start: network grid differs from the mesh's cells");
    hipLaunchKernelGGL(k_cells_to_unknowns, dim3(grid_for(n4)), dim3(kBlock), 0, c->s, c->nx,
                       c->ny, c->nz, out, c->V[7]);
    CK(hipGetLastError());
    for (int q = 0; q < 4; ++q) {
      CK(hipMemsetAsync(c->V[8], 0, sizeof(c128) * n4, c->s));
      CK(hipMemcpyAsync(c->V[8] + q * nn, c->V[7] + q * nn, sizeof(c128) * nn,
                        hipMemcpyDeviceToDevice, c->s));
      CK(launch_apply<double>(c->tile[0], c->op64, c->V[8], c->V[2 + q], ep, c->s));
      g.u[g.nb] = c->V[7];
      g.ublk[g.nb] = q;
      g.w[g.nb++] = c->V[2 + q];
    }
  }
  // the network's images are in V[2..5]; the previous solution is kept in
  // V[6] and its image in V[8] (which true_residual reuses, so it is
  // recomputed before a second fit)
  if (use_prev) {
    CK(hipMemcpyAsync(c->V[6], c->V[0], sizeof(c128) * n4, hipMemcpyDeviceToDevice, c->s));
    g.u[g.nb] = c->V[6];
    g.ublk[g.nb] = -1;
    g.w[g.nb++] = c->V[8];
  }
  double rel = 1.0, rn = 0.0, bn = 0.0;
  bool lsq = (flags & APHID_WS_LSQ) != 0;
  for (int attempt = 0; attempt < 2; ++attempt) {
    WsCoef cf{};
    bool ok = g.nb == 0;
    std::vector<c128> al;
    if (use_prev) CK(launch_apply<double>(c->tile[0], c->op64, c->V[6], c->V[8], ep, c->s));
    if (g.nb > 0) {
      rc = ws_fit(c, g, lsq, al, &ok);
      if (rc) return rc;
    }
    if (ok && g.nb > 0) {
      int q = 0;
      cf.use_net = net != nullptr;
      cf.use_prev = use_prev;
      if (net)
        for (int k = 0; k < 4; ++k) cf.a[k] = al[q++];
      if (use_prev) cf.prev = al[q++];
      hipLaunchKernelGGL(k_ws_combine, dim3(grid_for(n4)), dim3(kBlock), 0, c->s, nn,
                         (const c128*)c->V[7], (const c128*)c->V[6], c->V[0], cf);
      CK(hipGetLastError());
    } else {
      CK(hipMemsetAsync(c->V[0], 0, sizeof(c128) * n4, c->s));
    }
    rc = true_residual(c, &rel, &rn, &bn);
    if (rc) return rc;
    if (std::isfinite(rel) && (rel <= 1.0 || lsq)) break;
    lsq = true;  // the judge: fall back to the least-squares amplitudes
  }
  if (!std::isfinite(rel)) {
    CK(hipMemsetAsync(c->V[0], 0, sizeof(c128) * n4, c->s));
    CK(hipStreamSynchronize(c->s));
    rel = 1.0;
  }
  if (r0_rel) *r0_rel = rel;
  return 0;
}

int aphid_warm_residual_loss_grad(aphid_ctx* c, aphid_unet* net, const float* features,
                                  unsigned flags, double* loss, float* grad) {
  if (!c || !net || !features) return fail(APHID_E_ARG, "aphid_warm_residual_loss_grad: bad argument");
  if (!c->have_k2 || !c->have_diag) return fail(APHID_E_STATE, "operator not set up");
  if (!c->V[1]) return fail(APHID_E_STATE, "aphid_load_rhs first");
  CK(hipSetDevice(c->dev));
  int rc = autotune(c, 0);
  if (rc) return rc;
  const float* out = nullptr;
  int cx = 0, cy = 0, cz = 0;
  rc = unet_internal::forward_host(net, features, &out, &cx, &cy, &cz);
  if (rc) return rc;
  if (cx != c->nx - 1 || cy != c->ny - 1 || cz != c->nz - 1)
    return fail(APHID_E_ARG, "network grid differs from the mesh's cells");
  rc = unet_internal::zero_grads(net);
  if (rc) return rc;
  double L = 0.0;
  bool ok = false;
  rc = residual_out_grad(c, out, (flags & APHID_WS_JLSQ) != 0, unet_internal::out_grad(net), 1.0,
                         &L, &ok);
  if (rc) return rc;
  if (!ok) return fail(APHID_E_STATE, "the network's guess gives a singular fit");
  rc = unet_internal::backward_input(net, grad);
  if (rc) return rc;
  if (loss) *loss = L;
  return 0;
}

int aphid_warm_train_residual(aphid_ctx* c, aphid_unet* net, const float* sigma,
                              const double* omega, unsigned flags, int steps, int batch, float lr,
                              unsigned long long seed, float* losses, double* seconds) {
  if (!c || !net || !sigma || !omega || steps < 0 || batch < 1)
    return fail(APHID_E_ARG, "aphid_warm_train_residual: bad argument");
  if (need_mesh(c)) return fail(APHID_E_STATE, "aphid_set_mesh first");
  if (!c->V[1]) return fail(APHID_E_STATE, "aphid_load_rhs first");
  const int ns = unet_internal::samples(net);
  if (ns < 1) return fail(APHID_E_STATE, "the network holds no training samples");
  const auto t0 = std::chrono::steady_clock::now();
  CK(hipSetDevice(c->dev));
  const long long nc = (long long)(c->nx - 1) * (c->ny - 1) * (c->nz - 1);
  // the operators of all samples are rebuilt on the device from their cells
  float* dsig = nullptr;
  CK(hipMalloc(&dsig, sizeof(float) * nc * ns));
  struct Free {
    float* p;
    ~Free() { if (p) (void)hipFree(p); }
  } guard{dsig};
  CK(hipMemcpyAsync(dsig, sigma, sizeof(float) * nc * ns, hipMemcpyHostToDevice, c->s));
  if (!c->sig) CK(hipMalloc(&c->sig, sizeof(c128) * nc));
  int rc = ensure_vectors(c, false);
  if (rc) return rc;
  std::mt19937_64 pick(seed);
  for (int st = 0; st < steps; ++st) {
    rc = unet_internal::zero_grads(net);
    if (rc) return rc;
    double lsum = 0.0;
    int used = 0;
    for (int bq = 0; bq < batch; ++bq) {
      const int idx = (int)(pick() % (unsigned long long)ns);
      hipLaunchKernelGGL(k_sigma, dim3(grid_for(nc)), dim3(kBlock), 0, c->s, 'R', omega[idx],
                         (const float*)(dsig + nc * idx), (const float*)nullptr,
                         (const float*)nullptr, (const float*)nullptr, c->sig, dims(c));
      CK(hipGetLastError());
      rc = build_k2(c, omega[idx]);
      if (rc) return rc;
      if (bq == 0 && st == 0) {
        rc = autotune(c, 0);
        if (rc) return rc;
      }
      const float* out = nullptr;
      rc = unet_internal::forward_sample(net, idx, &out);
      if (rc) return rc;
      double L = 0.0;
      bool ok = false;
      rc = residual_out_grad(c, out, (flags & APHID_WS_JLSQ) != 0, unet_internal::out_grad(net),
                             1.0 / batch, &L, &ok);
      if (rc) return rc;
      if (!ok) continue;  // a singular fit contributes no gradient
      rc = unet_internal::backward_sample(net, idx);
      if (rc) return rc;
      lsum += L;
      ++used;
    }
    if (used) {
      rc = unet_internal::adam_step(net, lr);
      if (rc) return rc;
    }
    if (losses) losses[st] = used ? (float)(lsum / used) : NAN;
  }
  if (seconds)
    *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

int aphid_gen_fields(aphid_ctx* c, double omega, double* bx, double* by, double* bz, double* ex,
                     double* ey, double* ez) {
  if (!c || !bx || !by || !bz || !ex || !ey || !ez) return fail(APHID_E_ARG, "NULL argument");
  if (!c->V[0] || !c->dmet) return fail(APHID_E_STATE, "no resident solution");
  CK(hipSetDevice(c->dev));
  const long long nn = c->nn;
  const double* Dx = c->dmet;
  const double* Dy = Dx + (c->nx + 1);
  const double* Dz = Dy + (c->ny + 1);
  hipLaunchKernelGGL(k_fields, dim3(grid_for(nn)), dim3(kBlock), 0, c->s, (const c128*)c->V[0],
                     dims(c), Dx, Dy, Dz, omega, c->V[7], c->V[8]);
  CK(hipGetLastError());
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  struct Out { const c128* src; double* h; int n1, n2, n3; };
  const Out outs[6] = {{c->V[7], bx, nx, ny - 1, nz - 1},     {c->V[7] + nn, by, nx - 1, ny, nz - 1},
                       {c->V[7] + 2 * nn, bz, nx - 1, ny - 1, nz}, {c->V[8], ex, nx - 1, ny, nz},
                       {c->V[8] + nn, ey, nx, ny - 1, nz},      {c->V[8] + 2 * nn, ez, nx, ny, nz - 1}};
  for (const Out& o : outs) {
    const long long n = (long long)o.n1 * o.n2 * o.n3;
    hipLaunchKernelGGL(k_d2f<double>, dim3(grid_for(n)), dim3(kBlock), 0, c->s, o.src, o.n1, o.n2,
                       o.n3, c->stage, dims(c));
    CK(hipGetLastError());
    CK(hipMemcpyAsync(o.h, c->stage, sizeof(c128) * (size_t)n, hipMemcpyDeviceToHost, c->s));
    CK(hipStreamSynchronize(c->s));
  }
  return 0;
}

int aphid_get_sigma(aphid_ctx* c, double* sig) {
  if (!c || !sig) return fail(APHID_E_ARG, "NULL argument");
  if (!c->sig) return fail(APHID_E_STATE, "aphid_set_cells or aphid_set_sigma first");
  CK(hipSetDevice(c->dev));
  const int cx = c->nx - 1, cy = c->ny - 1, cz = c->nz - 1;
  const long long n = (long long)cx * cy * cz;
  hipLaunchKernelGGL(k_d2f<double>, dim3(grid_for(n)), dim3(kBlock), 0, c->s, (const c128*)c->sig,
                     cx, cy, cz, c->stage, Dims{cx, cy, cz, n});
  CK(hipGetLastError());
  CK(hipMemcpyAsync(sig, c->stage, sizeof(c128) * (size_t)n, hipMemcpyDeviceToHost, c->s));
  CK(hipStreamSynchronize(c->s));
  return 0;
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
  if (!tile_ok(c, tk, tj, chunk)) return fail(APHID_E_ARG, "unsupported tile, or too many blocks");
  if (precision == 0 && c->repro && (chunk == 0 || chunk % 4 != 0))
    return fail(APHID_E_ARG, "a reproducible context needs an x-march chunk that is a multiple of 4");
  c->tile[precision] = {tk, tj, chunk};
  c->tuned[precision] = true;
  if (c->gexec[precision]) (void)hipGraphExecDestroy(c->gexec[precision]), c->gexec[precision] = nullptr;
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

int aphid_set_reproducible(aphid_ctx* c, int on) {
  if (!c) return fail(APHID_E_ARG, "NULL context");
  const bool r = on != 0;
  if (r == c->repro) return 0;
  CK(hipSetDevice(c->dev));
  c->repro = r;
  // the iteration graph and the tile belong to the other mode
  if (c->gexec[0]) (void)hipGraphExecDestroy(c->gexec[0]), c->gexec[0] = nullptr;
  c->giters[0] = 0;
  c->tuned[0] = false;
  c->stuned[0] = false;
  if (!r && c->seg) {
    CK(hipFree(c->seg));
    c->seg = nullptr;
    c->seg_n = 0;
  }
  return 0;
}

int aphid_set_compact_k2(aphid_ctx* c, int on, int* active) {
  if (!c) return fail(APHID_E_ARG, "NULL context");
  c->k2c_on = on != 0;
  if (c->op64.k2 && c->k2c) select_k2c(c);
  if (active) *active = c->op64.k2rec != nullptr ? 1 : 0;
  return 0;
}

int aphid_time_traffic(aphid_ctx* c, int reps, double* seconds) {
  if (!c || !seconds || reps < 1) return fail(APHID_E_ARG, "bad argument");
  if (!c->have_k2) return fail(APHID_E_STATE, "operator not set up");
  CK(hipSetDevice(c->dev));
  int rc = ensure_vectors(c, false);
  if (rc) return rc;
  rc = autotune(c, 0);
  if (rc) return rc;
  const Tile& t = c->tile[0];
  if (t.chunk <= 0) return fail(APHID_E_STATE, "the tuned stencil is the direct variant");
  const OpArgs<double>& op = c->op64;
  const int ntk = (op.nz + t.tk - 1) / t.tk, ntj = (op.ny + t.tj - 1) / t.tj;
  const dim3 grid((unsigned)ntk * ntj * ((op.nx + t.chunk - 1) / t.chunk)), block(t.tk * t.tj);
  auto launch = [&]() -> hipError_t {
    const bool ck = op.k2rec != nullptr;  // the compact k^2 loads of the stencil
#define APHID_TRAFFIC(TK, TJ, NT, CKV)                                                        \
  hipLaunchKernelGGL((k_apply_traffic<double, TK, TJ, NT, CKV>), grid, block, 0, c->s, op,     \
                     c->V[3], c->V[8], t.chunk)
#define APHID_TILE(TK, TJ)                                                                  \
    if (t.tk == TK && t.tj == TJ) {                                                         \
      if (ck) {                                                                             \
        if (t.nt) APHID_TRAFFIC(TK, TJ, true, true); else APHID_TRAFFIC(TK, TJ, false, true); \
      } else {                                                                              \
        if (t.nt) APHID_TRAFFIC(TK, TJ, true, false);                                       \
        else APHID_TRAFFIC(TK, TJ, false, false);                                           \
      }                                                                                     \
      return hipGetLastError();                                                             \
    }
    APHID_TILE(16, 16)
    APHID_TILE(32, 8)
    APHID_TILE(64, 4)
    APHID_TILE(16, 8)
    APHID_TILE(32, 4)
    APHID_TILE(32, 16)
    APHID_TILE(64, 8)
    APHID_TILE(16, 32)
#undef APHID_TILE
#undef APHID_TRAFFIC
    return hipErrorInvalidValue;
  };
  CK(launch());  // warm-up
  CK(hipEventRecord(c->ev0, c->s));
  for (int r = 0; r < reps; ++r) CK(launch());
  CK(hipEventRecord(c->ev1, c->s));
  CK(hipEventSynchronize(c->ev1));
  float ms = 0.f;
  CK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  *seconds = 1e-3 * ms / reps;
  return 0;
}

int aphid_time_iteration(aphid_ctx* c, int precision, int reps, double* secs) {
  if (!c || !secs || reps < 1 || precision != 0) return fail(APHID_E_ARG, "bad argument");
  if (!c->have_k2 || !c->V[0]) return fail(APHID_E_STATE, "solve once first");
  CK(hipSetDevice(c->dev));
  if (int rc = ensure_seg(c)) return rc;
  // continue the Krylov process of the last solve with the stop test off
  QmrState hs = *c->hst;
  hs.done = 0;
  hs.status = ST_MAXIT;
  hs.iter = 2;  // the p-update then carries the d and x updates, as in a solve
  hs.max_it = 1 << 30;
  hs.target = 0.0;
  if (!(hs.rho_np > 0.0) || !std::isfinite(hs.rho_np) || is_zero(hs.del))
    return fail(APHID_E_STATE, "the last solve left no usable Krylov state");
  CK(hipMemcpyAsync(c->st, &hs, sizeof hs, hipMemcpyHostToDevice, c->s));
  const Vecs<double> w = vecs<double>(c);
  const long long n4 = 4 * c->nn;
  const Stream* sh = c->stream[0];
  // ev (optional): events recorded between a kernel and its finalize
  auto K0 = [&]() {
    return launch_pupdate<double>(sh[SK_PUPDATE], w, c->invd, n4, c->st, c->s);
  };
  auto K1 = [&](hipEvent_t mid) {
    return launch_stencil_eps<double>(c, c->tile[0], w.p, w.t, c->st, mid);
  };
  auto K2 = [&](hipEvent_t mid) {
    return launch_vtnorm<double>(sh[SK_VTNORM], w.t, w.v, n4, c->st, c->part, c->s, mid);
  };
  auto K3 = [&](hipEvent_t mid) {
    return launch_update<double>(sh[SK_UPDATE], w, c->invd, n4, c->st, c->part, nullptr, c->s,
                                 mid);
  };
  // each kernel's own duration inside the iteration's back-to-back
  // sequence (the events between kernels mark one's end and the next's
  // start), without the one-block finalize kernels: ev[0] | p-update |
  // ev[1] | stencil | ev[2] | fin | ev[3] | vtnorm | ev[4] | fin | ev[5] |
  // update | ev[6] | fin | ev[7]
  hipEvent_t ev[8] = {};
  for (auto& e : ev) CK(hipEventCreate(&e));
  float ms[5] = {0, 0, 0, 0, 0};
  hipError_t te = hipSuccess;
  for (int rep = 0; rep < reps + 1 && te == hipSuccess; ++rep) {  // first sweep warms up
    (void)hipEventRecord(ev[0], c->s);
    te = K0();
    (void)hipEventRecord(ev[1], c->s);
    if (te == hipSuccess) te = K1(ev[2]);
    (void)hipEventRecord(ev[3], c->s);
    if (te == hipSuccess) te = K2(ev[4]);
    (void)hipEventRecord(ev[5], c->s);
    if (te == hipSuccess) te = K3(ev[6]);
    (void)hipEventRecord(ev[7], c->s);
    if (te == hipSuccess) te = hipEventSynchronize(ev[7]);
    const int from[4] = {0, 1, 3, 5}, to[4] = {1, 2, 4, 6};
    for (int q = 0; q < 4 && te == hipSuccess && rep; ++q) {
      float t = 0.f;
      te = hipEventElapsedTime(&t, ev[from[q]], ev[to[q]]);
      ms[q] += t;
    }
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  if (te != hipSuccess) return fail(APHID_E_HIP, std::string("timing: ") + hipGetErrorString(te));
  CK(hipGetLastError());
  CK(hipEventRecord(c->ev0, c->s));
  for (int rep = 0; rep < reps; ++rep) {
    CK(K0());
    CK(K1(nullptr));
    CK(K2(nullptr));
    CK(K3(nullptr));
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
    (void)hipFree(a);
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
  (void)hipFree(a);
  (void)hipFree(b);
  if (e != hipSuccess) rc = fail(APHID_E_HIP, std::string("copy probe: ") + hipGetErrorString(e));
  *seconds = 1e-3 * ms / reps;
  return rc;
}

}  // extern "C"

// ===========================================================================
// Batched solve (include/aphid_gpu.h, aphid_batch_*): many systems, one
// launch per phase.  Each operator (mesh + cells + omega) is built by an
// internal single-system context, so the batch reuses the setup kernels
// unchanged; its k^2 stays there and 1/diag is gathered into one flat array.
#include "batch.hpp"

struct aphid_batch {
  int dev = 0;
  hipStream_t s = nullptr;
  std::vector<aphid_ctx*> ops;
  std::vector<long long> opoff;  // element offset of each op in the flat 1/diag
  struct Sys {
    int op;
    long long off, n4;
    bool x0;
    c128* staged;  // b (and x0) on the device until the flat arrays exist
    // second-wave system of a sweep: its initial guess is fitted from the
    // solutions of these systems (aphid_batch_set_sweep_guess)
    std::vector<int> from;
    double r0 = 1.0;  // ||b - K x0|| / ||b|| of its initial guess (NaN: not formed)
    int hist_len = 0;                            // history entries of the last solve
    std::vector<std::pair<int, double>> marks;   // (iterations done, seconds) of that solve
  };
  std::vector<Sys> sys;
  long long n4_tot = 0, op4_tot = 0;
  // device
  c128* V[9] = {};  // x b v p t d s r spare (flat over systems)
  c128* invd = nullptr;
  OpArgs<double>* dops = nullptr;
  // complex64 inner solve (APHID_F_MIXED): e bi v p t d s r, 1/diag, operators
  c64* F[8] = {};
  c64* invdf = nullptr;
  OpArgs<float>* dops32 = nullptr;
  double *scale = nullptr, *itol = nullptr;
  int* nopen = nullptr;
  int* budget = nullptr;  // per-system inner iteration budget of the sweep
  hipGraphExec_t gexec32 = nullptr;
  int giters32 = 0;
  long long *doff = nullptr, *dn4 = nullptr, *dopoff = nullptr;
  int *dop = nullptr, *dbstart = nullptr, *dtstart = nullptr;
  int *dbsys = nullptr, *dtsys = nullptr;  // work item / tile -> system
  QmrState* st = nullptr;
  std::vector<QmrState> hst;
  double* part = nullptr;
  double* res = nullptr;
  int nstream = 0, ntile = 0;
  // batched stencil: the x-march over flat chunks (xch planes per block, an
  // LDS window of wstride nodes per array) unless a system's z-rows are too
  // long for the window (or APHID_BATCH_DIRECT=1), then the
  // one-node-per-thread kernel
  bool march = true;
  int xch = 8, wstride = 0;
  bool built = false;
  bool solved = false;
  hipGraphExec_t gexec = nullptr;
  int giters = 0;
  c128* stage = nullptr;
  long long stage_n = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // non-temporal hints of the streaming kernels per precision (p-update,
  // ||t - beta v||, fused update), chosen by timing; scratch states for it
  int nts[2][SK_COUNT] = {};
  bool ntuned[2] = {false, false};
  QmrState* tst = nullptr;
  int* dskip = nullptr;  // [nsys] systems the batched launches leave alone
  // residual history per system (aphid_batch_set_history): [nsys][hist_cap]
  // (hist_cap: entries per system the solves keep; hist_alloc: per system
  // in the allocation, which the captured iteration graph refers to)
  double* hist = nullptr;
  int hist_cap = 0, hist_alloc = 0;
  const std::vector<int>* wave = nullptr;  // systems of the wave being solved (marks)
  std::chrono::steady_clock::time_point t_solve;  // start of the batched QMR (marks)
};

namespace {

constexpr int kBT = 128;  // batched stencil block (flat in-plane chunk of 128 nodes)
// LDS window strides of the batched x-march, by the longest z-row of the
// batch (nz <= 40, 72, 136, 228); the last keeps 7 arrays of complex128 in
// 64 KB
constexpr int kBWindow[4] = {kBT + 2 * 40, kBT + 2 * 72, kBT + 2 * 136, kBT + 2 * 228};

void batch_free_device(aphid_batch* B) {
  // operator contexts that borrowed slices of the flat vectors let go of
  // them (and of graphs captured on them) before the vectors are freed
  for (aphid_ctx* c : B->ops) {
    if (!c || !c->vborrowed) continue;
    for (int p = 0; p < 2; ++p) {
      if (c->gexec[p]) (void)hipGraphExecDestroy(c->gexec[p]);
      c->gexec[p] = nullptr;
      c->gvec[p] = nullptr;
    }
    for (auto& v : c->V) v = nullptr;
    c->vborrowed = false;
  }
  if (B->gexec) (void)hipGraphExecDestroy(B->gexec), B->gexec = nullptr;
  if (B->gexec32) (void)hipGraphExecDestroy(B->gexec32), B->gexec32 = nullptr;
  for (auto& v : B->F)
    if (v) (void)hipFree(v), v = nullptr;
  void* ps32[] = {B->invdf, B->dops32, B->scale, B->itol, B->nopen, B->budget};
  for (void* p : ps32)
    if (p) (void)hipFree(p);
  B->invdf = nullptr; B->dops32 = nullptr; B->scale = nullptr; B->itol = nullptr; B->nopen = nullptr;
  B->budget = nullptr;
  for (auto& v : B->V)
    if (v) (void)hipFree(v), v = nullptr;
  void* ps[] = {B->invd, B->dops, B->doff, B->dn4, B->dopoff, B->dop, B->dbstart, B->dtstart,
                B->dbsys, B->dtsys, B->st, B->part, B->res, B->tst, B->dskip};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  B->invd = nullptr; B->dops = nullptr; B->doff = nullptr; B->dn4 = nullptr; B->dopoff = nullptr;
  B->dop = nullptr; B->dbstart = nullptr; B->dtstart = nullptr; B->st = nullptr; B->part = nullptr;
  B->dbsys = nullptr; B->dtsys = nullptr;
  B->res = nullptr; B->tst = nullptr; B->dskip = nullptr;
  if (B->hist) (void)hipFree(B->hist), B->hist = nullptr;
  B->hist_alloc = 0;
  B->ntuned[0] = B->ntuned[1] = false;
  B->built = false;
}

BatchDev batch_dev(const aphid_batch* B) {
  return BatchDev{(int)B->sys.size(), B->doff, B->dn4, B->dop, B->dopoff, B->dbstart, B->dtstart,
                  B->st, B->dbsys, B->dtsys, B->hist, B->hist_alloc};
}

// (re)build the flat device arrays after systems were added
int batch_build(aphid_batch* B) {
  if (B->built) return 0;
  const int ns = (int)B->sys.size(), no = (int)B->ops.size();
  if (B->V[1]) return fail(APHID_E_STATE, "internal: rebuild with live vectors");
  std::vector<long long> off(ns), n4(ns), opoff(no);
  std::vector<int> op(ns), bstart(ns + 1), tstart(ns + 1);
  long long o = 0;
  for (int q = 0; q < no; ++q) {
    opoff[q] = o;
    o += 4 * B->ops[q]->nn;
  }
  B->op4_tot = o;
  o = 0;
  int nb = 0, nt = 0;
  int nzmax = 0;
  for (int q = 0; q < ns; ++q) nzmax = std::max(nzmax, B->ops[B->sys[q].op]->nz);
  B->wstride = 0;
  for (int w : kBWindow)
    if (!B->wstride && w >= kBT + 2 * nzmax) B->wstride = w;
  B->march = B->wstride > 0;
  if (const char* e = getenv("APHID_BATCH_DIRECT")) B->march = B->march && e[0] != '1';
  for (int q = 0; q < ns; ++q) {
    const aphid_ctx* c = B->ops[B->sys[q].op];
    off[q] = o;
    n4[q] = 4 * c->nn;
    op[q] = B->sys[q].op;
    B->sys[q].off = o;
    B->sys[q].n4 = n4[q];
    o += n4[q];
    bstart[q] = nb;
    nb += (int)((n4[q] + kBElems - 1) / kBElems);
    tstart[q] = nt;
    const int nch = (int)((c->ny * c->nz + kBT - 1) / kBT);
    nt += nch * (B->march ? (c->nx + B->xch - 1) / B->xch : c->nx);
  }
  bstart[ns] = nb;
  tstart[ns] = nt;
  B->n4_tot = o;
  B->nstream = nb;
  B->ntile = nt;
  for (auto& v : B->V) CK(hipMalloc(&v, sizeof(c128) * B->n4_tot));
  CK(hipMalloc(&B->invd, sizeof(c128) * B->op4_tot));
  CK(hipMalloc(&B->dops, sizeof(OpArgs<double>) * no));
  CK(hipMalloc(&B->doff, sizeof(long long) * ns));
  CK(hipMalloc(&B->dn4, sizeof(long long) * ns));
  CK(hipMalloc(&B->dopoff, sizeof(long long) * no));
  CK(hipMalloc(&B->dop, sizeof(int) * ns));
  CK(hipMalloc(&B->dbstart, sizeof(int) * (ns + 1)));
  CK(hipMalloc(&B->dtstart, sizeof(int) * (ns + 1)));
  CK(hipMalloc(&B->dbsys, sizeof(int) * (size_t)std::max(nb, 1)));
  CK(hipMalloc(&B->dtsys, sizeof(int) * (size_t)std::max(nt, 1)));
  CK(hipMalloc(&B->st, sizeof(QmrState) * ns));
  CK(hipMalloc(&B->part, sizeof(double) * 3 * (size_t)std::max(nb, nt)));
  CK(hipMalloc(&B->res, sizeof(double) * 2 * ns));
  std::vector<OpArgs<double>> hops(no);
  for (int q = 0; q < no; ++q) {
    hops[q] = B->ops[q]->op64;
    CK(hipMemcpy(B->invd + opoff[q], B->ops[q]->invd, sizeof(c128) * 4 * B->ops[q]->nn,
                 hipMemcpyDeviceToDevice));
  }
  CK(hipMemcpy(B->dops, hops.data(), sizeof(OpArgs<double>) * no, hipMemcpyHostToDevice));
  CK(hipMemcpy(B->doff, off.data(), sizeof(long long) * ns, hipMemcpyHostToDevice));
  CK(hipMemcpy(B->dn4, n4.data(), sizeof(long long) * ns, hipMemcpyHostToDevice));
  CK(hipMemcpy(B->dopoff, opoff.data(), sizeof(long long) * no, hipMemcpyHostToDevice));
  CK(hipMemcpy(B->dop, op.data(), sizeof(int) * ns, hipMemcpyHostToDevice));
  CK(hipMemcpy(B->dbstart, bstart.data(), sizeof(int) * (ns + 1), hipMemcpyHostToDevice));
  CK(hipMemcpy(B->dtstart, tstart.data(), sizeof(int) * (ns + 1), hipMemcpyHostToDevice));
  {
    std::vector<int> bsys((size_t)nb), tsys((size_t)nt);
    for (int q = 0; q < ns; ++q) {
      for (int b = bstart[q]; b < bstart[q + 1]; ++b) bsys[b] = q;
      for (int t = tstart[q]; t < tstart[q + 1]; ++t) tsys[t] = q;
    }
    if (nb) CK(hipMemcpy(B->dbsys, bsys.data(), sizeof(int) * nb, hipMemcpyHostToDevice));
    if (nt) CK(hipMemcpy(B->dtsys, tsys.data(), sizeof(int) * nt, hipMemcpyHostToDevice));
  }
  CK(hipMemset(B->V[0], 0, sizeof(c128) * B->n4_tot));
  B->opoff = opoff;
  B->hst.assign(ns, QmrState{});
  B->built = true;
  return 0;
}

// kernel k (0 p-update, 1 stencil with eps, 2 ||t - beta v||, 3 fused
// update) of a batch iteration, each with its finalize, with the
// non-temporal hints nts
// the batched stencil on every system: t = K u, with the eps partials when
// with_eps (x-march over flat chunks, or one node per thread, see batch_build)
template <typename V>
void batch_apply(aphid_batch* B, const BatchDev& D, const OpArgs<V>* ops, const cplx<V>* u,
                 cplx<V>* t, int with_eps) {
  if (B->march) {
    const size_t lds = 7 * (size_t)B->wstride * sizeof(cplx<V>);
#define APHID_KB_APPLY(WS)                                                                    \
  hipLaunchKernelGGL((kb_apply<V, kBT, WS>), dim3(B->ntile), dim3(kBT), lds, B->s, D, ops, u, t, \
                     with_eps, B->part, B->xch)
    switch (B->wstride) {
      case kBWindow[0]: APHID_KB_APPLY(kBWindow[0]); break;
      case kBWindow[1]: APHID_KB_APPLY(kBWindow[1]); break;
      case kBWindow[2]: APHID_KB_APPLY(kBWindow[2]); break;
      default: APHID_KB_APPLY(kBWindow[3]);
    }
#undef APHID_KB_APPLY
  } else
    hipLaunchKernelGGL((kb_apply_direct<V, kBT>), dim3(B->ntile), dim3(kBT), 0, B->s, D, ops, u,
                       t, with_eps, B->part);
}

template <typename V>
void batch_launch(aphid_batch* B, const BatchDev& D, int k, const int* nts,
                  const OpArgs<V>* ops, cplx<V>* x, cplx<V>* v, cplx<V>* p, cplx<V>* t, cplx<V>* d,
                  cplx<V>* s, cplx<V>* r, const cplx<V>* invd) {
  const int ns = (int)B->sys.size();
  const dim3 gs(B->nstream), bs(kBlock);
  switch (k) {
    case 0:
      if (nts[SK_PUPDATE])
        hipLaunchKernelGGL((kb_pupdate<V, true>), gs, bs, 0, B->s, D, v, invd, p, d, x);
      else
        hipLaunchKernelGGL((kb_pupdate<V, false>), gs, bs, 0, B->s, D, v, invd, p, d, x);
      break;
    case 1:
      batch_apply<V>(B, D, ops, p, t, 1);
      hipLaunchKernelGGL(kb_fin_eps, dim3(ns), bs, 0, B->s, D, B->part);
      break;
    case 2:
      if (nts[SK_VTNORM])
        hipLaunchKernelGGL((kb_vtnorm<V, true>), gs, bs, 0, B->s, D, t, v, B->part);
      else
        hipLaunchKernelGGL((kb_vtnorm<V, false>), gs, bs, 0, B->s, D, t, v, B->part);
      hipLaunchKernelGGL(kb_fin_vtnorm, dim3(ns), bs, 0, B->s, D, B->part);
      break;
    default:
      if (nts[SK_UPDATE])
        hipLaunchKernelGGL((kb_update<V, true>), gs, bs, 0, B->s, D, t, s, r, v, invd, B->part);
      else
        hipLaunchKernelGGL((kb_update<V, false>), gs, bs, 0, B->s, D, t, s, r, v, invd, B->part);
      hipLaunchKernelGGL(kb_fin_update, dim3(ns), bs, 0, B->s, D, B->part);
  }
}

template <typename V>
int batch_enqueue_iteration(aphid_batch* B) {
  const BatchDev D = batch_dev(B);
  const int ns = (int)B->sys.size();
  cplx<V>*x, *v, *p, *t, *d, *s, *r;
  const cplx<V>* invd;
  if constexpr (std::is_same<V, double>::value) {
    x = B->V[0]; v = B->V[2]; p = B->V[3]; t = B->V[4]; d = B->V[5]; s = B->V[6]; r = B->V[7];
    invd = B->invd;
  } else {
    x = B->F[0]; v = B->F[2]; p = B->F[3]; t = B->F[4]; d = B->F[5]; s = B->F[6]; r = B->F[7];
    invd = B->invdf;
  }
  const OpArgs<V>* ops;
  if constexpr (std::is_same<V, double>::value) ops = B->dops;
  else ops = B->dops32;
  const int* nts = B->nts[std::is_same<V, double>::value ? 0 : 1];
  for (int k = 0; k < 4; ++k)
    batch_launch<V>(B, D, k, nts, ops, x, v, p, t, d, s, r, invd);
  CK(hipGetLastError());
  return 0;
}

// Non-temporal hints of the batch's streaming kernels: all 8 combinations
// timed over the iteration's kernel sequence (each kernel with its
// finalize, scratch states reloaded before every launch, on vectors the
// solve overwrites before it reads them; the spare stands in for x).  A
// combination replaces plain loads and stores only when the sequence is at
// least 3% faster.
template <typename V>
int batch_tune_nt(aphid_batch* B) {
  constexpr int prec = std::is_same<V, double>::value ? 0 : 1;
  if (B->ntuned[prec]) return 0;
  const int ns = (int)B->sys.size();
  if (!B->tst) CK(hipMalloc(&B->tst, sizeof(QmrState) * ns));
  QmrState t1{};
  t1.k1 = {0.5, 0.25};
  t1.beta = {1.0, 0.5};
  t1.eta = {0.5, 0.0};
  t1.eps = {1.0, 0.0};
  t1.del = {1.0, 0.0};
  t1.k2s = 0.5;
  t1.rho_np = t1.rho_n = 1.0;
  t1.c_nm = 1.0;
  t1.max_it = 1 << 30;
  t1.iter = 1;
  QmrState t2 = t1;
  t2.iter = 2;
  // both kinds of update iteration (x and r stored or not), alternately
  const std::vector<QmrState> ts[2] = {std::vector<QmrState>(ns, t1), std::vector<QmrState>(ns, t2)};
  BatchDev D = batch_dev(B);
  D.st = B->tst;
  D.hist = nullptr;
  cplx<V>*x, *v, *p, *t, *d, *s, *r;
  const cplx<V>* invd;
  const OpArgs<V>* ops;
  x = reinterpret_cast<cplx<V>*>(B->V[8]);
  if constexpr (prec == 0) {
    v = B->V[2]; p = B->V[3]; t = B->V[4]; d = B->V[5]; s = B->V[6]; r = B->V[7];
    invd = B->invd; ops = B->dops;
  } else {
    v = B->F[2]; p = B->F[3]; t = B->F[4]; d = B->F[5]; s = B->F[6]; r = B->F[7];
    invd = B->invdf; ops = B->dops32;
  }
  const int reps = B->n4_tot > 8000000 ? 4 : 8;
  auto seq_time = [&](const int (&nt)[SK_COUNT], float* ms) -> int {
    float tot = 0.f;
    for (int rep = 0; rep <= reps; ++rep)
      for (int k = 0; k < 4; ++k) {
        CK(hipMemcpyAsync(B->tst, ts[rep & 1].data(), sizeof(QmrState) * ns, hipMemcpyHostToDevice,
                          B->s));
        CK(hipEventRecord(B->ev0, B->s));
        batch_launch<V>(B, D, k, nt, ops, x, v, p, t, d, s, r, invd);
        CK(hipGetLastError());
        CK(hipEventRecord(B->ev1, B->s));
        CK(hipEventSynchronize(B->ev1));
        float m = 0.f;
        CK(hipEventElapsedTime(&m, B->ev0, B->ev1));
        if (rep) tot += m;
      }
    *ms = tot / reps;
    return 0;
  };
  float def_ms = 0.f, best_ms = 0.f;
  int best[SK_COUNT] = {0, 0, 0};
  for (int m = 0; m < 8; ++m) {
    const int nt[SK_COUNT] = {m & 1, (m >> 1) & 1, (m >> 2) & 1};
    float ms = 0.f;
    const int rc = seq_time(nt, &ms);
    if (rc) return rc;
    if (tune_verbose())
      fprintf(stderr, "[aphid] batch nt prec %d: %d %d %d: %.1f us\n", prec, nt[0], nt[1], nt[2],
              1e3 * ms);
    if (m == 0) def_ms = best_ms = ms;
    if (ms < best_ms) {
      best_ms = ms;
      for (int q = 0; q < SK_COUNT; ++q) best[q] = nt[q];
    }
  }
  for (int q = 0; q < SK_COUNT; ++q) B->nts[prec][q] = best_ms < 0.97f * def_ms ? best[q] : 0;
  B->ntuned[prec] = true;
  return 0;
}

bool batch_all_done(const aphid_batch* B) {
  for (const QmrState& q : B->hst)
    if (!q.done) return false;
  return true;
}

int batch_fetch_states(aphid_batch* B) {
  CK(hipMemcpyAsync(B->hst.data(), B->st, sizeof(QmrState) * B->hst.size(),
                    hipMemcpyDeviceToHost, B->s));
  CK(hipStreamSynchronize(B->s));
  return 0;
}

}  // namespace

extern "C" {

int aphid_batch_create(aphid_batch** out, int device) {
  if (!out) return fail(APHID_E_ARG, "batch pointer is NULL");
  *out = nullptr;
  int ndev = 0;
  CK(hipGetDeviceCount(&ndev));
  if (ndev <= 0) return fail(APHID_E_HIP, "no HIP device");
  if (device < 0) CK(hipGetDevice(&device));
  if (device >= ndev) return fail(APHID_E_ARG, "device index out of range");
  CK(hipSetDevice(device));
  aphid_batch* B = new aphid_batch();
  B->dev = device;
  CK(hipStreamCreateWithFlags(&B->s, hipStreamNonBlocking));
  CK(hipEventCreate(&B->ev0));
  CK(hipEventCreate(&B->ev1));
  *out = B;
  return 0;
}

int aphid_batch_destroy(aphid_batch* B) {
  if (!B) return 0;
  (void)hipSetDevice(B->dev);
  if (B->s) (void)hipStreamSynchronize(B->s);
  batch_free_device(B);
  if (B->stage) (void)hipFree(B->stage);
  for (auto& q : B->sys)
    if (q.staged) (void)hipFree(q.staged);
  for (aphid_ctx* c : B->ops) aphid_ctx_destroy(c);
  if (B->ev0) (void)hipEventDestroy(B->ev0);
  if (B->ev1) (void)hipEventDestroy(B->ev1);
  if (B->s) (void)hipStreamDestroy(B->s);
  delete B;
  return 0;
}

int aphid_batch_add_operator(aphid_batch* B, int nx, int ny, int nz, const double* xn,
                             const double* yn, const double* zn, char model, double omega,
                             const float* a, const float* b, const float* c, const float* d) {
  if (!B) return fail(APHID_E_ARG, "NULL batch");
  if (!B->sys.empty()) return fail(APHID_E_STATE, "add every operator before the first system");
  if (nx < 3 || ny < 3 || nz < 3 || !nodes_fit(nx, ny, nz))
    return fail(APHID_E_ARG, "need 3 to APHID_MAX_NODES-1 nodes (4*nn*16 bytes below 2^32)");
  aphid_ctx* ctx = nullptr;
  int rc = aphid_ctx_create(&ctx, B->dev);
  if (rc) return rc;
  rc = aphid_set_mesh(ctx, nx, ny, nz, xn, yn, zn);
  if (!rc) rc = aphid_set_cells(ctx, model, omega, a, b, c, d);
  if (rc) {
    const std::string e = g_err;
    aphid_ctx_destroy(ctx);
    return fail(rc, e);
  }
  B->ops.push_back(ctx);
  return (int)B->ops.size() - 1;
}

int aphid_batch_add_system(aphid_batch* B, int op, const double* bx, const double* by,
                           const double* bz, const double* bp, const double* x0x,
                           const double* x0y, const double* x0z, const double* x0p) {
  if (!B || !bx || !by || !bz || !bp) return fail(APHID_E_ARG, "NULL argument");
  if (op < 0 || op >= (int)B->ops.size()) return fail(APHID_E_ARG, "no such operator");
  const bool have_x0 = x0x || x0y || x0z || x0p;
  if (have_x0 && !(x0x && x0y && x0z && x0p)) return fail(APHID_E_ARG, "x0 arrays incomplete");
  if (B->built) return fail(APHID_E_STATE, "add every system before the first solve");
  B->sys.push_back({op, 0, 0, have_x0, nullptr, {}, have_x0 ? NAN : 1.0});
  // upload right-hand side and guess into host-side staging until the build
  aphid_ctx* c = B->ops[op];
  const long long n4 = 4 * c->nn;
  c128* tmp = nullptr;
  CK(hipSetDevice(B->dev));
  CK(hipMalloc(&tmp, sizeof(c128) * n4 * (have_x0 ? 2 : 1)));
  const double* hb[4] = {bx, by, bz, bp};
  int rc = upload4(c, hb, tmp);
  if (!rc && have_x0) {
    const double* h0[4] = {x0x, x0y, x0z, x0p};
    rc = upload4(c, h0, tmp + n4);
  }
  if (!rc) {
    hipError_t e = hipStreamSynchronize(c->s);
    if (e != hipSuccess) rc = fail(APHID_E_HIP, hipGetErrorString(e));
  }
  if (rc) {
    (void)hipFree(tmp);
    B->sys.pop_back();
    return rc;
  }
  B->sys.back().staged = tmp;
  return (int)B->sys.size() - 1;
}

}  // extern "C"

namespace {

// complex64 copies of the operators' data for the mixed-precision inner solve
int batch_build_f32(aphid_batch* B) {
  if (B->F[0]) return 0;
  const int ns = (int)B->sys.size(), no = (int)B->ops.size();
  for (auto& v : B->F) CK(hipMalloc(&v, sizeof(c64) * B->n4_tot));
  CK(hipMalloc(&B->invdf, sizeof(c64) * B->op4_tot));
  CK(hipMalloc(&B->dops32, sizeof(OpArgs<float>) * no));
  CK(hipMalloc(&B->scale, sizeof(double) * ns));
  CK(hipMalloc(&B->itol, sizeof(double) * ns));
  CK(hipMalloc(&B->nopen, sizeof(int)));
  CK(hipMalloc(&B->budget, sizeof(int) * ns));
  std::vector<OpArgs<float>> hops(no);
  for (int q = 0; q < no; ++q) {
    hops[q] = B->ops[q]->op32;
    CK(hipMemcpy(B->invdf + B->opoff[q], B->ops[q]->invdf, sizeof(c64) * 4 * B->ops[q]->nn,
                 hipMemcpyDeviceToDevice));
  }
  CK(hipMemcpy(B->dops32, hops.data(), sizeof(OpArgs<float>) * no, hipMemcpyHostToDevice));
  return 0;
}

// One batched QMR solve at storage precision V: b, x and the Krylov vectors
// of that precision; per-system stop targets itol[s]*||b_s|| (or tol).
template <typename V>
int batch_qmr(aphid_batch* B, double tol, const double* itol, int max_it, bool use_x0,
              const int* maxits = nullptr, const int* skip = nullptr) {
  constexpr bool f64 = std::is_same<V, double>::value;
  int trc = batch_tune_nt<V>(B);
  if (trc) return trc;
  const BatchDev D = batch_dev(B);
  const int ns = (int)B->sys.size();
  cplx<V>*x, *b, *v, *p, *t, *d, *s_, *r;
  const cplx<V>* invd;
  const OpArgs<V>* ops;
  if constexpr (f64) {
    x = B->V[0]; b = B->V[1]; v = B->V[2]; p = B->V[3]; t = B->V[4]; d = B->V[5]; s_ = B->V[6];
    r = B->V[7]; invd = B->invd; ops = B->dops;
  } else {
    x = B->F[0]; b = B->F[1]; v = B->F[2]; p = B->F[3]; t = B->F[4]; d = B->F[5]; s_ = B->F[6];
    r = B->F[7]; invd = B->invdf; ops = B->dops32;
  }
  if (f64) B->t_solve = std::chrono::steady_clock::now();
  CK(hipMemsetAsync(B->st, 0, sizeof(QmrState) * ns, B->s));
  hipLaunchKernelGGL(kb_norm_b<V>, dim3(B->nstream), dim3(kBlock), 0, B->s, D, b, B->part);
  if (itol)
    hipLaunchKernelGGL(kb_fin_norm_b_tol, dim3(ns), dim3(kBlock), 0, B->s, D, B->part, itol);
  else
    hipLaunchKernelGGL(kb_fin_norm_b, dim3(ns), dim3(kBlock), 0, B->s, D, B->part, tol);
  if (use_x0) batch_apply<V>(B, D, ops, x, t, 0);
  hipLaunchKernelGGL(kb_init_r<V>, dim3(B->nstream), dim3(kBlock), 0, B->s, D, b,
                     use_x0 ? (const cplx<V>*)t : nullptr, r, B->part);
  hipLaunchKernelGGL(kb_fin_init_r, dim3(ns), dim3(kBlock), 0, B->s, D, B->part);
  hipLaunchKernelGGL(kb_init_v<V>, dim3(B->nstream), dim3(kBlock), 0, B->s, D, r, invd, v, p, d,
                     s_, B->part);
  hipLaunchKernelGGL(kb_fin_init_v, dim3(ns), dim3(kBlock), 0, B->s, D, B->part, max_it,
                     f64 ? B->hist_cap : 0, maxits, skip);
  CK(hipGetLastError());
  int rc = batch_fetch_states(B);
  if (rc) return rc;
  if (batch_all_done(B)) return 0;
  // wall-clock marks per system of this wave (complex128 solves with a
  // history only): (iterations done, seconds since this batched QMR began)
  auto mark = [&]() {
    if (!f64 || !B->hist_cap || !B->wave) return;
    const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() -
                                                   B->t_solve).count();
    for (int q = 0; q < ns; ++q)
      if ((*B->wave)[q]) B->sys[q].marks.push_back({B->hst[q].iter, t});
  };
  const double bytes = 22.0 * (f64 ? 16.0 : 8.0) * (double)B->n4_tot;
  // about 30 ms of work per replay (the host reads every system's state
  // after each one); iterations past a system's stop return at once
  const int k = std::max(4, std::min(128, (int)std::ceil(3e-2 / std::max(bytes / 4.0e12, 25e-6))));
  hipGraphExec_t& gx = f64 ? B->gexec : B->gexec32;
  int& gk = f64 ? B->giters : B->giters32;
  if (!gx || gk != k) {
    if (gx) (void)hipGraphExecDestroy(gx), gx = nullptr;
    if (use_graphs()) {
      hipGraph_t g;
      CK(hipStreamBeginCapture(B->s, hipStreamCaptureModeThreadLocal));
      for (int q = 0; q < k && !rc; ++q) rc = batch_enqueue_iteration<V>(B);
      hipError_t e = hipStreamEndCapture(B->s, &g);
      if (rc) return rc;
      if (e != hipSuccess) return fail(APHID_E_HIP, std::string("capture: ") + hipGetErrorString(e));
      CK(hipGraphInstantiate(&gx, g, nullptr, nullptr, 0));
      CK(hipGraphDestroy(g));
    }
    gk = k;
  }
  const long long max_launch = (long long)max_it / k + 3;
  for (long long l = 0;; ++l) {
    if (gx) {
      CK(hipGraphLaunch(gx, B->s));
    } else {
      for (int q = 0; q < k && !rc; ++q) rc = batch_enqueue_iteration<V>(B);
      if (rc) return rc;
    }
    rc = batch_fetch_states(B);
    if (rc) return rc;
    mark();
    if (batch_all_done(B)) {
      // every system's last d and x updates, and x and r left one step
      // behind by an odd last iteration
      hipLaunchKernelGGL(kb_finish<V>, dim3(B->nstream), dim3(kBlock), 0, B->s, D, x, d, p, r, s_);
      CK(hipGetLastError());
      return 0;
    }
    if (l > max_launch) return fail(APHID_E_STATE, "batched QMR state did not advance");
  }
}

// Systems of at least this many unknowns (4 nn) take the single-system
// path inside a batch solve: one such system alone fills the GPU's
// bandwidth, and the single-system kernels (tuned stencil tiles and stream
// shapes, and a working set that may stay in the Infinity Cache) beat the
// batched ones.  APHID_BATCH_SEQ_MIN overrides the threshold.
long long batch_seq_min() {
  const char* e = getenv("APHID_BATCH_SEQ_MIN");
  const long long v = e ? atoll(e) : 0;
  return v > 0 ? v : (1LL << 20);
}

// Solve system q on its operator's context, in place on the batch's
// vectors: the context borrows the system's slices of the flat arrays.
int batch_solve_single(aphid_batch* B, int q, double tol, int max_it, bool use_x0,
                       QmrState* out) {
  const auto& sy = B->sys[q];
  aphid_ctx* c = B->ops[sy.op];
  if (!c->vborrowed)
    for (auto& v : c->V)
      if (v) (void)hipFree(v), v = nullptr;
  for (int i = 0; i < 9; ++i) c->V[i] = B->V[i] + sy.off;
  c->vborrowed = true;
  int rc = autotune(c, 0);
  if (rc) return rc;
  if (B->hist_cap > c->hist_cap) {
    if (c->hist) CK(hipFree(c->hist));
    CK(hipMalloc(&c->hist, sizeof(double) * B->hist_cap));
    c->hist_cap = B->hist_cap;
    for (auto& g : c->gexec)
      if (g) (void)hipGraphExecDestroy(g), g = nullptr;  // graphs captured the old pointer
  }
  double secs = 0.0;
  rc = qmr_device<double>(c, tol, max_it, use_x0, &secs);
  if (rc) return rc;
  *out = *c->hst;
  if (B->hist_cap) {
    auto& sy2 = B->sys[q];
    const int executed = out->status == ST_MAXIT ? std::max(max_it, 1) : out->iter;
    sy2.hist_len = std::min(std::max(executed, 0), B->hist_cap);
    if (sy2.hist_len > 0)
      CK(hipMemcpy(B->hist + (size_t)q * B->hist_alloc, c->hist, sizeof(double) * sy2.hist_len,
                   hipMemcpyDeviceToDevice));
    sy2.marks = c->marks;
  }
  return 0;
}

// Solve the systems with in_wave[q] != 0 (complex128): those of at least
// batch_seq_min() unknowns one at a time on the single-system path, the
// rest in the batched launches with every other system masked out.
int batch_solve_wave(aphid_batch* B, const std::vector<int>& in_wave, double tol, int max_it,
                     bool use_x0, std::vector<QmrState>& last, std::vector<int>& iters) {
  const int ns = (int)B->sys.size();
  const long long seq_min = batch_seq_min();
  std::vector<int> skip(ns, 1);
  int nbat = 0, nseq = 0;
  for (int q = 0; q < ns; ++q)
    if (in_wave[q]) {
      if (B->sys[q].n4 >= seq_min) ++nseq;
      else skip[q] = 0, ++nbat;
    }
  if (nseq) {
    CK(hipStreamSynchronize(B->s));  // b, x0 and the zeroed x are in place
    for (int q = 0; q < ns; ++q)
      if (in_wave[q] && B->sys[q].n4 >= seq_min) {
        const int rc = batch_solve_single(B, q, tol, max_it, use_x0, &last[q]);
        if (rc) return rc;
        iters[q] = last[q].iter;
      }
  }
  if (nbat) {
    if (!B->dskip) CK(hipMalloc(&B->dskip, sizeof(int) * ns));
    CK(hipMemcpyAsync(B->dskip, skip.data(), sizeof(int) * ns, hipMemcpyHostToDevice, B->s));
    std::vector<int> marked(ns, 0);
    for (int q = 0; q < ns; ++q) marked[q] = !skip[q];
    B->wave = &marked;
    const int rc = batch_qmr<double>(B, tol, nullptr, max_it, use_x0, nullptr, B->dskip);
    B->wave = nullptr;
    if (rc) return rc;
    for (int q = 0; q < ns; ++q)
      if (!skip[q]) {
        last[q] = B->hst[q];
        iters[q] = last[q].iter;
        if (B->hist_cap) {
          const int executed = last[q].status == ST_MAXIT ? std::max(max_it, 1) : last[q].iter;
          B->sys[q].hist_len = std::min(std::max(executed, 0), B->hist_cap);
        }
      }
  }
  return 0;
}

// Initial guess of system q from the current solutions of the systems in
// B->sys[q].from (same mesh), x0 = sum_a al_a x_a, with the amplitudes from
// the Galerkin condition x_a^T (b - K x0) = 0 on q's own operator (the
// bilinear form of QMR), or, when that leaves ||b - K x0|| > ||b||, by
// least squares on the residual; x0 = 0 if neither fit is usable.  Runs on
// q's operator context, which borrows q's slices of the flat vectors (x, b
// and the spare Krylov vectors, which the solve rewrites before reading).
// wave of every system of a sweep schedule: 0 without sources, else one
// more than its latest source; false if the sources form a cycle
bool sweep_levels(const aphid_batch* B, std::vector<int>& level) {
  const int ns = (int)B->sys.size();
  level.assign(ns, 0);
  for (int pass = 0; pass <= ns; ++pass) {
    bool changed = false;
    for (int q = 0; q < ns; ++q)
      for (int f : B->sys[q].from)
        if (level[q] < level[f] + 1) {
          level[q] = level[f] + 1;
          changed = true;
        }
    if (!changed) return true;
  }
  return false;
}

int batch_fit_guess(aphid_batch* B, int q) {
  auto& sy = B->sys[q];
  aphid_ctx* c = B->ops[sy.op];
  if (!c->vborrowed)
    for (auto& v : c->V)
      if (v) (void)hipFree(v), v = nullptr;
  for (int i = 0; i < 9; ++i) c->V[i] = B->V[i] + sy.off;
  c->vborrowed = true;
  int rc = autotune(c, 0);
  if (rc) return rc;
  const long long nn = c->nn, n4 = 4 * nn;
  const ApplyEpilogue ep{nullptr, nullptr};
  GramArgs g{};
  g.b = c->V[1];
  g.n = n4;
  g.nn = nn;
  CombineArgs cb{};
  for (int from : sy.from) {
    const int a = g.nb++;
    const c128* u = B->V[0] + B->sys[from].off;
    CK(launch_apply<double>(c->tile[0], c->op64, u, c->V[2 + a], ep, c->s));  // K u_a
    cb.u[a] = u;
    g.u[a] = u;
    g.ublk[a] = -1;
    g.w[a] = c->V[2 + a];
  }
  cb.nb = g.nb;
  double rel = 1.0, rn = 0.0, bn = 0.0;
  bool lsq = false;
  for (int attempt = 0; attempt < 2; ++attempt) {
    std::vector<c128> al;
    bool ok = false;
    rc = ws_fit(c, g, lsq, al, &ok);
    if (rc) return rc;
    if (ok) {
      for (int a = 0; a < g.nb; ++a) cb.al[a] = al[a];
      hipLaunchKernelGGL(k_combine, dim3(grid_for(n4)), dim3(kBlock), 0, c->s, n4, cb, c->V[0]);
      CK(hipGetLastError());
    } else {
      CK(hipMemsetAsync(c->V[0], 0, sizeof(c128) * n4, c->s));
    }
    rc = true_residual(c, &rel, &rn, &bn);
    if (rc) return rc;
    if (std::isfinite(rel) && (rel <= 1.0 || lsq)) break;
    lsq = true;
  }
  if (!std::isfinite(rel) || rel > 1.0) {
    CK(hipMemsetAsync(c->V[0], 0, sizeof(c128) * n4, c->s));
    CK(hipStreamSynchronize(c->s));
    rel = 1.0;
  }
  sy.r0 = rel;
  return 0;
}

// per-system ||b - K x|| and ||b|| (complex128) into B->res
int batch_true_residual(aphid_batch* B) {
  const BatchDev D = batch_dev(B);
  const int ns = (int)B->sys.size();
  batch_apply<double>(B, D, B->dops, B->V[0], B->V[8], 0);
  hipLaunchKernelGGL(kb_resnorm<double>, dim3(B->nstream), dim3(kBlock), 0, B->s, D, B->V[1],
                     B->V[8], B->part);
  hipLaunchKernelGGL(kb_fin_resnorm, dim3(ns), dim3(kBlock), 0, B->s, D, B->part, B->res);
  CK(hipGetLastError());
  return 0;
}

}  // namespace

extern "C" {

int aphid_batch_solve(aphid_batch* B, double tol, int max_it, unsigned flags,
                      aphid_sys_report* reports, double* seconds) {
  if (!B || B->sys.empty()) return fail(APHID_E_ARG, "empty batch");
  if (!(tol >= 0.0) || max_it < 0) return fail(APHID_E_ARG, "tol and max_it must be >= 0");
  CK(hipSetDevice(B->dev));
  const int ns = (int)B->sys.size();
  bool any_x0 = false;
  if (!B->built) {
    std::vector<c128*> staged(ns);
    for (int q = 0; q < ns; ++q) staged[q] = B->sys[q].staged;
    int rc = batch_build(B);
    if (rc) return rc;
    for (int q = 0; q < ns; ++q) {
      const long long n4 = B->sys[q].n4;
      CK(hipMemcpy(B->V[1] + B->sys[q].off, staged[q], sizeof(c128) * n4, hipMemcpyDeviceToDevice));
      if (B->sys[q].x0)
        CK(hipMemcpy(B->V[0] + B->sys[q].off, staged[q] + n4, sizeof(c128) * n4,
                     hipMemcpyDeviceToDevice));
      CK(hipFree(staged[q]));
      B->sys[q].staged = nullptr;
    }
  }
  // a guess given with the system applies to the first solve; later solves
  // start from the current solutions only when asked (APHID_F_X0)
  if (!B->solved)
    for (const auto& q : B->sys) any_x0 |= q.x0;
  const bool use_x0 = (flags & APHID_F_X0) || any_x0;
  if (!use_x0) CK(hipMemsetAsync(B->V[0], 0, sizeof(c128) * B->n4_tot, B->s));
  B->solved = true;
  if (B->hist_cap > B->hist_alloc) {
    if (B->hist) {
      CK(hipFree(B->hist));
      B->hist = nullptr;
    }
    B->hist_alloc = 0;
    if (B->gexec) (void)hipGraphExecDestroy(B->gexec), B->gexec = nullptr;  // captured the old pointer
    CK(hipMalloc(&B->hist, sizeof(double) * (size_t)ns * B->hist_cap));
    B->hist_alloc = B->hist_cap;
  }
  for (auto& sy : B->sys) {
    sy.hist_len = 0;
    sy.marks.clear();
  }
  CK(hipStreamSynchronize(B->s));
  // wall clock: large systems run on their operator's own stream
  const auto t_start = std::chrono::steady_clock::now();
  std::vector<int> iters(ns, 0), sweeps(ns, 0);
  std::vector<QmrState> last(ns);
  int rc = 0;
  if (!(flags & APHID_F_MIXED)) {
    // one wave, or a sweep: first every system without a sweep guess,
    // then wave by wave the others, each from an initial guess fitted to
    // the solutions of its sources, which earlier waves solved
    std::vector<int> level;
    if (!sweep_levels(B, level)) return fail(APHID_E_STATE, "sweep guesses form a cycle");
    const int top = ns ? *std::max_element(level.begin(), level.end()) : 0;
    for (int q = 0; q < ns; ++q)
      if (level[q] == 0) B->sys[q].r0 = use_x0 ? NAN : 1.0;
    for (int lev = 0; lev <= top; ++lev) {
      std::vector<int> w(ns);
      for (int q = 0; q < ns; ++q) w[q] = level[q] == lev;
      if (lev > 0) {
        CK(hipStreamSynchronize(B->s));
        for (int q = 0; q < ns; ++q)
          if (w[q]) {
            rc = batch_fit_guess(B, q);
            if (rc) return rc;
          }
        for (int q = 0; q < ns; ++q)
          if (w[q]) CK(hipStreamSynchronize(B->ops[B->sys[q].op]->s));
      }
      rc = batch_solve_wave(B, w, tol, max_it, lev > 0 || use_x0, last, iters);
      if (rc) return rc;
    }
  } else {
    rc = batch_build_f32(B);
    if (rc) return rc;
    const BatchDev D = batch_dev(B);
    const double floor_tol = mixed_inner_floor();
    // each sweep may spend at most this many inner iterations per system (a
    // complex64 Krylov process that stagnates is restarted from the refined
    // FP64 iterate); every system also has its own total budget max_it
    const int sweep_cap = std::max(200, std::max(max_it, 1) / 8);
    std::vector<double> hscale(ns);
    std::vector<int> budget(ns);
    for (int sweep = 0; sweep < 64; ++sweep) {
      rc = batch_true_residual(B);
      if (rc) return rc;
      int most = 0;
      for (int q = 0; q < ns; ++q) {
        budget[q] = std::min(sweep_cap, std::max(max_it, 1) - iters[q]);
        most = std::max(most, budget[q]);
      }
      CK(hipMemcpyAsync(B->budget, budget.data(), sizeof(int) * ns, hipMemcpyHostToDevice, B->s));
      CK(hipMemsetAsync(B->nopen, 0, sizeof(int), B->s));
      hipLaunchKernelGGL(kb_inner_setup, dim3((ns + kBlock - 1) / kBlock), dim3(kBlock), 0, B->s, ns,
                         B->res, tol, floor_tol, B->budget, B->scale, B->itol, B->nopen);
      int nopen = 0;
      CK(hipMemcpyAsync(&nopen, B->nopen, sizeof(int), hipMemcpyDeviceToHost, B->s));
      CK(hipMemcpyAsync(hscale.data(), B->scale, sizeof(double) * ns, hipMemcpyDeviceToHost, B->s));
      CK(hipStreamSynchronize(B->s));
      if (nopen == 0 || most <= 0) break;
      hipLaunchKernelGGL(kb_inner_rhs, dim3(B->nstream), dim3(kBlock), 0, B->s, D, B->V[1],
                         B->V[8], B->scale, B->F[1], B->F[0]);
      CK(hipGetLastError());
      rc = batch_qmr<float>(B, 0.0, B->itol, most, false, B->budget);
      if (rc) return rc;
      for (int q = 0; q < ns; ++q) {
        if (hscale[q] == 0.0) continue;
        const QmrState& h = B->hst[q];
        iters[q] += h.status == ST_MAXIT ? std::max(budget[q], 1) : std::max(h.iter, 0);
        ++sweeps[q];
        last[q] = h;
      }
      hipLaunchKernelGGL(kb_refine, dim3(B->nstream), dim3(kBlock), 0, B->s, D, B->V[0], B->F[0],
                         B->res, B->scale);
      CK(hipGetLastError());
    }
  }
  CK(hipStreamSynchronize(B->s));
  const double wall =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  rc = batch_true_residual(B);
  if (rc) return rc;
  std::vector<double> res(2 * ns);
  CK(hipMemcpyAsync(res.data(), B->res, sizeof(double) * 2 * ns, hipMemcpyDeviceToHost, B->s));
  CK(hipStreamSynchronize(B->s));
  int worst = 0;
  for (int q = 0; q < ns; ++q) {
    const double rel = res[2 * q + 1] > 0 ? res[2 * q] / res[2 * q + 1] : res[2 * q];
    int status = last[q].status;
    if (flags & APHID_F_MIXED)
      status = (rel < tol) ? ST_CONV : (std::isfinite(rel) ? ST_MAXIT : ST_BREAK);
    if (reports) {
      reports[q].iters = iters[q];
      reports[q].status = status;
      reports[q].rnorm = (flags & APHID_F_MIXED) ? res[2 * q] : last[q].rnorm;
      reports[q].rhsnorm = res[2 * q + 1];
      reports[q].true_rel_res = rel;
      reports[q].r0_rel = (flags & APHID_F_MIXED) ? NAN : B->sys[q].r0;
    }
    worst = std::max(worst, status);
  }
  if (seconds) *seconds = wall;
  return worst;
}

int aphid_batch_get_solution(aphid_batch* B, int sys, double* x_x, double* x_y, double* x_z,
                             double* x_p) {
  if (!B || !x_x || !x_y || !x_z || !x_p) return fail(APHID_E_ARG, "NULL argument");
  if (sys < 0 || sys >= (int)B->sys.size() || !B->built) return fail(APHID_E_ARG, "no such solved system");
  CK(hipSetDevice(B->dev));
  aphid_ctx* c = B->ops[B->sys[sys].op];
  double* h[4] = {x_x, x_y, x_z, x_p};
  return download4(c, B->V[0] + B->sys[sys].off, h);
}

int aphid_batch_set_sweep_guess(aphid_batch* B, int sys, int nfrom, const int* from) {
  if (!B) return fail(APHID_E_ARG, "NULL batch");
  const int ns = (int)B->sys.size();
  if (sys < 0 || sys >= ns) return fail(APHID_E_ARG, "no such system");
  if (nfrom < 0 || nfrom > APHID_SWEEP_MAX_FROM || (nfrom > 0 && !from))
    return fail(APHID_E_ARG, "nfrom must be 0..APHID_SWEEP_MAX_FROM");
  const aphid_ctx* cs = B->ops[B->sys[sys].op];
  for (int a = 0; a < nfrom; ++a) {
    const int f = from[a];
    if (f < 0 || f >= ns || f == sys) return fail(APHID_E_ARG, "no such source system");
    const aphid_ctx* cf = B->ops[B->sys[f].op];
    if (cf->nx != cs->nx || cf->ny != cs->ny || cf->nz != cs->nz)
      return fail(APHID_E_ARG, "sweep guesses come from systems on the same mesh");
  }
  // a source may itself be fitted from earlier waves, but no system may
  // (through its sources' sources) depend on itself
  std::vector<int> saved = B->sys[sys].from;
  B->sys[sys].from.assign(from, from + nfrom);
  std::vector<int> level;
  if (!sweep_levels(B, level)) {
    B->sys[sys].from = saved;
    return fail(APHID_E_ARG, "sweep guesses would form a cycle");
  }
  return 0;
}

int aphid_batch_set_history(aphid_batch* B, int cap) {
  if (!B) return fail(APHID_E_ARG, "NULL batch");
  if (cap < 0) return fail(APHID_E_ARG, "cap must be >= 0");
  B->hist_cap = cap;
  return 0;
}

int aphid_batch_get_history(aphid_batch* B, int sys, double* hist, double* time, int cap,
                            int* len) {
  if (!B || !len || (cap > 0 && !hist)) return fail(APHID_E_ARG, "NULL argument");
  if (cap < 0) return fail(APHID_E_ARG, "cap must be >= 0");
  if (sys < 0 || sys >= (int)B->sys.size()) return fail(APHID_E_ARG, "no such system");
  const auto& sy = B->sys[sys];
  const int n = B->hist ? std::min(std::min(cap, sy.hist_len), B->hist_alloc) : 0;
  *len = n;
  if (n <= 0) return 0;
  CK(hipSetDevice(B->dev));
  CK(hipMemcpy(hist, B->hist + (size_t)sys * B->hist_alloc, sizeof(double) * n,
               hipMemcpyDeviceToHost));
  if (time) fill_hist_time(sy.marks, time, n);
  return 0;
}

int aphid_batch_set_guess(aphid_batch* B, int sys, int from_sys) {
  if (!B || !B->built) return fail(APHID_E_STATE, "solve the batch once first");
  const int ns = (int)B->sys.size();
  if (sys < 0 || sys >= ns || from_sys < 0 || from_sys >= ns) return fail(APHID_E_ARG, "no such system");
  if (B->sys[sys].op != B->sys[from_sys].op &&
      B->ops[B->sys[sys].op]->nn != B->ops[B->sys[from_sys].op]->nn)
    return fail(APHID_E_ARG, "guess must come from a system on the same mesh");
  if (sys == from_sys) return 0;
  CK(hipSetDevice(B->dev));
  CK(hipMemcpyAsync(B->V[0] + B->sys[sys].off, B->V[0] + B->sys[from_sys].off,
                    sizeof(c128) * B->sys[sys].n4, hipMemcpyDeviceToDevice, B->s));
  CK(hipStreamSynchronize(B->s));
  return 0;
}

}  // extern "C"
