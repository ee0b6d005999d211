// unet.hip -- host side of the warm-start network: a 3-level 3-D U-Net
// (16/32/64 channels) on the cell grid, trained and run on the GPU with the
// bf16 matrix-core kernels of unet.hpp.
//
// Input: 3 features per cell (normalised log10 sigma, skin-depth
// normalised log frequency, source mask), padded to 8 channels.  Output:
// 8 FP32 channels per cell (re, im of Ax, Ay, Az, Phi), a normalised field
// shape; the solver maps it to the staggered unknowns and fits amplitudes.
//
// Layers (ks = kernel size, all stride 1 with zero padding, ReLU except the
// last), levels 0/1/2 at full, half and quarter resolution:
//   0 e1a 3x3x3   3 -> 16  L0      7 d2a 3x3x3 [up 32 | skip 32] -> 32  L1
//   1 e1b 3x3x3  16 -> 16  L0 skip 8 d2b 3x3x3  32 -> 32  L1
//   2 e2a 3x3x3  16 -> 32  L1      9 r2  1x1    32 -> 16  L1, then up x2
//   3 e2b 3x3x3  32 -> 32  L1 skip 10 d1a 3x3x3 [up 16 | skip 16] -> 16 L0
//   4 b1  3x3x3  32 -> 64  L2      11 d1b 3x3x3  16 -> 16  L0
//   5 b2  3x3x3  64 -> 64  L2      12 out 1x1    16 -> 8   L0 (no ReLU)
//   6 r3  1x1    64 -> 32  L2, then up x2
// with 2x average pooling after the skips and nearest upsampling.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "aphid_gpu.h"
#include "lib_internal.hpp"
#include "unet.hpp"

#define UCK(call)                                                                       \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return aphid_fail(APHID_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

namespace {

enum Buf {
  B_IN, B_A1, B_CAT1, B_P1, B_A2, B_CAT2, B_P2, B_A3, B_B3, B_R3,
  B_D2A, B_D2B, B_R2, B_D1A, B_D1B, B_OUT, NBUF
};
struct BufSpec {
  int level, ch;
  bool f32;
};
const BufSpec kBufs[NBUF] = {
    {0, 8, false},  {0, 16, false}, {0, 32, false}, {1, 16, false}, {1, 32, false}, {1, 64, false},
    {2, 32, false}, {2, 64, false}, {2, 64, false}, {2, 32, false}, {1, 32, false}, {1, 32, false},
    {1, 16, false}, {0, 16, false}, {0, 16, false}, {0, 8, true}};

struct LayerSpec {
  int ks, cin, cinp, cout, mode, level, in, in_off, out, out_off;
};
const LayerSpec kLayers[] = {
    {3, 3, 8, 16, 0, 0, B_IN, 0, B_A1, 0},       // 0 e1a
    {3, 16, 16, 16, 0, 0, B_A1, 0, B_CAT1, 16},  // 1 e1b (skip 1)
    {3, 16, 16, 32, 0, 1, B_P1, 0, B_A2, 0},     // 2 e2a
    {3, 32, 32, 32, 0, 1, B_A2, 0, B_CAT2, 32},  // 3 e2b (skip 2)
    {3, 32, 32, 64, 0, 2, B_P2, 0, B_A3, 0},     // 4 b1
    {3, 64, 64, 64, 0, 2, B_A3, 0, B_B3, 0},     // 5 b2
    {1, 64, 64, 32, 0, 2, B_B3, 0, B_R3, 0},     // 6 r3
    {3, 64, 64, 32, 0, 1, B_CAT2, 0, B_D2A, 0},  // 7 d2a
    {3, 32, 32, 32, 0, 1, B_D2A, 0, B_D2B, 0},   // 8 d2b
    {1, 32, 32, 16, 0, 1, B_D2B, 0, B_R2, 0},    // 9 r2
    {3, 32, 32, 16, 0, 0, B_CAT1, 0, B_D1A, 0},  // 10 d1a
    {3, 16, 16, 16, 0, 0, B_D1A, 0, B_D1B, 0},   // 11 d1b
    {1, 16, 16, 8, 1, 0, B_D1B, 0, B_OUT, 0},    // 12 out
};
constexpr int NL = sizeof(kLayers) / sizeof(kLayers[0]);
constexpr int kBlock = 256;
constexpr int kBiasBlocks = 1024;
constexpr int kLossBlocks = 1024;

struct Layer {
  LayerSpec sp;
  int nt, kpad, coutp, cinp_t, kpad_t;
  size_t woff, boff;  // offsets in the flat parameter vector
  bf16_t* wf = nullptr;
  bf16_t* wt = nullptr;
  int nsplit = 1, vper = 32;
};

unsigned grid_for(long long n) {
  return (unsigned)std::min<long long>((n + kBlock - 1) / kBlock, 4096);
}

}  // namespace

struct aphid_unet {
  int dev = 0;
  hipStream_t s = nullptr;
  int X[3] = {0, 0, 0}, Y[3] = {0, 0, 0}, Z[3] = {0, 0, 0};
  int V[3] = {0, 0, 0};
  Layer L[NL];
  long long nparam = 0;
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // flat FP32
  long long adam_t = 0;
  void* buf[NBUF] = {};
  float* dbuf[NBUF] = {};
  bf16_t* gy = nullptr;      // bf16 dY (pre-activation) of the current layer
  float* dwpart = nullptr;   // weight-gradient split partials
  float* bpart = nullptr;    // bias-gradient block partials
  double* lpart = nullptr;   // loss block partials
  double* lacc = nullptr;    // loss accumulator (device)
  float* stage = nullptr;    // FP32 staging for host uploads [V0][8]
  bf16_t* tgt1 = nullptr;    // one staged target [V0][8]
  // dataset (bf16): inputs [n][V0][8], targets [n][V0][8]
  bf16_t* ds_in = nullptr;
  bf16_t* ds_tg = nullptr;
  int ds_cap = 0, ds_n = 0;
  std::mt19937_64 rng;
  hipEvent_t e0 = nullptr, e1 = nullptr;
};

namespace {

int lv_dims(const aphid_unet* n, int level, int& X, int& Y, int& Z) {
  X = n->X[level]; Y = n->Y[level]; Z = n->Z[level];
  return n->V[level];
}

// 3x3x3 layers use the LDS-brick kernel (APHID_CONV_DIRECT=1: the
// global-load kernel, kept as the reference point); 1x1 layers the latter
bool conv_lds() {
  static const bool on = [] {
    const char* e = getenv("APHID_CONV_DIRECT");
    return !(e && e[0] == '1');
  }();
  return on;
}

// Blocks sharing one brick of a 3x3x3 layer: on the deep, small levels,
// whose bricks alone leave most of the 256 CUs idle, the output-channel
// groups are split over up to 4 blocks per brick (1 where there are
// 384 bricks or more).
int conv_split(int bricks, int nb) {
  int sp = 1;
  while (sp < 4 && nb % (2 * sp) == 0 && (long long)bricks * sp < 384) sp *= 2;
  return sp;
}

#define CONV_CASE(KS, CI, CO, MODE)                                                     \
  if (ks == KS && cin == CI && cout == CO && mode == MODE) {                            \
    if (KS == 3 && conv_lds()) {                                                        \
      using Bk = Brick<CI, CO>;                                                         \
      const int bricks = brick_count(a.X, a.Y, a.Z, Bk::TX, Bk::TY);                    \
      const int sp = conv_split(bricks, Bk::NB);                                        \
      if constexpr (Bk::NB % 4 == 0) {                                                  \
        if (sp == 4) {                                                                  \
          hipLaunchKernelGGL((kc_conv3l<CI, CO, MODE, 4>), dim3(bricks, 4), dim3(256), 0, s, a); \
          return hipGetLastError();                                                     \
        }                                                                               \
      }                                                                                 \
      if constexpr (Bk::NB % 2 == 0) {                                                  \
        if (sp == 2) {                                                                  \
          hipLaunchKernelGGL((kc_conv3l<CI, CO, MODE, 2>), dim3(bricks, 2), dim3(256), 0, s, a); \
          return hipGetLastError();                                                     \
        }                                                                               \
      }                                                                                 \
      hipLaunchKernelGGL((kc_conv3l<CI, CO, MODE>), dim3(bricks), dim3(256), 0, s, a);  \
    } else {                                                                            \
      hipLaunchKernelGGL((kc_conv<KS, CI, CO, MODE>), grid, dim3(256), 0, s, a);        \
    }                                                                                   \
    return hipGetLastError();                                                           \
  }

hipError_t launch_conv(int ks, int cin, int cout, int mode, const ConvArgs& a, hipStream_t s) {
  const int nb = (cout + 15) / 16, mt = nb >= 8 ? 1 : 8 / nb;
  const long long waves = ((long long)a.V + 16LL * mt - 1) / (16LL * mt);
  const dim3 grid((unsigned)((waves + 3) / 4));
  // forward
  CONV_CASE(3, 8, 16, 0)
  CONV_CASE(3, 16, 16, 0)
  CONV_CASE(3, 16, 32, 0)
  CONV_CASE(3, 32, 32, 0)
  CONV_CASE(3, 32, 64, 0)
  CONV_CASE(3, 64, 64, 0)
  CONV_CASE(1, 64, 32, 0)
  CONV_CASE(3, 64, 32, 0)
  CONV_CASE(1, 32, 16, 0)
  CONV_CASE(3, 32, 16, 0)
  CONV_CASE(1, 16, 8, 1)
  // input gradients (cin = forward cout, cout = forward cin)
  CONV_CASE(1, 8, 16, 2)
  CONV_CASE(3, 16, 16, 2)
  CONV_CASE(3, 16, 32, 2)
  CONV_CASE(1, 16, 32, 2)
  CONV_CASE(3, 32, 32, 2)
  CONV_CASE(3, 32, 64, 2)
  CONV_CASE(1, 32, 64, 2)
  CONV_CASE(3, 64, 64, 2)
  CONV_CASE(3, 64, 32, 2)
  CONV_CASE(3, 32, 16, 2)
  return hipErrorInvalidValue;
}
#undef CONV_CASE

// 3x3x3 layers: one block per LDS brick (dw_splits gives their count);
// 1x1 layers (and APHID_CONV_DIRECT=1): the voxel-split global-gather kernel
#define DW_CASE(KS, CI, CO)                                                              \
  if (ks == KS && cin == CI && cout == CO) {                                             \
    if constexpr (KS == 3) {                                                             \
      if (conv_lds()) {                                                                  \
        hipLaunchKernelGGL((kc_dw3l<CI, CO>), dim3(grid.y), dim3(256), 0, s, a, a.vper);  \
        return hipGetLastError();                                                        \
      }                                                                                  \
    }                                                                                    \
    hipLaunchKernelGGL((kc_dw<KS, CI, CO>), grid, dim3(256), 0, s, a);                   \
    return hipGetLastError();                                                            \
  }
hipError_t launch_dw(int ks, int cin, int cout, dim3 grid, const DwArgs& a, hipStream_t s) {
  DW_CASE(3, 8, 16)
  DW_CASE(3, 16, 16)
  DW_CASE(3, 16, 32)
  DW_CASE(3, 32, 32)
  DW_CASE(3, 32, 64)
  DW_CASE(3, 64, 64)
  DW_CASE(1, 64, 32)
  DW_CASE(3, 64, 32)
  DW_CASE(1, 32, 16)
  DW_CASE(3, 32, 16)
  DW_CASE(1, 16, 8)
  return hipErrorInvalidValue;
}
#undef DW_CASE

int refresh_weights(aphid_unet* n) {
  for (int l = 0; l < NL; ++l) {
    Layer& L = n->L[l];
    const long long tot = (long long)L.coutp * L.kpad + (L.wt ? (long long)L.cinp_t * L.kpad_t : 0);
    hipLaunchKernelGGL(kc_wconv, dim3(grid_for(tot)), dim3(kBlock), 0, n->s, n->p + L.woff,
                       L.sp.cout, L.sp.cin, L.sp.cinp, L.nt, L.coutp, L.kpad, L.wf, L.cinp_t,
                       L.kpad_t, L.wt);
    UCK(hipGetLastError());
  }
  return 0;
}

int conv_fwd(aphid_unet* n, int l, const bf16_t* in_override) {
  const Layer& L = n->L[l];
  const LayerSpec& sp = L.sp;
  ConvArgs a;
  a.V = lv_dims(n, sp.level, a.X, a.Y, a.Z);
  a.in = in_override ? in_override : (const bf16_t*)n->buf[sp.in];
  a.in_ld = kBufs[sp.in].ch;
  a.in_off = sp.in_off;
  a.w = L.wf;
  a.bias = n->p + L.boff;
  a.out = n->buf[sp.out];
  a.out_ld = kBufs[sp.out].ch;
  a.out_off = sp.out_off;
  UCK(launch_conv(sp.ks, sp.cinp, sp.cout, sp.mode, a, n->s));
  return 0;
}

int forward(aphid_unet* n, const bf16_t* in) {
  int X, Y, Z, rc;
  if ((rc = conv_fwd(n, 0, in))) return rc;
  if ((rc = conv_fwd(n, 1, nullptr))) return rc;
  lv_dims(n, 0, X, Y, Z);
  hipLaunchKernelGGL(kc_pool, dim3(grid_for((long long)n->V[1] * 16)), dim3(kBlock), 0, n->s, X, Y,
                     Z, 16, (const bf16_t*)n->buf[B_CAT1], 32, 16, (bf16_t*)n->buf[B_P1]);
  UCK(hipGetLastError());
  if ((rc = conv_fwd(n, 2, nullptr))) return rc;
  if ((rc = conv_fwd(n, 3, nullptr))) return rc;
  lv_dims(n, 1, X, Y, Z);
  hipLaunchKernelGGL(kc_pool, dim3(grid_for((long long)n->V[2] * 32)), dim3(kBlock), 0, n->s, X, Y,
                     Z, 32, (const bf16_t*)n->buf[B_CAT2], 64, 32, (bf16_t*)n->buf[B_P2]);
  UCK(hipGetLastError());
  for (int l = 4; l <= 6; ++l)
    if ((rc = conv_fwd(n, l, nullptr))) return rc;
  hipLaunchKernelGGL(kc_up, dim3(grid_for((long long)n->V[1] * 32)), dim3(kBlock), 0, n->s, X, Y, Z,
                     32, (const bf16_t*)n->buf[B_R3], (bf16_t*)n->buf[B_CAT2], 64, 0);
  UCK(hipGetLastError());
  for (int l = 7; l <= 9; ++l)
    if ((rc = conv_fwd(n, l, nullptr))) return rc;
  lv_dims(n, 0, X, Y, Z);
  hipLaunchKernelGGL(kc_up, dim3(grid_for((long long)n->V[0] * 16)), dim3(kBlock), 0, n->s, X, Y, Z,
                     16, (const bf16_t*)n->buf[B_R2], (bf16_t*)n->buf[B_CAT1], 32, 0);
  UCK(hipGetLastError());
  for (int l = 10; l <= 12; ++l)
    if ((rc = conv_fwd(n, l, nullptr))) return rc;
  return 0;
}

// backward of one conv layer: dY -> gy (ReLU mask) -> bias grad, weight
// grad, and (need_dx) the input gradient written over dbuf[in]
int conv_bwd(aphid_unet* n, int l, const bf16_t* in_override, bool need_dx) {
  const Layer& L = n->L[l];
  const LayerSpec& sp = L.sp;
  int X, Y, Z;
  const int V = lv_dims(n, sp.level, X, Y, Z);
  const int C = sp.cout;
  const long long nel = (long long)V * C;
  if (C & (C - 1)) return aphid_fail(APHID_E_STATE, "relu backward needs a power-of-two channel count");
  const int nbb = (int)std::min<long long>(kBiasBlocks, (nel + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(kc_relu_bwd, dim3(nbb), dim3(kBlock), 0, n->s, V, C,
                     (const float*)n->dbuf[sp.out], kBufs[sp.out].ch, sp.out_off,
                     sp.mode == 0 ? (const bf16_t*)n->buf[sp.out] : (const bf16_t*)nullptr,
                     kBufs[sp.out].ch, sp.out_off, n->gy, n->bpart);
  UCK(hipGetLastError());
  hipLaunchKernelGGL(kc_bias_reduce, dim3(C), dim3(kBlock), 0, n->s, (const float*)n->bpart, nbb,
                     C, n->g + L.boff);
  UCK(hipGetLastError());
  DwArgs d;
  d.X = X; d.Y = Y; d.Z = Z; d.V = V;
  d.gy = n->gy;
  d.in = in_override ? in_override : (const bf16_t*)n->buf[sp.in];
  d.in_ld = kBufs[sp.in].ch;
  d.in_off = sp.in_off;
  d.part = n->dwpart;
  d.vper = L.vper;
  const dim3 grid((unsigned)((L.kpad + 255) / 256), (unsigned)L.nsplit);
  UCK(launch_dw(sp.ks, sp.cinp, sp.cout, grid, d, n->s));
  const long long ne = (long long)L.coutp * L.kpad;  // entries of one partial
  hipLaunchKernelGGL(kc_dw_reduce, dim3((unsigned)((ne + 63) / 64)), dim3(kBlock), 0, n->s,
                     (const float*)n->dwpart, L.nsplit, L.coutp, L.kpad, sp.cout, sp.cin, sp.cinp,
                     L.nt, n->g + L.woff);
  UCK(hipGetLastError());
  if (need_dx) {
    ConvArgs a;
    a.X = X; a.Y = Y; a.Z = Z; a.V = V;
    a.in = n->gy;
    a.in_ld = C;
    a.in_off = 0;
    a.w = L.wt;
    a.bias = nullptr;
    a.out = n->dbuf[sp.in];
    a.out_ld = kBufs[sp.in].ch;
    a.out_off = sp.in_off;
    UCK(launch_conv(sp.ks, sp.cout, sp.cinp, 2, a, n->s));
  }
  return 0;
}

int backward(aphid_unet* n, const bf16_t* in) {
  int X, Y, Z, rc;
  for (int l = 12; l >= 10; --l)
    if ((rc = conv_bwd(n, l, nullptr, true))) return rc;
  lv_dims(n, 0, X, Y, Z);
  hipLaunchKernelGGL(kc_up_bwd, dim3(grid_for((long long)n->V[1] * 16)), dim3(kBlock), 0, n->s, X,
                     Y, Z, 16, (const float*)n->dbuf[B_CAT1], 32, 0, n->dbuf[B_R2]);
  UCK(hipGetLastError());
  for (int l = 9; l >= 7; --l)
    if ((rc = conv_bwd(n, l, nullptr, true))) return rc;
  lv_dims(n, 1, X, Y, Z);
  hipLaunchKernelGGL(kc_up_bwd, dim3(grid_for((long long)n->V[2] * 32)), dim3(kBlock), 0, n->s, X,
                     Y, Z, 32, (const float*)n->dbuf[B_CAT2], 64, 0, n->dbuf[B_R3]);
  UCK(hipGetLastError());
  for (int l = 6; l >= 4; --l)
    if ((rc = conv_bwd(n, l, nullptr, true))) return rc;
  hipLaunchKernelGGL(kc_pool_bwd, dim3(grid_for((long long)n->V[1] * 32)), dim3(kBlock), 0, n->s,
                     X, Y, Z, 32, (const float*)n->dbuf[B_P2], n->dbuf[B_CAT2], 64, 32);
  UCK(hipGetLastError());
  for (int l = 3; l >= 2; --l)
    if ((rc = conv_bwd(n, l, nullptr, true))) return rc;
  lv_dims(n, 0, X, Y, Z);
  hipLaunchKernelGGL(kc_pool_bwd, dim3(grid_for((long long)n->V[0] * 16)), dim3(kBlock), 0, n->s,
                     X, Y, Z, 16, (const float*)n->dbuf[B_P1], n->dbuf[B_CAT1], 32, 16);
  UCK(hipGetLastError());
  if ((rc = conv_bwd(n, 1, nullptr, true))) return rc;
  return conv_bwd(n, 0, in, false);
}

// loss of the current output against tgt, accumulated into n->lacc with
// weight 1/(8 V batch); writes dbuf[OUT]
int loss(aphid_unet* n, const bf16_t* tgt, int batch) {
  const long long ne = (long long)n->V[0] * 8;
  const double scale = 1.0 / ((double)ne * batch);
  const int nb = (int)std::min<long long>(kLossBlocks, (ne + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(kc_loss, dim3(nb), dim3(kBlock), 0, n->s, ne, (const float*)n->buf[B_OUT], tgt,
                     (float)scale, n->dbuf[B_OUT], n->lpart);
  UCK(hipGetLastError());
  hipLaunchKernelGGL(kc_loss_reduce, dim3(1), dim3(kBlock), 0, n->s, (const double*)n->lpart, nb,
                     scale, n->lacc);
  UCK(hipGetLastError());
  return 0;
}

int zero_grads(aphid_unet* n) {
  hipLaunchKernelGGL(kc_fill, dim3(grid_for(n->nparam)), dim3(kBlock), 0, n->s, n->g, n->nparam, 0.f);
  UCK(hipGetLastError());
  UCK(hipMemsetAsync(n->lacc, 0, sizeof(double), n->s));
  return 0;
}

// upload host features [V0][3] (FP32) into buf[IN] as bf16 [V0][8]
int upload_features(aphid_unet* n, const float* feat, bf16_t* dst) {
  const long long V = n->V[0];
  UCK(hipMemcpyAsync(n->stage, feat, sizeof(float) * V * 3, hipMemcpyHostToDevice, n->s));
  hipLaunchKernelGGL(kc_pack_bf16, dim3(grid_for(V * 8)), dim3(kBlock), 0, n->s, V, 3, 8,
                     (const float*)n->stage, dst);
  UCK(hipGetLastError());
  return 0;
}

int upload_target(aphid_unet* n, const float* tgt, bf16_t* dst) {
  const long long V = n->V[0];
  UCK(hipMemcpyAsync(n->stage, tgt, sizeof(float) * V * 8, hipMemcpyHostToDevice, n->s));
  hipLaunchKernelGGL(kc_pack_bf16, dim3(grid_for(V * 8)), dim3(kBlock), 0, n->s, V, 8, 8,
                     (const float*)n->stage, dst);
  UCK(hipGetLastError());
  return 0;
}

void free_net(aphid_unet* n) {
  for (auto& L : n->L) {
    if (L.wf) (void)hipFree(L.wf);
    if (L.wt) (void)hipFree(L.wt);
    L.wf = L.wt = nullptr;
  }
  void* ptrs[] = {n->p, n->g, n->m, n->v, n->gy, n->dwpart, n->bpart, n->lpart, n->lacc, n->stage,
                  n->tgt1, n->ds_in, n->ds_tg};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  for (auto& b : n->buf)
    if (b) (void)hipFree(b), b = nullptr;
  for (auto& b : n->dbuf)
    if (b) (void)hipFree(b), b = nullptr;
  if (n->e0) (void)hipEventDestroy(n->e0);
  if (n->e1) (void)hipEventDestroy(n->e1);
  if (n->s) (void)hipStreamDestroy(n->s);
}

int create(aphid_unet* n, int dev, int nx, int ny, int nz, unsigned long long seed) {
  n->dev = dev;
  UCK(hipSetDevice(dev));
  UCK(hipStreamCreateWithFlags(&n->s, hipStreamNonBlocking));
  UCK(hipEventCreate(&n->e0));
  UCK(hipEventCreate(&n->e1));
  for (int l = 0; l < 3; ++l) {
    n->X[l] = nx >> l;
    n->Y[l] = ny >> l;
    n->Z[l] = nz >> l;
    n->V[l] = n->X[l] * n->Y[l] * n->Z[l];
  }
  // layers: parameter offsets, kernel-layout weights, weight-gradient splits
  size_t off = 0, part_max = 0, gy_max = 0;
  for (int l = 0; l < NL; ++l) {
    Layer& L = n->L[l];
    L.sp = kLayers[l];
    L.nt = L.sp.ks * L.sp.ks * L.sp.ks;
    L.kpad = cpad(L.nt * L.sp.cinp, 32);
    L.coutp = cpad(L.sp.cout, 16);
    L.cinp_t = cpad(L.sp.cinp, 16);
    L.kpad_t = cpad(L.nt * L.sp.cout, 32);
    L.woff = off;
    off += (size_t)L.sp.cout * L.sp.cin * L.nt;
    L.boff = off;
    off += L.sp.cout;
    const int V = n->V[L.sp.level];
    if (L.sp.ks == 3 && conv_lds()) {  // blocks of LDS bricks, one partial per block
      const int nvox = L.sp.cinp <= 16 ? 512 : L.sp.cinp <= 32 ? 256 : 128;
      const int tx = nvox == 128 ? 2 : 4, ty = nvox / 16 / tx;
      const int nbr = brick_count(n->X[L.sp.level], n->Y[L.sp.level], n->Z[L.sp.level], tx, ty);
      L.nsplit = std::min(nbr, 512);
      L.vper = nbr;  // the kernel's brick count
    } else {
      const int ntile = (L.kpad + 255) / 256;
      const int want = std::max(1, 512 / ntile);  // ~2 blocks per CU
      const int maxs = std::max(1, (V + 31) / 32);
      int ns = std::min(want, maxs);
      L.vper = cpad((V + ns - 1) / ns, 32);
      L.nsplit = (V + L.vper - 1) / L.vper;
    }
    part_max = std::max(part_max, (size_t)L.nsplit * L.coutp * L.kpad);
    gy_max = std::max(gy_max, (size_t)V * L.sp.cout);
    UCK(hipMalloc(&L.wf, sizeof(bf16_t) * L.coutp * L.kpad));
    if (l > 0) UCK(hipMalloc(&L.wt, sizeof(bf16_t) * L.cinp_t * L.kpad_t));
  }
  n->nparam = (long long)off;
  for (float** q : {&n->p, &n->g, &n->m, &n->v}) {
    UCK(hipMalloc(q, sizeof(float) * n->nparam));
    UCK(hipMemset(*q, 0, sizeof(float) * n->nparam));
  }
  for (int b = 0; b < NBUF; ++b) {
    const size_t ne = (size_t)n->V[kBufs[b].level] * kBufs[b].ch;
    UCK(hipMalloc(&n->buf[b], ne * (kBufs[b].f32 ? 4 : 2)));
    UCK(hipMemset(n->buf[b], 0, ne * (kBufs[b].f32 ? 4 : 2)));
    if (b != B_IN) {
      UCK(hipMalloc(&n->dbuf[b], ne * sizeof(float)));
      UCK(hipMemset(n->dbuf[b], 0, ne * sizeof(float)));
    }
  }
  UCK(hipMalloc(&n->gy, sizeof(bf16_t) * gy_max));
  UCK(hipMalloc(&n->dwpart, sizeof(float) * part_max));
  UCK(hipMalloc(&n->bpart, sizeof(float) * kBiasBlocks * 64));
  UCK(hipMalloc(&n->lpart, sizeof(double) * kLossBlocks));
  UCK(hipMalloc(&n->lacc, sizeof(double)));
  UCK(hipMalloc(&n->stage, sizeof(float) * n->V[0] * 8));
  UCK(hipMalloc(&n->tgt1, sizeof(bf16_t) * n->V[0] * 8));
  UCK(hipMemset(n->tgt1, 0, sizeof(bf16_t) * n->V[0] * 8));
  // He-normal weights, zero biases
  n->rng.seed(seed);
  std::vector<float> hp(n->nparam, 0.f);
  for (int l = 0; l < NL; ++l) {
    const Layer& L = n->L[l];
    std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (float)(L.sp.cin * L.nt)));
    for (size_t q = 0; q < (size_t)L.sp.cout * L.sp.cin * L.nt; ++q) hp[L.woff + q] = nd(n->rng);
  }
  UCK(hipMemcpy(n->p, hp.data(), sizeof(float) * n->nparam, hipMemcpyHostToDevice));
  int rc = refresh_weights(n);
  if (rc) return rc;
  UCK(hipStreamSynchronize(n->s));
  return 0;
}

double forward_flops(const aphid_unet* n) {
  double f = 0.0;
  for (int l = 0; l < NL; ++l) {
    const LayerSpec& sp = n->L[l].sp;
    f += 2.0 * n->V[sp.level] * sp.cout * sp.cin * n->L[l].nt;
  }
  return f;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
extern "C" {

aphid_unet* aphid_unet_create(int device, int nx, int ny, int nz, unsigned long long seed) {
  if (nx < 4 || ny < 4 || nz < 4 || nx % 4 || ny % 4 || nz % 4) {
    aphid_fail(APHID_E_ARG, "aphid_unet_create: cell grid dims must be multiples of 4 (>= 4)");
    return nullptr;
  }
  if ((long long)nx * ny * nz > (1LL << 27)) {
    aphid_fail(APHID_E_ARG, "aphid_unet_create: cell grid too large");
    return nullptr;
  }
  aphid_unet* n = new aphid_unet();
  if (create(n, device, nx, ny, nz, seed)) {
    free_net(n);
    delete n;
    return nullptr;
  }
  return n;
}

void aphid_unet_destroy(aphid_unet* n) {
  if (!n) return;
  (void)hipSetDevice(n->dev);
  (void)hipStreamSynchronize(n->s);
  free_net(n);
  delete n;
}

long long aphid_unet_param_count(const aphid_unet* n) { return n ? n->nparam : -1; }

double aphid_unet_forward_flops(const aphid_unet* n) { return n ? forward_flops(n) : 0.0; }

int aphid_unet_get_params(aphid_unet* n, float* out) {
  if (!n || !out) return aphid_fail(APHID_E_ARG, "aphid_unet_get_params: null argument");
  UCK(hipSetDevice(n->dev));
  UCK(hipMemcpyAsync(out, n->p, sizeof(float) * n->nparam, hipMemcpyDeviceToHost, n->s));
  UCK(hipStreamSynchronize(n->s));
  return 0;
}

int aphid_unet_set_params(aphid_unet* n, const float* in) {
  if (!n || !in) return aphid_fail(APHID_E_ARG, "aphid_unet_set_params: null argument");
  UCK(hipSetDevice(n->dev));
  UCK(hipMemcpyAsync(n->p, in, sizeof(float) * n->nparam, hipMemcpyHostToDevice, n->s));
  int rc = refresh_weights(n);
  if (rc) return rc;
  UCK(hipStreamSynchronize(n->s));
  return 0;
}

int aphid_unet_forward(aphid_unet* n, const float* feat, float* out) {
  if (!n || !feat) return aphid_fail(APHID_E_ARG, "aphid_unet_forward: null argument");
  UCK(hipSetDevice(n->dev));
  int rc = upload_features(n, feat, (bf16_t*)n->buf[B_IN]);
  if (!rc) rc = forward(n, nullptr);
  if (rc) return rc;
  if (out)
    UCK(hipMemcpyAsync(out, n->buf[B_OUT], sizeof(float) * n->V[0] * 8, hipMemcpyDeviceToHost, n->s));
  UCK(hipStreamSynchronize(n->s));
  return 0;
}

int aphid_unet_loss_grad(aphid_unet* n, const float* feat, const float* target, double* loss_out,
                         float* grad) {
  if (!n || !feat || !target) return aphid_fail(APHID_E_ARG, "aphid_unet_loss_grad: null argument");
  UCK(hipSetDevice(n->dev));
  int rc = zero_grads(n);
  if (!rc) rc = upload_features(n, feat, (bf16_t*)n->buf[B_IN]);
  if (!rc) {
    UCK(hipStreamSynchronize(n->s));  // the staging buffer is reused
    rc = upload_target(n, target, n->tgt1);
  }
  if (!rc) rc = forward(n, nullptr);
  if (!rc) rc = loss(n, n->tgt1, 1);
  if (!rc) rc = backward(n, nullptr);
  if (rc) return rc;
  if (loss_out) UCK(hipMemcpyAsync(loss_out, n->lacc, sizeof(double), hipMemcpyDeviceToHost, n->s));
  if (grad) UCK(hipMemcpyAsync(grad, n->g, sizeof(float) * n->nparam, hipMemcpyDeviceToHost, n->s));
  UCK(hipStreamSynchronize(n->s));
  return 0;
}

int aphid_unet_activation(aphid_unet* n, int which, int grad, float* out) {
  if (!n || !out || which < 0 || which >= NBUF || (grad && which == B_IN))
    return aphid_fail(APHID_E_ARG, "aphid_unet_activation: bad argument");
  UCK(hipSetDevice(n->dev));
  const long long ne = (long long)n->V[kBufs[which].level] * kBufs[which].ch;
  if (grad || kBufs[which].f32) {
    UCK(hipMemcpyAsync(out, grad ? (void*)n->dbuf[which] : n->buf[which], sizeof(float) * ne,
                       hipMemcpyDeviceToHost, n->s));
    UCK(hipStreamSynchronize(n->s));
    return 0;
  }
  std::vector<bf16_t> h(ne);
  UCK(hipMemcpyAsync(h.data(), n->buf[which], sizeof(bf16_t) * ne, hipMemcpyDeviceToHost, n->s));
  UCK(hipStreamSynchronize(n->s));
  for (long long q = 0; q < ne; ++q) {
    const unsigned u = (unsigned)h[q] << 16;
    std::memcpy(&out[q], &u, 4);
  }
  return 0;
}

int aphid_unet_reserve(aphid_unet* n, int cap) {
  if (!n || cap < 0) return aphid_fail(APHID_E_ARG, "aphid_unet_reserve: bad argument");
  if (n->ds_cap || n->ds_n) return aphid_fail(APHID_E_STATE, "aphid_unet_reserve: already reserved");
  UCK(hipSetDevice(n->dev));
  const size_t per = (size_t)n->V[0] * 8;
  hipError_t e = hipMalloc(&n->ds_in, sizeof(bf16_t) * per * (size_t)std::max(cap, 1));
  if (e == hipSuccess) e = hipMalloc(&n->ds_tg, sizeof(bf16_t) * per * (size_t)std::max(cap, 1));
  if (e != hipSuccess) {
    if (n->ds_in) (void)hipFree(n->ds_in);
    n->ds_in = nullptr;
    return aphid_fail(APHID_E_NOMEM, "aphid_unet_reserve: dataset does not fit in device memory");
  }
  n->ds_cap = cap;
  return 0;
}

int aphid_unet_add_sample(aphid_unet* n, const float* feat, const float* target) {
  if (!n || !feat || !target) return aphid_fail(APHID_E_ARG, "aphid_unet_add_sample: null argument");
  if (n->ds_n >= n->ds_cap) return aphid_fail(APHID_E_STATE, "aphid_unet_add_sample: dataset full");
  UCK(hipSetDevice(n->dev));
  const size_t per = (size_t)n->V[0] * 8;
  int rc = upload_features(n, feat, n->ds_in + per * n->ds_n);
  if (!rc) {
    UCK(hipStreamSynchronize(n->s));  // the staging buffer is reused
    rc = upload_target(n, target, n->ds_tg + per * n->ds_n);
  }
  if (rc) return rc;
  UCK(hipStreamSynchronize(n->s));
  return n->ds_n++;
}

int aphid_unet_train(aphid_unet* n, int steps, int batch, float lr, unsigned long long seed,
                     float* losses, double* seconds) {
  if (!n || steps < 0 || batch < 1) return aphid_fail(APHID_E_ARG, "aphid_unet_train: bad argument");
  if (n->ds_n < 1) return aphid_fail(APHID_E_STATE, "aphid_unet_train: empty dataset");
  UCK(hipSetDevice(n->dev));
  std::mt19937_64 pick(seed);
  const size_t per = (size_t)n->V[0] * 8;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  UCK(hipEventRecord(n->e0, n->s));
  for (int st = 0; st < steps; ++st) {
    int rc = zero_grads(n);
    if (rc) return rc;
    for (int b = 0; b < batch; ++b) {
      const size_t idx = (size_t)(pick() % (unsigned long long)n->ds_n);
      const bf16_t* in = n->ds_in + per * idx;
      if ((rc = forward(n, in))) return rc;
      if ((rc = loss(n, n->ds_tg + per * idx, batch))) return rc;
      if ((rc = backward(n, in))) return rc;
    }
    ++n->adam_t;
    const float bc1 = 1.f - std::pow(b1, (float)n->adam_t), bc2 = 1.f - std::pow(b2, (float)n->adam_t);
    hipLaunchKernelGGL(kc_adam, dim3(grid_for(n->nparam)), dim3(kBlock), 0, n->s, n->nparam, n->p,
                       (const float*)n->g, n->m, n->v, lr, b1, b2, eps, bc1, bc2, 1.f);
    UCK(hipGetLastError());
    if ((rc = refresh_weights(n))) return rc;
    if (losses) {
      double l = 0.0;
      UCK(hipMemcpyAsync(&l, n->lacc, sizeof(double), hipMemcpyDeviceToHost, n->s));
      UCK(hipStreamSynchronize(n->s));
      losses[st] = (float)l;
    }
  }
  UCK(hipEventRecord(n->e1, n->s));
  UCK(hipEventSynchronize(n->e1));
  float ms = 0.f;
  UCK(hipEventElapsedTime(&ms, n->e0, n->e1));
  if (seconds) *seconds = 1e-3 * ms;
  return 0;
}

int aphid_unet_time(aphid_unet* n, int reps, double* fwd_seconds, double* step_seconds) {
  if (!n || reps < 1) return aphid_fail(APHID_E_ARG, "aphid_unet_time: bad argument");
  UCK(hipSetDevice(n->dev));
  const bf16_t* in = (const bf16_t*)n->buf[B_IN];
  int rc = forward(n, in);
  if (rc) return rc;
  float ms = 0.f;
  UCK(hipEventRecord(n->e0, n->s));
  for (int r = 0; r < reps; ++r)
    if ((rc = forward(n, in))) return rc;
  UCK(hipEventRecord(n->e1, n->s));
  UCK(hipEventSynchronize(n->e1));
  UCK(hipEventElapsedTime(&ms, n->e0, n->e1));
  if (fwd_seconds) *fwd_seconds = 1e-3 * ms / reps;
  if (step_seconds) {
    // forward + loss + backward against the staged target (no update)
    UCK(hipEventRecord(n->e0, n->s));
    for (int r = 0; r < reps && !rc; ++r) {
      rc = zero_grads(n);
      if (!rc) rc = forward(n, in);
      if (!rc) rc = loss(n, n->tgt1, 1);
      if (!rc) rc = backward(n, in);
    }
    UCK(hipEventRecord(n->e1, n->s));
    UCK(hipEventSynchronize(n->e1));
    if (rc) return rc;
    UCK(hipEventElapsedTime(&ms, n->e0, n->e1));
    *step_seconds = 1e-3 * ms / reps;
  }
  UCK(hipStreamSynchronize(n->s));
  return 0;
}

}  // extern "C"

namespace unet_internal {
int samples(const aphid_unet* n) { return n ? n->ds_n : 0; }

int zero_grads(aphid_unet* n) {
  UCK(hipSetDevice(n->dev));
  const int rc = ::zero_grads(n);
  if (rc) return rc;
  UCK(hipStreamSynchronize(n->s));
  return 0;
}

int forward_sample(aphid_unet* n, int idx, const float** out) {
  if (idx < 0 || idx >= n->ds_n) return aphid_fail(APHID_E_ARG, "no such training sample");
  UCK(hipSetDevice(n->dev));
  const bf16_t* in = n->ds_in + (size_t)n->V[0] * 8 * idx;
  const int rc = forward(n, in);
  if (rc) return rc;
  UCK(hipStreamSynchronize(n->s));
  *out = (const float*)n->buf[B_OUT];
  return 0;
}

float* out_grad(aphid_unet* n) { return n->dbuf[B_OUT]; }

int backward_sample(aphid_unet* n, int idx) {
  if (idx < 0 || idx >= n->ds_n) return aphid_fail(APHID_E_ARG, "no such training sample");
  UCK(hipSetDevice(n->dev));
  const int rc = backward(n, n->ds_in + (size_t)n->V[0] * 8 * idx);
  if (rc) return rc;
  UCK(hipStreamSynchronize(n->s));
  return 0;
}

int backward_input(aphid_unet* n, float* grad) {
  UCK(hipSetDevice(n->dev));
  const int rc = backward(n, nullptr);
  if (rc) return rc;
  if (grad) UCK(hipMemcpyAsync(grad, n->g, sizeof(float) * n->nparam, hipMemcpyDeviceToHost, n->s));
  UCK(hipStreamSynchronize(n->s));
  return 0;
}

int adam_step(aphid_unet* n, float lr) {
  UCK(hipSetDevice(n->dev));
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  ++n->adam_t;
  const float bc1 = 1.f - std::pow(b1, (float)n->adam_t), bc2 = 1.f - std::pow(b2, (float)n->adam_t);
  hipLaunchKernelGGL(kc_adam, dim3(grid_for(n->nparam)), dim3(kBlock), 0, n->s, n->nparam, n->p,
                     (const float*)n->g, n->m, n->v, lr, b1, b2, eps, bc1, bc2, 1.f);
  UCK(hipGetLastError());
  const int rc = refresh_weights(n);
  if (rc) return rc;
  UCK(hipStreamSynchronize(n->s));
  return 0;
}

int forward_host(aphid_unet* n, const float* feat, const float** out, int* nx, int* ny, int* nz) {
  int rc = aphid_unet_forward(n, feat, nullptr);
  if (rc) return rc;
  *out = (const float*)n->buf[B_OUT];
  *nx = n->X[0];
  *ny = n->Y[0];
  *nz = n->Z[0];
  return 0;
}
}  // namespace unet_internal
