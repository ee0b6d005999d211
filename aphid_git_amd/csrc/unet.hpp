// unet.hpp -- kernels of the warm-start network: 3-D convolutions as
// implicit GEMMs on the bf16 matrix cores (v_mfma_f32_16x16x32_bf16, FP32
// accumulate), their input and weight gradients, 2x pooling / upsampling,
// the loss and the Adam update.
//
// Tensors are channels-last on a cell grid with z fastest:
//   t[((x*Y + y)*Z + z)*ld + off + c]
// so that the 8 channels of one (voxel, tap) pair are one 16-byte load.
// Activations are bf16 (stored as raw 16-bit patterns), activation
// gradients FP32, parameters FP32 with bf16 copies for the matrix cores.
//
// Convolution weights in kernel layout: wk[co][kk], kk = tap*CIN + ci
// (K-major per output channel), rows padded to a multiple of 16 and K to a
// multiple of 32 with zeros.  tap = (dx+1)*9 + (dy+1)*3 + (dz+1) for a 3x3x3
// kernel (cross-correlation, zero padding), tap = 0 for a 1x1 kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short bf16_t;  // raw bf16 bit pattern

__device__ __forceinline__ float bf2f(bf16_t h) { return __uint_as_float((unsigned)h << 16); }
// round to nearest even; NaN stays NaN
__device__ __forceinline__ bf16_t f2bf(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (bf16_t)0x7fc0;
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16_t)(u >> 16);
}

__device__ __forceinline__ bf16x8 ld8(const bf16_t* p) {
  return __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(p));
}
__device__ __forceinline__ bf16x8 zero8() {
  const u32x4 z = {0u, 0u, 0u, 0u};
  return __builtin_bit_cast(bf16x8, z);
}

constexpr int cpad(int n, int m) { return (n + m - 1) / m * m; }

// ---------------------------------------------------------------------------
// Convolution, forward and input-gradient (the same implicit GEMM):
//   out[v][co] = sum_{tap,ci} w[co][tap*CIN+ci] * in[v + off(tap)][ci]
// MODE 0: + bias, ReLU, bf16 out.  MODE 1: + bias, FP32 out.
// MODE 2: FP32 out, no bias (input gradient with flipped, transposed weights).
struct ConvArgs {
  int X, Y, Z, V;
  const bf16_t* in;
  int in_ld, in_off;
  const bf16_t* w;     // [cpad(COUT,16)][cpad(NT*CIN,32)]
  const float* bias;   // [COUT] (MODE 0/1)
  void* out;
  int out_ld, out_off;
};

template <int KS, int CIN, int COUT, int MODE>
__global__ __launch_bounds__(256) void kc_conv(ConvArgs a) {
  constexpr int NT = KS * KS * KS;
  constexpr int KPAD = cpad(NT * CIN, 32);
  constexpr int NB = (COUT + 15) / 16;
  constexpr int MT = NB >= 8 ? 1 : 8 / NB;  // 16-voxel tiles per wave
  static_assert(CIN % 8 == 0, "CIN must be a multiple of 8");
  const int lane = threadIdx.x % 64, r = lane % 16, g = lane / 16;
  const int wave = blockIdx.x * 4 + threadIdx.x / 64;
  const int YZ = a.Y * a.Z;
  const int v0 = wave * MT * 16;
  if (v0 >= a.V) return;

  int vx[MT], vy[MT], vz[MT];
  bool vv[MT];
#pragma unroll
  for (int t = 0; t < MT; ++t) {
    const int v = v0 + 16 * t + r;
    vv[t] = v < a.V;
    const int vc = vv[t] ? v : 0;
    vx[t] = vc / YZ;
    const int rem = vc - vx[t] * YZ;
    vy[t] = rem / a.Z;
    vz[t] = rem - vy[t] * a.Z;
  }
  f32x4 acc[MT][NB];
#pragma unroll
  for (int t = 0; t < MT; ++t)
#pragma unroll
    for (int n = 0; n < NB; ++n) acc[t][n] = f32x4{0.f, 0.f, 0.f, 0.f};

  for (int ks = 0; ks < KPAD / 32; ++ks) {
    const int kk = ks * 32 + 8 * g;
    const int tap = kk / CIN, ci = kk - tap * CIN;
    int dx = 0, dy = 0, dz = 0;
    if (KS == 3) {
      dx = tap / 9 - 1;
      dy = (tap / 3) % 3 - 1;
      dz = tap % 3 - 1;
    }
    const bool tv = tap < NT;
    bf16x8 bw[NB];
#pragma unroll
    for (int n = 0; n < NB; ++n) bw[n] = ld8(a.w + (size_t)(16 * n + r) * KPAD + kk);
#pragma unroll
    for (int t = 0; t < MT; ++t) {
      const int x = vx[t] + dx, y = vy[t] + dy, z = vz[t] + dz;
      const bool ok = tv && vv[t] && x >= 0 && x < a.X && y >= 0 && y < a.Y && z >= 0 && z < a.Z;
      const int vn = (x * a.Y + y) * a.Z + z;
      const bf16x8 av = ok ? ld8(a.in + (size_t)vn * a.in_ld + a.in_off + ci) : zero8();
#pragma unroll
      for (int n = 0; n < NB; ++n)
        acc[t][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bw[n], acc[t][n], 0, 0, 0);
    }
  }
  // D[i][j]: lane holds rows i = 4g+e (voxels), column j = r (output channel)
#pragma unroll
  for (int t = 0; t < MT; ++t)
#pragma unroll
    for (int n = 0; n < NB; ++n) {
      const int co = 16 * n + r;
      if (co >= COUT) continue;
      const float b = MODE == 2 ? 0.f : a.bias[co];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int v = v0 + 16 * t + 4 * g + e;
        if (v >= a.V) continue;
        float val = acc[t][n][e] + b;
        const size_t o = (size_t)v * a.out_ld + a.out_off + co;
        if (MODE == 0) {
          val = fmaxf(val, 0.f);
          reinterpret_cast<bf16_t*>(a.out)[o] = f2bf(val);
        } else {
          reinterpret_cast<float*>(a.out)[o] = val;
        }
      }
    }
}

// ---------------------------------------------------------------------------
// The same 3x3x3 convolution with the input staged in LDS.  A block owns an
// output brick of TX x TY x 16 voxels (4 waves x MT rows of 16 z-voxels;
// TX*TY = 4*MT) and loads the brick with a one-voxel halo once, zero outside
// the grid; every tap's A fragment is then one 16-byte LDS read, so each
// input value leaves global memory once per block instead of once per tap.
template <int CIN, int COUT>
struct Brick {
  static constexpr int NB = (COUT + 15) / 16;
  static constexpr int MT = NB >= 8 ? 1 : 8 / NB;
  static constexpr int TX = MT >= 8 ? 4 : MT >= 4 ? 4 : 2;
  static constexpr int TY = 4 * MT / TX;
  static constexpr int BX = TX + 2, BY = TY + 2, BZ = 18;
  static constexpr int NV = BX * BY * BZ;  // staged voxels
  static_assert(TX * TY == 4 * MT, "brick shape");
};

inline int brick_count(int X, int Y, int Z, int tx, int ty) {
  return ((X + tx - 1) / tx) * ((Y + ty - 1) / ty) * ((Z + 15) / 16);
}

// SPLIT > 1 (layers with fewer bricks than CUs): gridDim.y = SPLIT blocks
// share a brick, each computing NB/SPLIT of its 16-channel output groups
// (group blockIdx.y*NB/SPLIT + n); every output's sum runs in the same
// order as with one block per brick.
template <int CIN, int COUT, int MODE, int SPLIT = 1>
__global__ __launch_bounds__(256) void kc_conv3l(ConvArgs a) {
  using Bk = Brick<CIN, COUT>;
  constexpr int NBT = Bk::NB, MT = Bk::MT, TX = Bk::TX, TY = Bk::TY;
  static_assert(NBT % SPLIT == 0, "SPLIT must divide the output channel groups");
  constexpr int NB = NBT / SPLIT;
  const int nb0 = (int)blockIdx.y * NB;  // first output channel group of this block
  constexpr int BY = Bk::BY, BZ = Bk::BZ, NV = Bk::NV;
  constexpr int KPAD = cpad(27 * CIN, 32);
  constexpr int PCS = CIN / 8;  // 16-byte pieces per voxel
  static_assert(CIN % 8 == 0, "CIN must be a multiple of 8");
  __shared__ __attribute__((aligned(16))) bf16_t lds[NV * CIN];
  const int tid = threadIdx.x, lane = tid % 64, r = lane % 16, g = lane / 16, w = tid / 64;
  const int nbz = (a.Z + 15) / 16, nby = (a.Y + TY - 1) / TY;
  const int b = blockIdx.x;
  const int z0 = (b % nbz) * 16, y0 = ((b / nbz) % nby) * TY, x0 = (b / (nbz * nby)) * TX;

  // stage the brick with its halo (coalesced: pieces of a voxel, then z);
  // SB loads are in flight before their LDS stores
  constexpr int SB = 4;
  for (int p0 = tid; p0 < NV * PCS; p0 += SB * 256) {
    u32x4 q[SB];
#pragma unroll
    for (int u = 0; u < SB; ++u) {
      const int p = p0 + u * 256;
      const int vox = p / PCS, pc = p - vox * PCS;
      const int lz = vox % BZ, ly = (vox / BZ) % BY, lx = vox / (BZ * BY);
      const int x = x0 + lx - 1, y = y0 + ly - 1, z = z0 + lz - 1;
      const bool in = p < NV * PCS && x >= 0 && x < a.X && y >= 0 && y < a.Y && z >= 0 && z < a.Z;
      const size_t src = in ? (size_t)((x * a.Y + y) * a.Z + z) * a.in_ld + a.in_off + 8 * pc : 0;
      q[u] = *reinterpret_cast<const u32x4*>(a.in + src);
      if (!in) q[u] = u32x4{0u, 0u, 0u, 0u};
    }
#pragma unroll
    for (int u = 0; u < SB; ++u) {
      const int p = p0 + u * 256;
      if (p < NV * PCS) {
        const int vox = p / PCS, pc = p - vox * PCS;
        *reinterpret_cast<u32x4*>(&lds[vox * CIN + 8 * pc]) = q[u];
      }
    }
  }
  __syncthreads();

  f32x4 acc[MT][NB];
#pragma unroll
  for (int t = 0; t < MT; ++t)
#pragma unroll
    for (int n = 0; n < NB; ++n) acc[t][n] = f32x4{0.f, 0.f, 0.f, 0.f};

  // the weight fragments of step ks+1 are loaded while step ks runs its
  // MFMAs, so the global-load latency is paid once, not once per step
  constexpr int NK = KPAD / 32;
  bf16x8 bw[NB], bwn[NB];
#pragma unroll
  for (int n = 0; n < NB; ++n) bw[n] = ld8(a.w + (size_t)(16 * (nb0 + n) + r) * KPAD + 8 * g);
  for (int ks = 0; ks < NK; ++ks) {
    const int kk = ks * 32 + 8 * g;
    const int tap = kk / CIN, ci = kk - tap * CIN;
    const bool tv = tap < 27;
    const int dx = tap / 9, dy = (tap / 3) % 3, dz = tap % 3;  // offsets + 1
    const int kn = ks + 1 < NK ? kk + 32 : kk;  // the last step reloads its own
#pragma unroll
    for (int n = 0; n < NB; ++n) bwn[n] = ld8(a.w + (size_t)(16 * (nb0 + n) + r) * KPAD + kn);
#pragma unroll
    for (int t = 0; t < MT; ++t) {
      const int row = w * MT + t, tx = row / TY, ty = row - tx * TY;
      const int vox = ((tx + dx) * BY + (ty + dy)) * BZ + (r + dz);
      const bf16x8 av = tv ? ld8(&lds[vox * CIN + ci]) : zero8();
#pragma unroll
      for (int n = 0; n < NB; ++n)
        acc[t][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bw[n], acc[t][n], 0, 0, 0);
    }
#pragma unroll
    for (int n = 0; n < NB; ++n) bw[n] = bwn[n];
  }
#pragma unroll
  for (int t = 0; t < MT; ++t) {
    const int row = w * MT + t, tx = row / TY, ty = row - tx * TY;
    const int x = x0 + tx, y = y0 + ty;
    if (x >= a.X || y >= a.Y) continue;
#pragma unroll
    for (int n = 0; n < NB; ++n) {
      const int co = 16 * (nb0 + n) + r;
      if (co >= COUT) continue;
      const float bias = MODE == 2 ? 0.f : a.bias[co];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int z = z0 + 4 * g + e;
        if (z >= a.Z) continue;
        float val = acc[t][n][e] + bias;
        const size_t o = (size_t)((x * a.Y + y) * a.Z + z) * a.out_ld + a.out_off + co;
        if (MODE == 0) {
          val = fmaxf(val, 0.f);
          reinterpret_cast<bf16_t*>(a.out)[o] = f2bf(val);
        } else {
          reinterpret_cast<float*>(a.out)[o] = val;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Weight gradient: dw[co][kk] = sum_v gy[v][co] * in[v + off(tap(kk))][ci(kk)]
// A GEMM with K running over voxels.  A block owns 256 columns kk (4 waves x
// 4 blocks of 16) and one voxel split; each 32-voxel step stages gy and the
// gathered inputs transposed ([row][voxel]) in LDS, so that every lane reads
// its 8 consecutive voxels as one 16-byte LDS load.  Per-split partials are
// summed in a fixed order by kc_dw_reduce.
struct DwArgs {
  int X, Y, Z, V;
  const bf16_t* gy;  // [V][COUT]
  const bf16_t* in;
  int in_ld, in_off;
  float* part;       // [nsplit][cpad(COUT,16)][KPAD]
  int vper;          // voxels per split, a multiple of 32
};

constexpr int kLdsPitch = 40;  // bf16 per LDS row (32 voxels + 8 pad)

template <int KS, int CIN, int COUT>
__global__ __launch_bounds__(256) void kc_dw(DwArgs a) {
  constexpr int NT = KS * KS * KS;
  constexpr int KTOT = NT * CIN;
  constexpr int KPAD = cpad(KTOT, 32);
  constexpr int COUTP = cpad(COUT, 16);
  constexpr int MB = COUTP / 16;
  __shared__ __attribute__((aligned(16))) bf16_t gs[COUTP][kLdsPitch];
  __shared__ __attribute__((aligned(16))) bf16_t xs[256][kLdsPitch];
  const int tid = threadIdx.x, lane = tid % 64, r = lane % 16, g = lane / 16, w = tid / 64;
  const int kbase = blockIdx.x * 256;
  const int s = blockIdx.y;
  const int vb0 = s * a.vper;
  const int vb1 = min(a.V, vb0 + a.vper);
  const int YZ = a.Y * a.Z;

  f32x4 acc[MB][4];
#pragma unroll
  for (int m = 0; m < MB; ++m)
#pragma unroll
    for (int n = 0; n < 4; ++n) acc[m][n] = f32x4{0.f, 0.f, 0.f, 0.f};

  for (int vb = vb0; vb < vb1; vb += 32) {
    // stage gy: 32 voxels x COUTP/8 pieces of 8 channels
    if (tid < 32 * (COUTP / 8)) {
      const int vl = tid % 32, pc = tid / 32, v = vb + vl;
      bf16_t h[8];
      if (v < vb1 && 8 * pc < COUT) {
        const uint4 q = *reinterpret_cast<const uint4*>(a.gy + (size_t)v * COUT + 8 * pc);
        const unsigned qq[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          h[2 * e] = (bf16_t)(qq[e] & 0xffffu);
          h[2 * e + 1] = (bf16_t)(qq[e] >> 16);
        }
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) h[e] = 0;
      }
#pragma unroll
      for (int e = 0; e < 8; ++e) gs[8 * pc + e][vl] = h[e];
    }
    // stage the gathered inputs: 32 pieces of 8 columns x 32 voxels
#pragma unroll
    for (int q4 = 0; q4 < 4; ++q4) {
      const int idx = tid + 256 * q4, vl = idx % 32, pc = idx / 32;
      const int kk0 = kbase + 8 * pc, v = vb + vl;
      bf16_t h[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) h[e] = 0;
      if (kk0 < KTOT && v < vb1) {
        const int tap = kk0 / CIN, ci = kk0 - tap * CIN;
        int dx = 0, dy = 0, dz = 0;
        if (KS == 3) {
          dx = tap / 9 - 1;
          dy = (tap / 3) % 3 - 1;
          dz = tap % 3 - 1;
        }
        const int x0 = v / YZ, rem = v - x0 * YZ, y0 = rem / a.Z, z0 = rem - y0 * a.Z;
        const int x = x0 + dx, y = y0 + dy, z = z0 + dz;
        if (x >= 0 && x < a.X && y >= 0 && y < a.Y && z >= 0 && z < a.Z) {
          const int vn = (x * a.Y + y) * a.Z + z;
          const uint4 q =
              *reinterpret_cast<const uint4*>(a.in + (size_t)vn * a.in_ld + a.in_off + ci);
          const unsigned qq[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            h[2 * e] = (bf16_t)(qq[e] & 0xffffu);
            h[2 * e + 1] = (bf16_t)(qq[e] >> 16);
          }
        }
      }
#pragma unroll
      for (int e = 0; e < 8; ++e) xs[8 * pc + e][vl] = h[e];
    }
    __syncthreads();
    bf16x8 bx[4];
#pragma unroll
    for (int n = 0; n < 4; ++n) bx[n] = ld8(&xs[w * 64 + 16 * n + r][8 * g]);
#pragma unroll
    for (int m = 0; m < MB; ++m) {
      const bf16x8 ag = ld8(&gs[16 * m + r][8 * g]);
#pragma unroll
      for (int n = 0; n < 4; ++n)
        acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ag, bx[n], acc[m][n], 0, 0, 0);
    }
    __syncthreads();
  }
  float* P = a.part + (size_t)s * COUTP * KPAD;
#pragma unroll
  for (int m = 0; m < MB; ++m)
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      const int kk = kbase + w * 64 + 16 * n + r;
      if (kk >= KPAD) continue;
#pragma unroll
      for (int e = 0; e < 4; ++e) P[(size_t)(16 * m + 4 * g + e) * KPAD + kk] = acc[m][n][e];
    }
}

// ---------------------------------------------------------------------------
// Weight gradient of a 3x3x3 layer from LDS bricks of TX x TY x 16 output
// voxels (NVOX = 512, 256 or 128 for CIN <= 16, 32, 64).  Block b
// accumulates bricks b, b + gridDim, ... in registers (a wave holds all of
// its output columns: n-blocks w, w+4, ...) and writes one partial dw to
// part[b].  For each brick it stages
//   gt[co][vox]          dY of its voxels, transposed (voxel-contiguous);
//   bt[ci][bx][by][bz]   its input brick with a one-voxel halo, transposed
//                        (z-contiguous, rows padded to 24);
// so that both matrix-core operands are 16-byte LDS reads of 8 consecutive
// voxels.  K runs over the block's voxels in steps of 32 = two 16-voxel rows;
// a lane's 8 voxels are z0..z0+7 (z0 = 0 or 8) of one row, whose inputs for
// a tap with z offset dz are bt[...][z0+dz .. z0+dz+7]: two aligned reads and
// a 16- or 32-bit funnel shift.
template <int CIN>
struct DwBrick {
  static constexpr int NVOX = CIN <= 16 ? 512 : CIN <= 32 ? 256 : 128;
  static constexpr int TX = NVOX == 512 ? 4 : NVOX == 256 ? 4 : 2;
  static constexpr int TY = NVOX / 16 / TX;
  static constexpr int BX = TX + 2, BY = TY + 2, BZP = 24;
};

template <int CIN, int COUT>
__global__ __launch_bounds__(256) void kc_dw3l(DwArgs a, int nbricks) {
  using Bk = DwBrick<CIN>;
  constexpr int NVOX = Bk::NVOX, TX = Bk::TX, TY = Bk::TY, BX = Bk::BX, BY = Bk::BY;
  constexpr int BZP = Bk::BZP, ROW = 18;  // staged z extent of a brick row
  constexpr int KTOT = 27 * CIN, KPAD = cpad(KTOT, 32);
  constexpr int COUTP = cpad(COUT, 16), MB = COUTP / 16;
  constexpr int NBN = KPAD / 16;            // n-blocks of 16 columns
  constexpr int NPW = (NBN + 3) / 4;        // n-blocks of one wave (w, w+4, ...)
  constexpr int NCMAX = 64 / MB;            // accumulators <= 256 per lane
  constexpr int NC = NPW < NCMAX ? NPW : NCMAX;
  constexpr int NPASS = (NPW + NC - 1) / NC;
  constexpr int BPL = BX * BY * BZP;        // elements per channel of the brick
  static_assert(CIN % 8 == 0 && COUT % 8 == 0, "channel counts");
  __shared__ __attribute__((aligned(16))) bf16_t gt[COUTP * NVOX];
  __shared__ __attribute__((aligned(16))) bf16_t bt[CIN * BPL];
  const int tid = threadIdx.x, lane = tid % 64, r = lane % 16, g = lane / 16, w = tid / 64;
  const int nbz = (a.Z + 15) / 16, nby = (a.Y + TY - 1) / TY;
  float* P = a.part + (size_t)blockIdx.x * COUTP * KPAD;

  for (int pass = 0; pass < NPASS; ++pass) {
    f32x4 acc[MB][NC];
#pragma unroll
    for (int m = 0; m < MB; ++m)
#pragma unroll
      for (int c = 0; c < NC; ++c) acc[m][c] = f32x4{0.f, 0.f, 0.f, 0.f};

    for (int br = blockIdx.x; br < nbricks; br += gridDim.x) {
      const int z0b = (br % nbz) * 16, y0b = ((br / nbz) % nby) * TY, x0b = (br / (nbz * nby)) * TX;
      // dY of the brick's voxels, transposed; zero outside the grid and for
      // padded output channels
      // (SB loads in flight before their LDS stores, in both staging loops)
      constexpr int SB = 4;
      for (int p0 = tid; p0 < NVOX * (COUTP / 8); p0 += SB * 256) {
        u32x4 qs[SB];
#pragma unroll
        for (int u = 0; u < SB; ++u) {
          const int p = p0 + u * 256;
          const int vox = p % NVOX, pc = p / NVOX;
          const int row = vox / 16, z = z0b + vox % 16, x = x0b + row / TY, y = y0b + row % TY;
          const bool in = p < NVOX * (COUTP / 8) && 8 * pc < COUT && x < a.X && y < a.Y && z < a.Z;
          qs[u] = *reinterpret_cast<const u32x4*>(
              a.gy + (in ? (size_t)((x * a.Y + y) * a.Z + z) * COUT + 8 * pc : 0));
          if (!in) qs[u] = u32x4{0u, 0u, 0u, 0u};
        }
#pragma unroll
        for (int u = 0; u < SB; ++u) {
          const int p = p0 + u * 256;
          if (p >= NVOX * (COUTP / 8)) continue;
          const int vox = p % NVOX, pc = p / NVOX;
          const u32x4 q = qs[u];
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            gt[(8 * pc + 2 * e) * NVOX + vox] = (bf16_t)(q[e] & 0xffffu);
            gt[(8 * pc + 2 * e + 1) * NVOX + vox] = (bf16_t)(q[e] >> 16);
          }
        }
      }
      // the input brick with its halo, transposed; zero outside the grid
      constexpr int PCS = CIN / 8;
      for (int p0 = tid; p0 < BX * BY * ROW * PCS; p0 += SB * 256) {
        u32x4 qs[SB];
#pragma unroll
        for (int u = 0; u < SB; ++u) {
          const int p = p0 + u * 256;
          const int pc = p % PCS, vox = p / PCS;
          const int lz = vox % ROW, ly = (vox / ROW) % BY, lx = vox / (ROW * BY);
          const int x = x0b + lx - 1, y = y0b + ly - 1, z = z0b + lz - 1;
          const bool in = p < BX * BY * ROW * PCS && x >= 0 && x < a.X && y >= 0 && y < a.Y &&
                          z >= 0 && z < a.Z;
          qs[u] = *reinterpret_cast<const u32x4*>(
              a.in + (in ? (size_t)((x * a.Y + y) * a.Z + z) * a.in_ld + a.in_off + 8 * pc : 0));
          if (!in) qs[u] = u32x4{0u, 0u, 0u, 0u};
        }
#pragma unroll
        for (int u = 0; u < SB; ++u) {
          const int p = p0 + u * 256;
          if (p >= BX * BY * ROW * PCS) continue;
          const int pc = p % PCS, vox = p / PCS;
          const int lz = vox % ROW, ly = (vox / ROW) % BY, lx = vox / (ROW * BY);
          const int base = (lx * BY + ly) * BZP + lz;
          const u32x4 q = qs[u];
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            bt[(8 * pc + 2 * e) * BPL + base] = (bf16_t)(q[e] & 0xffffu);
            bt[(8 * pc + 2 * e + 1) * BPL + base] = (bf16_t)(q[e] >> 16);
          }
        }
      }
      __syncthreads();

      for (int ks = 0; ks < NVOX / 32; ++ks) {
        const int v8 = ks * 32 + 8 * g;  // first of this lane's 8 voxels
        const int row = v8 / 16, z0 = v8 % 16, tx = row / TY, ty = row % TY;
        bf16x8 ag[MB];
#pragma unroll
        for (int m = 0; m < MB; ++m) ag[m] = ld8(&gt[(16 * m + r) * NVOX + v8]);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          const int nb = w + 4 * (pass * NC + c);
          const int col = 16 * nb + r;
          const int tap = col / CIN, ci = col - tap * CIN;
          bf16x8 bv = zero8();
          if (nb < NBN && tap < 27) {
            const int dx = tap / 9, dy = (tap / 3) % 3, dz = tap % 3;
            const bf16_t* src = &bt[ci * BPL + ((tx + dx) * BY + (ty + dy)) * BZP + z0];
            const u32x4 q0 = *reinterpret_cast<const u32x4*>(src);
            const u32x4 q1 = *reinterpret_cast<const u32x4*>(src + 8);
            u32x4 q;
            if (dz == 0) {
              q = q0;
            } else if (dz == 1) {
              q[0] = __builtin_amdgcn_alignbit(q0[1], q0[0], 16);
              q[1] = __builtin_amdgcn_alignbit(q0[2], q0[1], 16);
              q[2] = __builtin_amdgcn_alignbit(q0[3], q0[2], 16);
              q[3] = __builtin_amdgcn_alignbit(q1[0], q0[3], 16);
            } else {
              q[0] = q0[1];
              q[1] = q0[2];
              q[2] = q0[3];
              q[3] = q1[0];
            }
            bv = __builtin_bit_cast(bf16x8, q);
          }
#pragma unroll
          for (int m = 0; m < MB; ++m)
            acc[m][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ag[m], bv, acc[m][c], 0, 0, 0);
        }
      }
      __syncthreads();
    }
#pragma unroll
    for (int m = 0; m < MB; ++m)
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const int nb = w + 4 * (pass * NC + c);
        if (nb >= NBN) continue;
        const int kk = 16 * nb + r;
#pragma unroll
        for (int e = 0; e < 4; ++e) P[(size_t)(16 * m + 4 * g + e) * KPAD + kk] = acc[m][c][e];
      }
  }
}

// grad[(co*cin_real + ci)*NT + tap] += sum_s part[s][co][tap*cinp + ci]
// A block of 256 threads owns 64 consecutive entries of the [COUTP][KPAD]
// partial layout, so every wave's loads are one contiguous 256-byte run;
// the 4 threads of an entry sum the splits congruent to their index mod 4,
// then combine in a fixed order (the same order for every entry, so the
// result is deterministic), and the weight (co, ci, tap) of a real entry
// receives the total.  Padding entries (co >= cout, ci >= cin_real, tap >=
// nt) are summed and dropped.
__global__ __launch_bounds__(256) void kc_dw_reduce(const float* __restrict__ part, int nsplit,
                                                    int coutp, int kpad, int cout, int cin_real,
                                                    int cinp, int nt, float* __restrict__ grad) {
  __shared__ float red[4][64];
  const int o = threadIdx.x % 64, g = threadIdx.x / 64;
  const size_t stride = (size_t)coutp * kpad;
  const size_t e = (size_t)blockIdx.x * 64 + o;
  float acc = 0.f;
  if (e < stride) {
    int s = g;
    for (; s + 12 < nsplit; s += 16)
      acc += (part[s * stride + e] + part[(s + 4) * stride + e]) +
             (part[(s + 8) * stride + e] + part[(s + 12) * stride + e]);
    for (; s < nsplit; s += 4) acc += part[s * stride + e];
  }
  red[g][o] = acc;
  __syncthreads();
  if (g == 0 && e < stride) {
    const int co = (int)(e / kpad), k = (int)(e % kpad), tap = k / cinp, ci = k % cinp;
    if (co < cout && tap < nt && ci < cin_real)
      grad[((size_t)co * cin_real + ci) * nt + tap] += (red[0][o] + red[1][o]) + (red[2][o] + red[3][o]);
  }
}

// ---------------------------------------------------------------------------
// ReLU backward and bias gradient:
//   gy[v][c] = bf16(dy[v][c] * (y[v][c] > 0))   (y == nullptr: no ReLU)
//   part[block][c] = sum over the block's voxels of gy (FP32, of the
//   bf16-rounded values)
__global__ __launch_bounds__(256) void kc_relu_bwd(int V, int C, const float* __restrict__ dy,
                                                   int dy_ld, int dy_off, const bf16_t* __restrict__ y,
                                                   int y_ld, int y_off, bf16_t* __restrict__ gy,
                                                   float* __restrict__ part) {
  // a thread owns 4 consecutive channels (one 16-byte load of dy); C is a
  // power of two (8..64), so every thread keeps its channels in the
  // grid-stride loop
  __shared__ float red[4][256];
  const int P = C >> 2, lp = __builtin_ctz((unsigned)P);
  const int tid = threadIdx.x, pc = tid & (P - 1);
  float s[4] = {0.f, 0.f, 0.f, 0.f};
  const int n = V * P;
  for (int q = blockIdx.x * 256 + tid; q < n; q += gridDim.x * 256) {
    const int v = q >> lp;
    const float4 d4 = *reinterpret_cast<const float4*>(dy + (size_t)v * dy_ld + dy_off + 4 * pc);
    float d[4] = {d4.x, d4.y, d4.z, d4.w};
    if (y) {
      const uint2 yy = *reinterpret_cast<const uint2*>(y + (size_t)v * y_ld + y_off + 4 * pc);
      const float yv[4] = {bf2f((bf16_t)(yy.x & 0xffffu)), bf2f((bf16_t)(yy.x >> 16)),
                           bf2f((bf16_t)(yy.y & 0xffffu)), bf2f((bf16_t)(yy.y >> 16))};
#pragma unroll
      for (int e = 0; e < 4; ++e)
        if (!(yv[e] > 0.f)) d[e] = 0.f;
    }
    bf16_t h[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      h[e] = f2bf(d[e]);
      s[e] += bf2f(h[e]);
    }
    *reinterpret_cast<uint2*>(gy + (size_t)q * 4) =
        uint2{(unsigned)h[0] | ((unsigned)h[1] << 16), (unsigned)h[2] | ((unsigned)h[3] << 16)};
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) red[e][tid] = s[e];
  __syncthreads();
  if (tid < C) {  // channel c: the threads that own it, in increasing order
    const int e = tid & 3, p = tid >> 2;
    float t = 0.f;
    for (int u = p; u < 256; u += P) t += red[e][u];
    part[(size_t)blockIdx.x * C + tid] = t;
  }
}

// db[c] += sum_b part[b][c]: one block of 256 threads per channel, then a
// fixed-order tree (deterministic)
__global__ __launch_bounds__(256) void kc_bias_reduce(const float* __restrict__ part, int nb, int C,
                                                      float* __restrict__ db) {
  __shared__ float red[256];
  const int c = blockIdx.x;
  float t = 0.f;
  for (int b = threadIdx.x; b < nb; b += 256) t += part[(size_t)b * C + c];
  red[threadIdx.x] = t;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[c] += red[0];
}

// ---------------------------------------------------------------------------
// 2x average pooling and nearest upsampling between a fine grid (X,Y,Z)
// and its coarse grid (X/2,Y/2,Z/2).
__device__ __forceinline__ int fine_of(int xc, int yc, int zc, int a, int Y, int Z) {
  return ((2 * xc + (a >> 2)) * Y + (2 * yc + ((a >> 1) & 1))) * Z + (2 * zc + (a & 1));
}

// The four kernels below move 16 bytes per thread and step: 8 bf16 or 4
// FP32 channels of one voxel (C, the leading dimensions and the channel
// offsets are multiples of 8), with 32-bit index arithmetic.  Per channel
// the arithmetic and its order are those of an element-wise loop.
__device__ __forceinline__ float bf_lo(unsigned w) { return bf2f((bf16_t)(w & 0xffffu)); }
__device__ __forceinline__ float bf_hi(unsigned w) { return bf2f((bf16_t)(w >> 16)); }
__device__ __forceinline__ unsigned bf_pack(float lo, float hi) {
  return (unsigned)f2bf(lo) | ((unsigned)f2bf(hi) << 16);
}

__global__ void kc_pool(int X, int Y, int Z, int C, const bf16_t* __restrict__ in, int in_ld,
                        int in_off, bf16_t* __restrict__ out) {
  const int Xc = X / 2, Yc = Y / 2, Zc = Z / 2, P = C / 8;
  const int n = Xc * Yc * Zc * P;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    const int pc = q % P, vc = q / P;
    const int zc = vc % Zc, yc = (vc / Zc) % Yc, xc = vc / (Zc * Yc);
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const u32x4 w = *reinterpret_cast<const u32x4*>(
          in + (size_t)fine_of(xc, yc, zc, a, Y, Z) * in_ld + in_off + 8 * pc);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        s[2 * e] += bf_lo(w[e]);
        s[2 * e + 1] += bf_hi(w[e]);
      }
    }
    u32x4 o;
#pragma unroll
    for (int e = 0; e < 4; ++e) o[e] = bf_pack(s[2 * e] * 0.125f, s[2 * e + 1] * 0.125f);
    *reinterpret_cast<u32x4*>(out + (size_t)vc * C + 8 * pc) = o;
  }
}

// d_in[fine][c] += d_out[coarse][c] / 8
__global__ void kc_pool_bwd(int X, int Y, int Z, int C, const float* __restrict__ dout,
                            float* __restrict__ din, int din_ld, int din_off) {
  const int Yc = Y / 2, Zc = Z / 2, P = C / 4;
  const int n = X * Y * Z * P;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    const int pc = q % P, v = q / P;
    const int z = v % Z, y = (v / Z) % Y, x = v / (Z * Y);
    const int vc = ((x / 2) * Yc + y / 2) * Zc + z / 2;
    const float4 d = *reinterpret_cast<const float4*>(dout + (size_t)vc * C + 4 * pc);
    float4* o = reinterpret_cast<float4*>(din + (size_t)v * din_ld + din_off + 4 * pc);
    float4 cur = *o;
    cur.x += 0.125f * d.x;
    cur.y += 0.125f * d.y;
    cur.z += 0.125f * d.z;
    cur.w += 0.125f * d.w;
    *o = cur;
  }
}

__global__ void kc_up(int X, int Y, int Z, int C, const bf16_t* __restrict__ in,
                      bf16_t* __restrict__ out, int out_ld, int out_off) {
  const int Yc = Y / 2, Zc = Z / 2, P = C / 8;
  const int n = X * Y * Z * P;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    const int pc = q % P, v = q / P;
    const int z = v % Z, y = (v / Z) % Y, x = v / (Z * Y);
    const int vc = ((x / 2) * Yc + y / 2) * Zc + z / 2;
    *reinterpret_cast<u32x4*>(out + (size_t)v * out_ld + out_off + 8 * pc) =
        *reinterpret_cast<const u32x4*>(in + (size_t)vc * C + 8 * pc);
  }
}

// d_in[coarse][c] = sum of the 8 children's d_out
__global__ void kc_up_bwd(int X, int Y, int Z, int C, const float* __restrict__ dout, int dout_ld,
                          int dout_off, float* __restrict__ din) {
  const int Xc = X / 2, Yc = Y / 2, Zc = Z / 2, P = C / 4;
  const int n = Xc * Yc * Zc * P;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    const int pc = q % P, vc = q / P;
    const int zc = vc % Zc, yc = (vc / Zc) % Yc, xc = vc / (Zc * Yc);
    float4 s = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const float4 d = *reinterpret_cast<const float4*>(
          dout + (size_t)fine_of(xc, yc, zc, a, Y, Z) * dout_ld + dout_off + 4 * pc);
      s.x += d.x;
      s.y += d.y;
      s.z += d.z;
      s.w += d.w;
    }
    *reinterpret_cast<float4*>(din + (size_t)vc * C + 4 * pc) = s;
  }
}

// ---------------------------------------------------------------------------
// Mean squared error over V voxels x 8 channels; dout = scale * 2 (out - t)
// with scale = 1 / (8 V batch); part[block] = sum (out - t)^2 (FP64).
__global__ __launch_bounds__(256) void kc_loss(long long n, const float* __restrict__ out,
                                               const bf16_t* __restrict__ tgt, float scale,
                                               float* __restrict__ dout, double* __restrict__ part) {
  __shared__ double red[256];
  double s = 0.0;
  for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long long)gridDim.x * 256) {
    const float d = out[q] - bf2f(tgt[q]);
    dout[q] = 2.f * scale * d;
    s += (double)d * d;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// *acc += scale * sum_b part[b]: one block of 256 threads, fixed-order tree
__global__ __launch_bounds__(256) void kc_loss_reduce(const double* __restrict__ part, int nb,
                                                      double scale, double* __restrict__ acc) {
  __shared__ double red[256];
  double t = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) t += part[b];
  red[threadIdx.x] = t;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *acc += red[0] * scale;
}

// ---------------------------------------------------------------------------
// Parameters, gradients and data movement
__global__ void kc_fill(float* __restrict__ p, long long n, float v) {
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (long long)gridDim.x * blockDim.x)
    p[q] = v;
}

// Adam (Kingma & Ba 2015) on the flat FP32 parameters; g is the summed
// gradient of the batch, gscale = 1/batch; bc1 = 1 - b1^t, bc2 = 1 - b2^t.
__global__ void kc_adam(long long n, float* __restrict__ p, const float* __restrict__ g,
                        float* __restrict__ m, float* __restrict__ v, float lr, float b1, float b2,
                        float eps, float bc1, float bc2, float gscale) {
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (long long)gridDim.x * blockDim.x) {
    const float gq = g[q] * gscale;
    const float mq = b1 * m[q] + (1.f - b1) * gq;
    const float vq = b2 * v[q] + (1.f - b2) * gq * gq;
    m[q] = mq;
    v[q] = vq;
    p[q] -= lr * (mq / bc1) / (sqrtf(vq / bc2) + eps);
  }
}

// Kernel-layout bf16 weights of one layer from the flat FP32 parameters
// w[(co*cin_real + ci)*nt + tap]:
//   forward  wf[co][tap*cinp + ci]             rows coutp, K kpad
//   backward wt[ci][tap'*cout + co] = w[co][ci][nt-1-tap']   rows cinp_t, K kpad_t
__global__ void kc_wconv(const float* __restrict__ w, int cout, int cin_real, int cinp, int nt,
                         int coutp, int kpad, bf16_t* __restrict__ wf, int cinp_t, int kpad_t,
                         bf16_t* __restrict__ wt) {
  const int nf = coutp * kpad, nb = cinp_t * kpad_t;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nf + nb; q += gridDim.x * blockDim.x) {
    if (q < nf) {
      const int co = q / kpad, kk = q % kpad, tap = kk / cinp, ci = kk % cinp;
      float val = 0.f;
      if (co < cout && tap < nt && ci < cin_real) val = w[((size_t)co * cin_real + ci) * nt + tap];
      wf[q] = f2bf(val);
    } else if (wt) {
      const int e = q - nf;
      const int ci = e / kpad_t, kk = e % kpad_t, tp = kk / cout, co = kk % cout;
      float val = 0.f;
      if (ci < cin_real && tp < nt) val = w[((size_t)co * cin_real + ci) * nt + (nt - 1 - tp)];
      wt[e] = f2bf(val);
    }
  }
}

// f32 -> bf16 with channel padding: dst[v][0..cdst) = src[v][0..csrc), 0 beyond
__global__ void kc_pack_bf16(long long V, int csrc, int cdst, const float* __restrict__ src,
                             bf16_t* __restrict__ dst) {
  const long long n = V * cdst;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (long long)gridDim.x * blockDim.x) {
    const long long v = q / cdst;
    const int c = (int)(q % cdst);
    dst[q] = c < csrc ? f2bf(src[v * csrc + c]) : (bf16_t)0;
  }
}

// dst[v][0..cdst) = src[v][0..csrc) (bf16), zero channels beyond csrc
__global__ void kc_widen(long long V, int csrc, int cdst, const bf16_t* __restrict__ src,
                         bf16_t* __restrict__ dst) {
  const long long n = V * cdst;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (long long)gridDim.x * blockDim.x) {
    const long long v = q / cdst;
    const int c = (int)(q % cdst);
    dst[q] = c < csrc ? src[v * csrc + c] : (bf16_t)0;
  }
}
