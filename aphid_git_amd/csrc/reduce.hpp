// reduce.hpp -- wave64 / block / grid reductions with FP64 accumulation and
// the device-resident QMR scalar state.
//
// Grid reductions are two launches: every block of the streaming kernel
// writes its partial sums to part[block] (plain stores), and a one-block
// finalize kernel sums them in a fixed order (deterministic, independent of
// scheduling) and applies the scalar recurrence on the device.  The kernel
// boundary makes the partials visible; an in-kernel last-block scheme would
// need an agent-scope release fence per block, which on this GPU writes back
// the L2 and costs more than the extra launch.
#pragma once
#include "cplx.hpp"

constexpr int kWave = 64;

template <int NV>
__device__ __forceinline__ void wave_sum(double (&v)[NV]) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1)
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] += __shfl_xor(v[q], off, kWave);
}

// Sum NV doubles over the block; the result is valid in thread 0.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
  __shared__ double red[NV][16];  // up to 1024 threads = 16 waves
  wave_sum<NV>(v);
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  const int nw = (blockDim.x + kWave - 1) / kWave;
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < NV; ++q) red[q][w] = v[q];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double s = 0.0;
      for (int t = 0; t < nw; ++t) s += red[q][t];
      v[q] = s;
    }
  }
}

// Publish this block's NV partials (valid in thread 0) to part[block].
template <int NV>
__device__ __forceinline__ void block_partial(const double (&v)[NV], double* part) {
  if (threadIdx.x == 0) {
    const unsigned bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
#pragma unroll
    for (int q = 0; q < NV; ++q) part[(size_t)bid * NV + q] = v[q];
  }
}

// In a one-block finalize kernel: the fixed-order total of nb partials
// (thread t sums blocks t, t+T, ...; then a block tree), valid in thread 0.
// The loads of up to 8 partials per thread are issued together before any
// is added (one memory round trip for 2048 partials at 256 threads instead
// of one per partial); the additions keep the order b = t, t+T, t+2T, ...
template <int NV>
__device__ __forceinline__ void finalize_sum(const double* __restrict__ part, int nb, double (&tot)[NV]) {
  constexpr int U = 8;
#pragma unroll
  for (int q = 0; q < NV; ++q) tot[q] = 0.0;
  const int T = blockDim.x;
  for (int b0 = threadIdx.x; b0 < nb; b0 += U * T) {
    double x[U][NV];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int b = b0 + u * T;
#pragma unroll
      for (int q = 0; q < NV; ++q) x[u][q] = b < nb ? part[(size_t)b * NV + q] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (b0 + u * T < nb)
#pragma unroll
        for (int q = 0; q < NV; ++q) tot[q] += x[u][q];
  }
  block_sum<NV>(tot);
}

// Scalars of the coupled two-term recurrence (src/aphid_qmr.f90:63-65),
// kept in device memory for the whole solve.
struct QmrState {
  c128 del, eps, beta, eta, k1;
  double rho_n, rho_np;       // ||vt|| of the current and next Lanczos vector
  double c_n, c_nm;           // c_n and c_{n-1}
  double theta_n, theta_nm;   // theta_n and theta_{n-1}
  double k2s;                 // theta_{n-1} c_n
  double rnorm, rhsnorm, target;
  int iter, max_it, status, done;
  int hist_cap;
  int lag;  // 1: the last update left r one iteration behind (k_update)
  // the d and x updates of iteration m run in the p-update of iteration
  // m+1 (k_pupdate): upd_it = the last iteration whose fused update ran,
  // dx_done = the last iteration whose d and x updates ran
  int upd_it, dx_done;
};

enum : int { ST_CONV = 0, ST_MAXIT = 1, ST_BREAK = 2 };

__host__ __device__ __forceinline__ bool finite_c(c128 a) { return std::isfinite(a.re) && std::isfinite(a.im); }
