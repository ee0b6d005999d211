// reduce.hpp -- wave64 / block / grid reductions with FP64 accumulation and
// the device-resident QMR scalar state.
//
// Grid reductions use the last-block-finalises pattern: every block writes
// its partial sums to part[block], publishes them with an agent-scope fence
// and takes a ticket; the block holding the last ticket sums the partials in
// a fixed order (deterministic, independent of scheduling), applies the
// scalar recurrence and resets the ticket.  No second launch, no host
// round-trip.
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

__device__ __forceinline__ double ld_agent(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Publish this block's NV partials (valid in thread 0) and return true in
// every thread of the block that finishes last; that block then holds the
// fixed-order grid totals in tot[] (all threads).
template <int NV>
__device__ __forceinline__ bool grid_finish(double (&v)[NV], double* part, unsigned* ticket,
                                            double (&tot)[NV]) {
  __shared__ int am_last;
  __shared__ double sh_tot[NV];
  const unsigned nb = gridDim.x * gridDim.y * gridDim.z;
  const unsigned bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) st_agent(&part[(size_t)bid * NV + q], v[q]);
    __threadfence();
    unsigned t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    am_last = (t == nb - 1);
  }
  __syncthreads();
  if (!am_last) return false;
  __threadfence();
  // fixed-order two-level sum: thread t sums blocks t, t+T, ...; then block_sum
  double s[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) s[q] = 0.0;
  for (unsigned b = threadIdx.x; b < nb; b += blockDim.x)
#pragma unroll
    for (int q = 0; q < NV; ++q) s[q] += ld_agent(&part[(size_t)b * NV + q]);
  block_sum<NV>(s);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) sh_tot[q] = s[q];
    *ticket = 0u;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NV; ++q) tot[q] = sh_tot[q];
  return true;
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
  int hist_cap, pad0;
  unsigned ticket, pad1;
};

enum : int { ST_CONV = 0, ST_MAXIT = 1, ST_BREAK = 2 };

__device__ __forceinline__ bool finite_c(c128 a) { return isfinite(a.re) && isfinite(a.im); }
