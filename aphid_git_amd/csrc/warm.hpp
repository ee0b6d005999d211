// warm.hpp -- initial guesses for the QMR solve: the network's cell-grid
// output mapped onto the staggered unknowns, and the amplitudes of the
// candidate guesses, fitted by a Galerkin condition or by least squares on
// the residual.
#pragma once
#include "reduce.hpp"

// Unknown (c, node) takes channels (2c, 2c+1) of the cell it belongs to:
// each cell (i, j, k) owns the x-, y- and z-edges leaving its lower corner
// node (i, j, k) and that node's Phi.  Edges on the last j / k planes (no
// cell of their own) take the nearest cell across; Phi on boundary nodes
// (whose operator rows are zero) and the padding entries of the edge
// components stay 0.  Device layout: 4 blocks of nn = nx*ny*nz, z fastest;
// cells (nx-1, ny-1, nz-1), z fastest.
__global__ void k_cells_to_unknowns(int nx, int ny, int nz, const float* __restrict__ out,
                                    c128* __restrict__ x) {
  const long long nn = (long long)nx * ny * nz;
  const int cx = nx - 1, cy = ny - 1, cz = nz - 1;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < 4 * nn;
       q += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(q / nn);
    const long long node = q - c * nn;
    const int k = (int)(node % nz), j = (int)((node / nz) % ny), i = (int)(node / ((long long)nz * ny));
    bool valid;
    if (c == 3)
      valid = i > 0 && i < nx - 1 && j > 0 && j < ny - 1 && k > 0 && k < nz - 1;
    else
      valid = (c == 0 ? i : c == 1 ? j : k) < (c == 0 ? cx : c == 1 ? cy : cz);
    c128 v = {0.0, 0.0};
    if (valid) {
      const long long cell = ((long long)min(i, cx - 1) * cy + min(j, cy - 1)) * cz + min(k, cz - 1);
      v = {(double)out[cell * 8 + 2 * c], (double)out[cell * 8 + 2 * c + 1]};
    }
    x[q] = v;
  }
}

constexpr int kWsMax = 5;  // candidate vectors in one fit
constexpr int kWsPairs = kWsMax * (kWsMax + 1) / 2 + kWsMax;

struct GramArgs {
  const c128* w[kWsMax];  // K u_a
  const c128* u[kWsMax];  // u_a, restricted to block ublk[a] (-1: whole vector)
  int ublk[kWsMax];
  const c128* b;
  int nb;
  long long n, nn;
};

// Per-block partials of the fit's normal equations, packed as the upper
// triangle (a <= b, row-major) of G then h:
//   GALERKIN: G_ab = u_a^T K u_b, h_a = u_a^T b (bilinear, as QMR's own
//             inner product; K is complex symmetric so G is symmetric);
//   else      G_ab = (K u_a)^H K u_b, h_a = (K u_a)^H b (least squares on
//             the residual).
template <bool GALERKIN>
__global__ __launch_bounds__(256) void k_gram(GramArgs g, double* part) {
  double acc[2 * kWsPairs];
#pragma unroll
  for (int q = 0; q < 2 * kWsPairs; ++q) acc[q] = 0.0;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < g.n;
       q += (long long)gridDim.x * blockDim.x) {
    const int blk = (int)(q / g.nn);
    c128 wv[kWsMax], lv[kWsMax];
#pragma unroll
    for (int a = 0; a < kWsMax; ++a) {
      wv[a] = a < g.nb ? g.w[a][q] : c128{0.0, 0.0};
      if (GALERKIN)
        lv[a] = (a < g.nb && (g.ublk[a] < 0 || g.ublk[a] == blk)) ? g.u[a][q] : c128{0.0, 0.0};
      else
        lv[a] = {wv[a].re, -wv[a].im};  // conj(K u_a)
    }
    const c128 bv = g.b[q];
    int p = 0;
#pragma unroll
    for (int a = 0; a < kWsMax; ++a)
#pragma unroll
      for (int b = a; b < kWsMax; ++b, ++p) {
        acc[2 * p] += lv[a].re * wv[b].re - lv[a].im * wv[b].im;
        acc[2 * p + 1] += lv[a].re * wv[b].im + lv[a].im * wv[b].re;
      }
#pragma unroll
    for (int a = 0; a < kWsMax; ++a, ++p) {
      acc[2 * p] += lv[a].re * bv.re - lv[a].im * bv.im;
      acc[2 * p + 1] += lv[a].re * bv.im + lv[a].im * bv.re;
    }
  }
  block_sum<2 * kWsPairs>(acc);
  block_partial<2 * kWsPairs>(acc, part);
}

// x = sum_a al[a] u[a] over whole vectors (a < nb)
struct CombineArgs {
  const c128* u[kWsMax];
  c128 al[kWsMax];
  int nb;
};
__global__ void k_combine(long long n, CombineArgs g, c128* __restrict__ x) {
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (long long)gridDim.x * blockDim.x) {
    c128 v = {0.0, 0.0};
    for (int a = 0; a < g.nb; ++a) v += g.al[a] * g.u[a][q];
    x[q] = v;
  }
}

struct WsCoef {
  c128 a[4];  // amplitude of each component of the mapped prediction
  c128 prev;  // amplitude of the previous solution
  int use_net, use_prev;
};

// x = sum_c a_c u_c + prev * xp   (u_c: block c of u)
__global__ void k_ws_combine(long long nn, const c128* __restrict__ u,
                             const c128* __restrict__ xp, c128* __restrict__ x, WsCoef cf) {
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < 4 * nn;
       q += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(q / nn);
    c128 v = {0.0, 0.0};
    if (cf.use_net) v = cf.a[c] * u[q];
    if (cf.use_prev) v += cf.prev * xp[q];
    x[q] = v;
  }
}
