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

// x[q] = d[q] * x[q] (the Jacobi weight W = M^-1 = 1/diag)
__global__ void k_scale_diag(long long n, const c128* __restrict__ d, c128* __restrict__ x) {
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (long long)gridDim.x * blockDim.x)
    x[q] = d[q] * x[q];
}

// Residual of the fitted guess in the W-weighted norm (W = 1/diag, or the
// identity when invd is null):
//   r = W b - sum_c al_c W K P_c u  (w[c] = W K P_c u, bw = W b),
// y = W conj(r) (the input of the adjoint apply), and per-block partials of
// ||r||^2 and ||W b||^2.
__global__ __launch_bounds__(256) void k_wres(long long n, CombineArgs g,
                                              const c128* __restrict__ bw,
                                              const c128* __restrict__ invd,
                                              c128* __restrict__ y, double* part) {
  double acc[2] = {0.0, 0.0};
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (long long)gridDim.x * blockDim.x) {
    const c128 b = bw[q];
    c128 r = b;
    for (int a = 0; a < g.nb; ++a) {
      const c128 t = g.al[a] * g.u[a][q];
      r = {r.re - t.re, r.im - t.im};
    }
    y[q] = invd ? invd[q] * c128{r.re, -r.im} : c128{r.re, -r.im};
    acc[0] += abs2(r);
    acc[1] += abs2(b);
  }
  block_sum<2>(acc);
  block_partial<2>(acc, part);
}

// d loss / d output of the network for the loss L = ||r||^2 / ||W b||^2 of
// the fitted guess x0 = sum_c al_c P_c u, u = the cells mapped onto the
// unknowns (k_cells_to_unknowns).  With z = K W conj(r) (K complex
// symmetric), dL = -(2/||Wb||^2) Re(sum_q al_c(q) z_q du_q), so
//   dL/dRe u_q = -s Re(al_c z_q),  dL/dIm u_q = s Im(al_c z_q),
// s = 2 / (||Wb||^2 batch).  Each cell gathers the unknowns the mapping
// fed from it (its lower corner node; on the last cell planes also the
// nodes beyond).  dout[cell][2c], dout[cell][2c+1] are overwritten.
__global__ void k_adj_cells(int nx, int ny, int nz, const c128* __restrict__ z, CombineArgs g,
                            double s, float* __restrict__ dout) {
  const long long nn = (long long)nx * ny * nz;
  const int cx = nx - 1, cy = ny - 1, cz = nz - 1;
  const long long ncell = (long long)cx * cy * cz;
  for (long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x; cell < ncell;
       cell += (long long)gridDim.x * blockDim.x) {
    const int k = (int)(cell % cz), j = (int)((cell / cz) % cy), i = (int)(cell / ((long long)cz * cy));
    double gr[4] = {0.0, 0.0, 0.0, 0.0}, gi[4] = {0.0, 0.0, 0.0, 0.0};
    for (int di = 0; di <= (i == cx - 1 ? 1 : 0); ++di)
      for (int dj = 0; dj <= (j == cy - 1 ? 1 : 0); ++dj)
        for (int dk = 0; dk <= (k == cz - 1 ? 1 : 0); ++dk) {
          const int ii = i + di, jj = j + dj, kk = k + dk;
          const long long node = ((long long)ii * ny + jj) * nz + kk;
          for (int c = 0; c < 4; ++c) {
            bool valid;
            if (c == 3)
              valid = ii > 0 && ii < nx - 1 && jj > 0 && jj < ny - 1 && kk > 0 && kk < nz - 1;
            else
              valid = (c == 0 ? ii : c == 1 ? jj : kk) < (c == 0 ? cx : c == 1 ? cy : cz);
            if (!valid) continue;
            const c128 t = g.al[c] * z[c * nn + node];
            gr[c] -= s * t.re;
            gi[c] += s * t.im;
          }
        }
    for (int c = 0; c < 4; ++c) {
      dout[cell * 8 + 2 * c] = (float)gr[c];
      dout[cell * 8 + 2 * c + 1] = (float)gi[c];
    }
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
