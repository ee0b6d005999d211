// stencil.hpp -- the matrix-free A-Phi operator K u on the device.
//
// One fused kernel computes all four row blocks of the reference's
// inprod_x / inprod_y / inprod_z / inprod_p (src/aphid.f90:803-1282) for a
// 2-D tile of (y, z) columns, marching along x:
//   - device layout is z-fastest, idx = (i*ny + j)*nz + k on the common
//     node grid; each component (Ax, Ay, Az, Phi) and each k^2 grid is a
//     separate nn = nx*ny*nz block; the entries beyond an edge component's
//     extent (i = nx-1 for Ax, j = ny-1 for Ay, k = nz-1 for Az) are zero
//     padding and stay zero;
//   - a block owns TJ x TK columns and `chunk` x-planes; lanes run along z,
//     so every global access is a coalesced run of TK complex values;
//   - the x-neighbours (i-1, i+1) of every field live in a per-thread
//     register queue; the in-plane neighbours (j+-1, k+-1) come from the
//     current plane staged in LDS with a one-cell halo (2*(TJ+TK) halo
//     cells, corners are never needed by a 7-point stencil);
//   - outside the domain the staged value is 0, which is what the
//     reference's clamped-index-times-zero-coefficient terms contribute.
#pragma once
#include "reduce.hpp"

template <typename R>
struct OpArgs {
  int nx, ny, nz;
  long long nn;               // nx*ny*nz, stride between component blocks
  const cplx<R>* k2;          // 4*nn: k2x, k2y, k2z, k2p (device layout)
  // 1-D tables per axis a (0=x,1=y,2=z), 0-based:
  //  edge e in [0,n-2] between nodes e and e+1 (Fortran edge e+1):
  //   eP[e] = (E/(n-1)-1)/dc(E+1)/d(E),  eM[e] = ((n-E)/(n-1)-1)/dc(E)/d(E)
  //   dE[e] = d(E), rdE[e] = 1/d(E)                      (E = e+1)
  //  node q in [0,n-1] (Fortran node Q = q+1):
  //   nP[q] = (Q/n-1)/d(Q)/dc(Q),  nM[q] = ((n-Q)/(n-1)-1)/d(Q-1)/dc(Q)
  //   dN[q] = dc(Q), rdN[q] = 1/dc(Q)
  // i.e. gen_inprod_{x,y,z,p}_coefs (src/aphid.f90:750-797,878-925,
  // 1007-1054,1135-1185) with the integer-division wall masks.
  const R* eP[3];
  const R* eM[3];
  const R* nP[3];
  const R* nM[3];
  const R* dE[3];
  const R* rdE[3];
  const R* dN[3];
  const R* rdN[3];
};

// Optional epilogue of the apply: eps = in^T out (bilinear, FP64) reduced
// over the grid, then the QMR scalars beta = eps/del (src/aphid_qmr.f90:329-330).
struct ApplyEpilogue {
  QmrState* st;     // nullptr: plain apply
  double* part;     // per-block partials (2 doubles per block)
};

// 32-bit byte offsets from a base pointer: lets the compiler address every
// array as SGPR base + one 32-bit VGPR offset (all buffers are < 4 GiB)
template <typename T>
__device__ __forceinline__ T ldo(const T* __restrict__ base, unsigned boff) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + boff);
}
template <typename T>
__device__ __forceinline__ void sto(T* __restrict__ base, unsigned boff, T v) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + boff) = v;
}

template <typename R, int TK, int TJ>
__global__ __launch_bounds__(TK* TJ) void k_apply(OpArgs<R> op, const cplx<R>* __restrict__ in,
                                                  cplx<R>* __restrict__ out, int chunk,
                                                  ApplyEpilogue ep) {
  using C = cplx<R>;
  constexpr int LK = TK + 2, LJ = TJ + 2, LS = LK * LJ;
  constexpr int NH = 2 * TK + 2 * TJ;
  constexpr unsigned CS = sizeof(C);
  static_assert(NH <= TK * TJ, "halo larger than the block");
  __shared__ C sm[7][LS];  // Ax, Ay, Az, Phi, k2y, k2z, k2p of plane i
  // per-tile 1-D tables: [0]nP [1]nM [2]eP [3]eM [4]dN [5]dE [6]rdN [7]rdE
  __shared__ R tyj[8][TJ];
  __shared__ R tzk[8][TK];

  if (ep.st && ep.st->done) return;

  const int nx = op.nx, ny = op.ny, nz = op.nz;
  const unsigned pl = (unsigned)ny * nz;
  const unsigned nnb = (unsigned)op.nn * CS;  // component stride in bytes
  const int tk = threadIdx.x % TK, tj = threadIdx.x / TK;
  const int k = blockIdx.x * TK + tk, j = blockIdx.y * TJ + tj;
  const int i0 = blockIdx.z * chunk;
  const int i1 = min(nx, i0 + chunk);
  const bool own = (j < ny) && (k < nz);
  const unsigned colb = own ? ((unsigned)j * nz + k) * CS : 0u;
  const unsigned plb = pl * CS;
  const C zero = {R(0), R(0)};

  if (threadIdx.x < TJ) {
    const int jj = blockIdx.y * TJ + threadIdx.x;
    const bool v = jj < ny;
    tyj[0][threadIdx.x] = v ? op.nP[1][jj] : R(0);
    tyj[1][threadIdx.x] = v ? op.nM[1][jj] : R(0);
    tyj[2][threadIdx.x] = v ? op.eP[1][jj] : R(0);
    tyj[3][threadIdx.x] = v ? op.eM[1][jj] : R(0);
    tyj[4][threadIdx.x] = v ? op.dN[1][jj] : R(0);
    tyj[5][threadIdx.x] = v ? op.dE[1][jj] : R(0);
    tyj[6][threadIdx.x] = v ? op.rdN[1][jj] : R(0);
    tyj[7][threadIdx.x] = v ? op.rdE[1][jj] : R(0);
  }
  if (threadIdx.x < TK) {
    const int kk = blockIdx.x * TK + threadIdx.x;
    const bool v = kk < nz;
    tzk[0][threadIdx.x] = v ? op.nP[2][kk] : R(0);
    tzk[1][threadIdx.x] = v ? op.nM[2][kk] : R(0);
    tzk[2][threadIdx.x] = v ? op.eP[2][kk] : R(0);
    tzk[3][threadIdx.x] = v ? op.eM[2][kk] : R(0);
    tzk[4][threadIdx.x] = v ? op.dN[2][kk] : R(0);
    tzk[5][threadIdx.x] = v ? op.dE[2][kk] : R(0);
    tzk[6][threadIdx.x] = v ? op.rdN[2][kk] : R(0);
    tzk[7][threadIdx.x] = v ? op.rdE[2][kk] : R(0);
  }

  // own value of component c (0..3) or k2 grid c (0..3) at plane i
  auto ldc = [&](const C* __restrict__ base, int c, int i) -> C {
    return (own && i >= 0 && i < nx) ? ldo(base, c * nnb + (unsigned)i * plb + colb) : zero;
  };

  // halo cell of this thread (if any): rows j=-1 and j=TJ over the tile's
  // k range, then columns k=-1 and k=TK over its j range
  int hj = 0, hk = 0;
  const bool hv = threadIdx.x < NH;
  if (hv) {
    const int h = threadIdx.x;
    if (h < TK) { hj = -1; hk = h; }
    else if (h < 2 * TK) { hj = TJ; hk = h - TK; }
    else if (h < 2 * TK + TJ) { hj = h - 2 * TK; hk = -1; }
    else { hj = h - 2 * TK - TJ; hk = TK; }
  }
  const int gj = blockIdx.y * TJ + hj, gk = blockIdx.x * TK + hk;
  const bool hin = hv && gj >= 0 && gj < ny && gk >= 0 && gk < nz;
  const unsigned hcolb = hin ? ((unsigned)gj * nz + gk) * CS : 0u;
  const int hl = (hj + 1) * LK + (hk + 1);
  const int ol = (tj + 1) * LK + (tk + 1);

  C pa[4], ca[4], na[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    pa[c] = ldc(in, c, i0 - 1);
    ca[c] = ldc(in, c, i0);
  }
  const C* __restrict__ K2 = op.k2;
  C k2x_p = ldc(K2, 0, i0 - 1), k2x_c = ldc(K2, 0, i0), k2p_c = ldc(K2, 3, i0);

  double eacc_re = 0.0, eacc_im = 0.0;

  for (int i = i0; i < i1; ++i) {
    // plane i+1 of the own column: the x-neighbours of this step
#pragma unroll
    for (int c = 0; c < 4; ++c) na[c] = ldc(in, c, i + 1);
    const C k2x_n = ldc(K2, 0, i + 1), k2p_n = ldc(K2, 3, i + 1);

    // stage plane i: own values, then the halo
#pragma unroll
    for (int c = 0; c < 4; ++c) sm[c][ol] = ca[c];
    sm[4][ol] = ldc(K2, 1, i);
    sm[5][ol] = ldc(K2, 2, i);
    sm[6][ol] = k2p_c;
    if (hv) {
      const unsigned g = (unsigned)i * plb + hcolb;
#pragma unroll
      for (int c = 0; c < 4; ++c) sm[c][hl] = hin ? ldo(in, c * nnb + g) : zero;
#pragma unroll
      for (int c = 1; c < 4; ++c) sm[3 + c][hl] = hin ? ldo(K2, c * nnb + g) : zero;
    }
    __syncthreads();

    if (own) {
      const unsigned ob = (unsigned)i * plb + colb;
      const R nPx = op.nP[0][i], nMx = op.nM[0][i], dNx = op.dN[0][i];
      const R nPy = tyj[0][tj], nMy = tyj[1][tj], nPz = tzk[0][tk], nMz = tzk[1][tk];
      const R dNy = tyj[4][tj], dNz = tzk[4][tk];
      const C k2y_c = sm[4][ol], k2z_c = sm[5][ol];
      C rx = zero, ry = zero, rz = zero, rp = zero;
      // ---- x-edge row, inprod_x (src/aphid.f90:841-866)
      if (i < nx - 1) {
        const R cE = op.eP[0][i], cW = op.eM[0][i];
        const R sm6 = cE + cW + nPy + nMy + nPz + nMz;
        const C c1 = {k2x_c.re - sm6, k2x_c.im};
        C acc = c1 * ca[0];
        acc += cE * na[0];
        acc += cW * pa[0];
        acc += nPy * sm[0][ol + LK];
        acc += nMy * sm[0][ol - LK];
        acc += nPz * sm[0][ol + 1];
        acc += nMz * sm[0][ol - 1];
        const C cpl = na[3] * (k2p_n - k2x_c) - ca[3] * (k2p_c - k2x_c);
        const R vol = op.dE[0][i] * dNy * dNz;
        rx = (acc + cpl * op.rdE[0][i]) * vol;
      }
      sto(out, ob, rx);
      // ---- y-edge row, inprod_y (src/aphid.f90:970-995)
      if (j < ny - 1) {
        const R cN = tyj[2][tj], cS = tyj[3][tj];
        const R sm6 = cN + cS + nPx + nMx + nPz + nMz;
        const C c1 = {k2y_c.re - sm6, k2y_c.im};
        C acc = c1 * ca[1];
        acc += cN * sm[1][ol + LK];
        acc += cS * sm[1][ol - LK];
        acc += nPx * na[1];
        acc += nMx * pa[1];
        acc += nPz * sm[1][ol + 1];
        acc += nMz * sm[1][ol - 1];
        const C cpl = sm[3][ol + LK] * (sm[6][ol + LK] - k2y_c) - ca[3] * (k2p_c - k2y_c);
        const R vol = dNx * tyj[5][tj] * dNz;
        ry = (acc + cpl * tyj[7][tj]) * vol;
      }
      sto(out, nnb + ob, ry);
      // ---- z-edge row, inprod_z (src/aphid.f90:1098-1123)
      if (k < nz - 1) {
        const R cU = tzk[2][tk], cD = tzk[3][tk];
        const R sm6 = cU + cD + nPx + nMx + nPy + nMy;
        const C c1 = {k2z_c.re - sm6, k2z_c.im};
        C acc = c1 * ca[2];
        acc += cU * sm[2][ol + 1];
        acc += cD * sm[2][ol - 1];
        acc += nPx * na[2];
        acc += nMx * pa[2];
        acc += nPy * sm[2][ol + LK];
        acc += nMy * sm[2][ol - LK];
        const C cpl = sm[3][ol + 1] * (sm[6][ol + 1] - k2z_c) - ca[3] * (k2p_c - k2z_c);
        const R vol = dNx * dNy * tzk[5][tk];
        rz = (acc + cpl * tzk[7][tk]) * vol;
      }
      sto(out, 2 * nnb + ob, rz);
      // ---- interior node row, inprod_p (src/aphid.f90:1229-1278)
      if (j > 0 && j < ny - 1 && k > 0 && k < nz - 1 && i > 0 && i < nx - 1) {
        const C s1 = k2p_c, s2 = k2x_c, s3 = k2x_p;
        const C s4 = k2y_c, s5 = sm[4][ol - LK];
        const C s6 = k2z_c, s7 = sm[5][ol - 1];
        const C c2 = nPx * s2, c3 = nMx * s3, c4 = nPy * s4;
        const C c5 = nMy * s5, c6 = nPz * s6, c7 = nMz * s7;
        const C c1 = s1 * s1 - (c2 + c3 + c4 + c5 + c6 + c7);
        C acc = c1 * ca[3];
        acc += c2 * na[3];
        acc += c3 * pa[3];
        acc += c4 * sm[3][ol + LK];
        acc += c5 * sm[3][ol - LK];
        acc += c6 * sm[3][ol + 1];
        acc += c7 * sm[3][ol - 1];
        const R rx_ = op.rdN[0][i], ry_ = tyj[6][tj], rz_ = tzk[6][tk];
        C cp = ((s2 - s1) * rx_) * ca[0];
        cp += (-((s3 - s1) * rx_)) * pa[0];
        cp += ((s4 - s1) * ry_) * ca[1];
        cp += (-((s5 - s1) * ry_)) * sm[1][ol - LK];
        cp += ((s6 - s1) * rz_) * ca[2];
        cp += (-((s7 - s1) * rz_)) * sm[2][ol - 1];
        const R vol = dNx * dNy * dNz;
        rp = acc * vol + cp * vol;
      }
      sto(out, 3 * nnb + ob, rp);
      if (ep.st) {  // eps partial: sum_c in_c * out_c in FP64
        const cplx<double> a0 = ccast<double>(ca[0]) * ccast<double>(rx);
        const cplx<double> a1 = ccast<double>(ca[1]) * ccast<double>(ry);
        const cplx<double> a2 = ccast<double>(ca[2]) * ccast<double>(rz);
        const cplx<double> a3 = ccast<double>(ca[3]) * ccast<double>(rp);
        eacc_re += (a0.re + a1.re) + (a2.re + a3.re);
        eacc_im += (a0.im + a1.im) + (a2.im + a3.im);
      }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      pa[c] = ca[c];
      ca[c] = na[c];
    }
    k2x_p = k2x_c;
    k2x_c = k2x_n;
    k2p_c = k2p_n;
  }

  if (ep.st) {
    double v[2] = {eacc_re, eacc_im};
    block_sum<2>(v);
    block_partial<2>(v, ep.part);
  }
}

// eps = p^T K p (the partials of k_apply), beta = eps/del (src/aphid_qmr.f90:329-330)
__global__ void k_fin_eps(const double* part, int nb, QmrState* st) {
  if (st->done) return;
  double tot[2];
  finalize_sum<2>(part, nb, tot);
  if (threadIdx.x == 0) {
    st->eps = {tot[0], tot[1]};
    st->beta = cdiv(st->eps, st->del);
    if (is_zero(st->eps) || !finite_c(st->beta)) {
      st->status = ST_BREAK;
      st->done = 1;
    }
  }
}
