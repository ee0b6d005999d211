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
//     cells, corners are never needed by a 7-point stencil; a halo cell
//     holds the four fields and the one k^2 grid its side needs);
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
  // Compact k^2 (complex128 x-march; nullptr: off).  When every cell's
  // sigma-hat has the same imaginary part e (model R: sigma + i omega eps0),
  // the real parts of the edge grids depend on two coordinates only (the
  // transverse metric weights of gen_k2_{x,y,z}grid, src/aphid.f90:1306-1440,
  // times e), and the real part of the node grid k2p is -omega mu0 e times
  // a weighted mean that is 1 up to rounding (gen_k2_pgrid, :1465-1531): its
  // values lie within a few units in the last place of each other.  The
  // stencil then reads two 16-byte pairs per node, k2rec[2q .. 2q+1] =
  // {k2x.im, k2p.im} and, k2yz_off bytes further, the pair {k2y.im, k2z.im},
  // and one byte k2pc[q]; k2p's real part is the double with the bit
  // pattern k2p_base + k2pc[q], and the edge grids' real parts come from
  // the 2-D tables k2re[0](j,k) = [j*nz + k],
  // k2re[1](i,k) = [i*nz + k], k2re[2](i,j) = [i*ny + j].  Every real part
  // is checked against the stored grids bit for bit before the compact form
  // is used (k_k2c_check), so the stencil sees the same complex values
  // either way.
  const R* k2rec = nullptr;
  unsigned k2yz_off = 0;
  const R* k2re[3] = {nullptr, nullptr, nullptr};
  const unsigned char* k2pc = nullptr;
  unsigned long long k2p_base = 0;
};

// Optional epilogue of the apply: eps = in^T out (bilinear, FP64) reduced
// over the grid, then the QMR scalars beta = eps/del (src/aphid_qmr.f90:329-330).
struct ApplyEpilogue {
  QmrState* st;     // nullptr: plain apply
  double* part;     // per-block partials (2 doubles per block)
  // reproducible reduction (x-march with chunks that are multiples of 4):
  // instead of per-block partials, the partial of every (column, 4-plane
  // x-segment), seg[2*(s*ny*nz + j*nz + k) + {0,1}], each summed in
  // ascending i by the thread that owns the column; k_seg_reduce then sums
  // them in a fixed order, so eps does not depend on the tile or chunk
  double* seg = nullptr;
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
// the same store with the non-temporal hint (the nt bit of the global
// store): the output streams to memory without taking L2 lines from the
// inputs whose planes and halos the march re-reads
template <bool NT, typename R>
__device__ __forceinline__ void sto_h(cplx<R>* __restrict__ base, unsigned boff, cplx<R> v) {
  if constexpr (NT) {
    typedef R V2 __attribute__((ext_vector_type(2)));
    V2* p = reinterpret_cast<V2*>(reinterpret_cast<char*>(base) + boff);
    const V2 w = {v.re, v.im};
    __builtin_nontemporal_store(w, p);
  } else {
    sto(base, boff, v);
  }
}

// Loads through pointers that address global memory.  A pointer read from
// a table in device memory (a batch's operator records) is generic to the
// compiler, which then issues flat loads; those also count against the LDS
// counter, so every LDS wait and barrier would wait for them as well.
template <typename R>
__device__ __forceinline__ cplx<R> ldg(const cplx<R>* base, unsigned boff) {
  typedef R V2 __attribute__((ext_vector_type(2)));
  typedef const __attribute__((address_space(1))) V2 GV2;
  const V2 v = *(GV2*)(reinterpret_cast<const char*>(base) + boff);
  return cplx<R>{v.x, v.y};
}
template <typename R>
__device__ __forceinline__ R ldg_r(const R* p) {
  typedef const __attribute__((address_space(1))) R GR;
  return *(GR*)p;
}

// 1-D table values of one axis at one coordinate
template <typename R>
struct Ax1 {
  R nP, nM, eP, eM, dN, dE, rdN, rdE;
};
template <typename R>
__device__ __forceinline__ Ax1<R> axis_at(const OpArgs<R>& op, int a, int q) {
  return {op.nP[a][q], op.nM[a][q], op.eP[a][q], op.eM[a][q],
          op.dN[a][q], op.dE[a][q], op.rdN[a][q], op.rdE[a][q]};
}

// Block-uniform reads of the axis tables, which no kernel writes, through
// the constant address space: the compiler then uses scalar loads into
// SGPRs, which take no vector-memory slot and no wait on the vector loads.
template <typename R>
__device__ __forceinline__ R ldk_r(const R* p) {
  typedef const __attribute__((address_space(4))) R KR;
  return *(KR*)p;
}
template <typename R>
__device__ __forceinline__ Ax1<R> axis_at_k(const OpArgs<R>& op, int a, int q) {
  return {ldk_r(op.nP[a] + q), ldk_r(op.nM[a] + q), ldk_r(op.eP[a] + q), ldk_r(op.eM[a] + q),
          ldk_r(op.dN[a] + q), ldk_r(op.dE[a] + q), ldk_r(op.rdN[a] + q), ldk_r(op.rdE[a] + q)};
}
template <typename R>
__device__ __forceinline__ Ax1<R> axis_at_g(const OpArgs<R>& op, int a, int q) {
  return {ldg_r(op.nP[a] + q), ldg_r(op.nM[a] + q), ldg_r(op.eP[a] + q), ldg_r(op.eM[a] + q),
          ldg_r(op.dN[a] + q), ldg_r(op.dE[a] + q), ldg_r(op.rdN[a] + q), ldg_r(op.rdE[a] + q)};
}

// The four rows of K u at node (i,j,k).  `a` gives field values
// a.fv(c, d) (c: 0 Ax, 1 Ay, 2 Az, 3 Phi) and k^2 values a.kv(c, d)
// (c: 0 k2x, 1 k2y, 2 k2z, 3 k2p) at direction d: 0 centre, 1 i+1, 2 i-1,
// 3 j+1, 4 j-1, 5 k+1, 6 k-1, with 0 outside the domain.  Shared by both
// kernels so that they perform identical arithmetic.
template <typename R, typename A>
__device__ __forceinline__ void rows4(int i, int j, int k, int nx, int ny, int nz, const Ax1<R>& X,
                                      const Ax1<R>& Y, const Ax1<R>& Z, A& a, cplx<R>& rx,
                                      cplx<R>& ry, cplx<R>& rz, cplx<R>& rp) {
  using C = cplx<R>;
  const C zero = {R(0), R(0)};
  rx = ry = rz = rp = zero;
  const C k2p_c = a.kv(3, 0);
  // ---- x-edge row, inprod_x (src/aphid.f90:841-866)
  if (i < nx - 1) {
    const C k2x_c = a.kv(0, 0);
    const R sm6 = X.eP + X.eM + Y.nP + Y.nM + Z.nP + Z.nM;
    const C c1 = {k2x_c.re - sm6, k2x_c.im};
    C acc = c1 * a.fv(0, 0);
    acc += X.eP * a.fv(0, 1);
    acc += X.eM * a.fv(0, 2);
    acc += Y.nP * a.fv(0, 3);
    acc += Y.nM * a.fv(0, 4);
    acc += Z.nP * a.fv(0, 5);
    acc += Z.nM * a.fv(0, 6);
    const C cpl = a.fv(3, 1) * (a.kv(3, 1) - k2x_c) - a.fv(3, 0) * (k2p_c - k2x_c);
    const R vol = X.dE * Y.dN * Z.dN;
    rx = (acc + cpl * X.rdE) * vol;
  }
  // ---- y-edge row, inprod_y (src/aphid.f90:970-995)
  if (j < ny - 1) {
    const C k2y_c = a.kv(1, 0);
    const R sm6 = Y.eP + Y.eM + X.nP + X.nM + Z.nP + Z.nM;
    const C c1 = {k2y_c.re - sm6, k2y_c.im};
    C acc = c1 * a.fv(1, 0);
    acc += Y.eP * a.fv(1, 3);
    acc += Y.eM * a.fv(1, 4);
    acc += X.nP * a.fv(1, 1);
    acc += X.nM * a.fv(1, 2);
    acc += Z.nP * a.fv(1, 5);
    acc += Z.nM * a.fv(1, 6);
    const C cpl = a.fv(3, 3) * (a.kv(3, 3) - k2y_c) - a.fv(3, 0) * (k2p_c - k2y_c);
    const R vol = X.dN * Y.dE * Z.dN;
    ry = (acc + cpl * Y.rdE) * vol;
  }
  // ---- z-edge row, inprod_z (src/aphid.f90:1098-1123)
  if (k < nz - 1) {
    const C k2z_c = a.kv(2, 0);
    const R sm6 = Z.eP + Z.eM + X.nP + X.nM + Y.nP + Y.nM;
    const C c1 = {k2z_c.re - sm6, k2z_c.im};
    C acc = c1 * a.fv(2, 0);
    acc += Z.eP * a.fv(2, 5);
    acc += Z.eM * a.fv(2, 6);
    acc += X.nP * a.fv(2, 1);
    acc += X.nM * a.fv(2, 2);
    acc += Y.nP * a.fv(2, 3);
    acc += Y.nM * a.fv(2, 4);
    const C cpl = a.fv(3, 5) * (a.kv(3, 5) - k2z_c) - a.fv(3, 0) * (k2p_c - k2z_c);
    const R vol = X.dN * Y.dN * Z.dE;
    rz = (acc + cpl * Z.rdE) * vol;
  }
  // ---- interior node row, inprod_p (src/aphid.f90:1229-1278)
  if (j > 0 && j < ny - 1 && k > 0 && k < nz - 1 && i > 0 && i < nx - 1) {
    const C s1 = k2p_c, s2 = a.kv(0, 0), s3 = a.kv(0, 2);
    const C s4 = a.kv(1, 0), s5 = a.kv(1, 4);
    const C s6 = a.kv(2, 0), s7 = a.kv(2, 6);
    const C c2 = X.nP * s2, c3 = X.nM * s3, c4 = Y.nP * s4;
    const C c5 = Y.nM * s5, c6 = Z.nP * s6, c7 = Z.nM * s7;
    const C c1 = s1 * s1 - (c2 + c3 + c4 + c5 + c6 + c7);
    C acc = c1 * a.fv(3, 0);
    acc += c2 * a.fv(3, 1);
    acc += c3 * a.fv(3, 2);
    acc += c4 * a.fv(3, 3);
    acc += c5 * a.fv(3, 4);
    acc += c6 * a.fv(3, 5);
    acc += c7 * a.fv(3, 6);
    C cp = ((s2 - s1) * X.rdN) * a.fv(0, 0);
    cp += (-((s3 - s1) * X.rdN)) * a.fv(0, 2);
    cp += ((s4 - s1) * Y.rdN) * a.fv(1, 0);
    cp += (-((s5 - s1) * Y.rdN)) * a.fv(1, 4);
    cp += ((s6 - s1) * Z.rdN) * a.fv(2, 0);
    cp += (-((s7 - s1) * Z.rdN)) * a.fv(2, 6);
    const R vol = X.dN * Y.dN * Z.dN;
    rp = acc * vol + cp * vol;
  }
}

template <typename R>
__device__ __forceinline__ void eps_acc(const cplx<R> (&in4)[4], cplx<R> rx, cplx<R> ry,
                                        cplx<R> rz, cplx<R> rp, double& re, double& im) {
  const cplx<double> a0 = ccast<double>(in4[0]) * ccast<double>(rx);
  const cplx<double> a1 = ccast<double>(in4[1]) * ccast<double>(ry);
  const cplx<double> a2 = ccast<double>(in4[2]) * ccast<double>(rz);
  const cplx<double> a3 = ccast<double>(in4[3]) * ccast<double>(rp);
  re += (a0.re + a1.re) + (a2.re + a3.re);
  im += (a0.im + a1.im) + (a2.im + a3.im);
}

// The hardware hands consecutive block ids to the 8 XCDs round-robin; this
// bijection gives XCD x the contiguous range of logical tiles
// [x*q + min(x,r), ...), so that tiles next to each other share one L2.
__device__ __forceinline__ unsigned xcd_remap(unsigned b, unsigned nb) {
  const unsigned q = nb / 8, r = nb % 8, x = b % 8, loc = b / 8;
  return x * q + min(x, r) + loc;
}

// ---------------------------------------------------------------------------
// Variant 0: x-march.  A block owns TJ x TK columns and `chunk` x-planes;
// x-neighbours in a per-thread register queue, in-plane neighbours from
// the current plane staged in LDS.
template <typename R, int TK, int TJ>
struct MarchAcc {
  using C = cplx<R>;
  static constexpr int LK = TK + 2;
  const C (*sm)[(TK + 2) * (TJ + 2)];
  const C* pa;
  const C* ca;
  const C* na;
  int ol;
  C k2x_p, k2x_c, k2p_c, k2p_n;
  __device__ __forceinline__ C fv(int c, int d) const {
    switch (d) {
      case 0: return ca[c];
      case 1: return na[c];
      case 2: return pa[c];
      case 3: return sm[c][ol + LK];
      case 4: return sm[c][ol - LK];
      case 5: return sm[c][ol + 1];
      default: return sm[c][ol - 1];
    }
  }
  __device__ __forceinline__ C kv(int c, int d) const {
    if (c == 0) return d == 2 ? k2x_p : k2x_c;
    if (c == 3) {
      if (d == 0) return k2p_c;
      if (d == 1) return k2p_n;
      return d == 3 ? sm[6][ol + LK] : sm[6][ol + 1];
    }
    if (c == 1) return d == 4 ? sm[4][ol - LK] : sm[4][ol];
    return d == 6 ? sm[5][ol - 1] : sm[5][ol];
  }
};

template <typename R, int TK, int TJ, bool NTS, bool SEG = false, bool CK = false>
__global__ __launch_bounds__(TK* TJ, CK ? 3 : 1) void k_apply(OpArgs<R> op, const cplx<R>* __restrict__ in,
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
  // 1-D grid; tiles ordered k-tile fastest, then j-tile, then x-chunk, and
  // handed to the XCDs in contiguous ranges (xcd_remap) so neighbouring
  // tiles share one L2
  const int ntk = (nz + TK - 1) / TK, ntj = (ny + TJ - 1) / TJ;
  const unsigned t = xcd_remap(blockIdx.x, gridDim.x);
  const int bk = (int)(t % ntk), bj = (int)((t / ntk) % ntj);
  const int bi = (int)(t / ((unsigned)ntk * ntj));
  const int tk = threadIdx.x % TK, tj = threadIdx.x / TK;
  const int k = bk * TK + tk, j = bj * TJ + tj;
  const int i0 = bi * chunk;
  const int i1 = min(nx, i0 + chunk);
  const bool own = (j < ny) && (k < nz);
  const unsigned colb = own ? ((unsigned)j * nz + k) * CS : 0u;
  const unsigned plb = pl * CS;

  if (threadIdx.x < TJ) {
    const int jj = bj * TJ + threadIdx.x;
    const bool v = jj < ny;
    const Ax1<R> t = v ? axis_at(op, 1, jj) : Ax1<R>{};
    tyj[0][threadIdx.x] = t.nP; tyj[1][threadIdx.x] = t.nM;
    tyj[2][threadIdx.x] = t.eP; tyj[3][threadIdx.x] = t.eM;
    tyj[4][threadIdx.x] = t.dN; tyj[5][threadIdx.x] = t.dE;
    tyj[6][threadIdx.x] = t.rdN; tyj[7][threadIdx.x] = t.rdE;
  }
  if (threadIdx.x < TK) {
    const int kk = bk * TK + threadIdx.x;
    const bool v = kk < nz;
    const Ax1<R> t = v ? axis_at(op, 2, kk) : Ax1<R>{};
    tzk[0][threadIdx.x] = t.nP; tzk[1][threadIdx.x] = t.nM;
    tzk[2][threadIdx.x] = t.eP; tzk[3][threadIdx.x] = t.eM;
    tzk[4][threadIdx.x] = t.dN; tzk[5][threadIdx.x] = t.dE;
    tzk[6][threadIdx.x] = t.rdN; tzk[7][threadIdx.x] = t.rdE;
  }

  // own value of component c (0..3) or k2 grid c (0..3) at plane i
  auto ldc = [&](const C* __restrict__ base, int c, int i) -> C {
    return (own && i >= 0 && i < nx) ? ldo(base, c * nnb + (unsigned)i * plb + colb)
                                     : C{R(0), R(0)};
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
  const int gj = bj * TJ + hj, gk = bk * TK + hk;
  const bool hin = hv && gj >= 0 && gj < ny && gk >= 0 && gk < nz;
  // the one k^2 grid a halo cell serves: the j-1 row feeds k2y(j-1), the
  // k-1 column k2z(k-1), the j+1 row and the k+1 column k2p (rows4 reads
  // no other k^2 value off the tile)
  const int hc = hj == -1 ? 1 : (hk == -1 ? 2 : 3);
  const unsigned hcolb = hin ? ((unsigned)gj * nz + gk) * CS : 0u;
  const int hl = (hj + 1) * LK + (hk + 1);
  const int ol = (tj + 1) * LK + (tk + 1);

  // Compact k^2 (CK): the pairs of node (i, column) are at byte
  // i*plb + colb of k2rec (half = 0) and k2yz_off bytes further
  // (half = k2yz_off), its k2p code at byte (i*plb + colb)/16 of k2pc.  The
  // first pair {k2x.im, k2p.im} and the code are read one plane ahead (the
  // x-queues), the second pair {k2y.im, k2z.im} at the plane itself (0
  // outside the domain).  Real parts: k2x from its table (one value per
  // column, rex), k2y and k2z from theirs, k2p from its code (kpre).  In
  // the edge grids' padding (i = nx-1 for k2x, j = ny-1 for k2y, k = nz-1
  // for k2z) a table gives a nonzero real part where the whole grid holds
  // 0; rows4 never reads those entries.  The ry and rz tables are
  // contiguous (k2re[2] = k2re[1] + nx*nz), so both are read through one
  // base with 32-bit byte offsets.
  constexpr unsigned RS = sizeof(R);
  const unsigned rzb = (unsigned)nx * nz * RS;
  const R rex = (CK && own) ? ldo(op.k2re[0], ((unsigned)j * nz + k) * RS) : R(0);
  const unsigned ryo = (unsigned)min(k, nz - 1) * RS, rzo = (unsigned)min(j, ny - 1) * RS;
  auto ldr = [&](int i, unsigned half) -> C {
    return (own && i >= 0 && i < nx)
               ? ldo(reinterpret_cast<const C*>(op.k2rec), half + (unsigned)i * plb + colb)
               : C{R(0), R(0)};
  };
  // k2p's real part from its code
  auto kpre = [&](unsigned code) -> R {
    return __longlong_as_double((long long)(op.k2p_base + code));
  };
  // compact k^2, halo cell: byte offset of its real part's table entry at
  // plane 0 (hto; hc = 3 reads the entry of k = 0, unused; a cell off the
  // grid reads a clamped, in-bounds entry)
  const unsigned hto = hc == 1 ? (unsigned)min(max(gk, 0), nz - 1) * RS
                               : (hc == 2 ? rzb + (unsigned)min(max(gj, 0), ny - 1) * RS : 0u);
  // the k^2 value a halo cell stages (grid hc of the cell at (gj, gk));
  // hcode: the cell's k2p code at plane i (compact k^2, hc = 3)
  auto ldh = [&](unsigned g, int i, unsigned hcode) -> C {
    if constexpr (!CK) {
      return ldo(op.k2, hc * nnb + g);
    } else {
      // the pairs' k2p.im, k2y.im or k2z.im and the table entry (k2y,
      // k2z), both loaded unconditionally (in bounds for every hc)
      const R im = ldo(op.k2rec, g + (hc == 3 ? 8u : (hc == 1 ? op.k2yz_off : op.k2yz_off + 8u)));
      const R tre = ldo(op.k2re[1], (unsigned)i * (hc == 2 ? (unsigned)ny : (unsigned)nz) * RS + hto);
      return C{hc == 3 ? kpre(hcode) : tre, im};
    }
  };
  // a halo cell's k2p code at plane i (compact k^2), loaded at the end of
  // the step before: loaded next to its use, the compiler issues it only
  // for hc = 3, after a wait for the other halo loads
  auto ldhc = [&](int i) -> unsigned {
    if constexpr (CK) return op.k2pc[((unsigned)min(i, nx - 1) * plb + hcolb) / CS];
    else return 0u;
  };

  C pa[4], ca[4], na[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    pa[c] = ldc(in, c, i0 - 1);
    ca[c] = ldc(in, c, i0);
  }
  const C* __restrict__ K2 = op.k2;
  // the x-neighbour queue of k2x: whole values, or (CK) imaginary parts
  // only, all sharing the column's table value rex
  using KX = std::conditional_t<CK, R, C>;
  auto kxv = [&](KX v) -> C {
    if constexpr (CK) return C{rex, v};
    else return v;
  };
  // k2x and k2p of the own column at plane i (the queues' loads); with
  // the compact k^2, k2p's real part stays a code (kc) until the value is
  // used (kfix): a conversion next to the load would make the wave wait
  // for every load issued before it, the plane i+1 fields included
  auto ldxp = [&](int i, KX& kx, C& kp, unsigned& kc) {
    if constexpr (CK) {
      const C r = ldr(i, 0u);
      kx = r.re;
      kp = C{R(0), r.im};
      // unconditional, from an in-bounds node (plane nx-1 past the last
      // plane, column 0 for threads off the grid): a load under a branch
      // would have its merge value formed, and waited for, at once
      kc = op.k2pc[((unsigned)min(i, nx - 1) * plb + colb) / CS];
    } else {
      kx = ldc(K2, 0, i);
      kp = ldc(K2, 3, i);
    }
  };
  auto kfix = [&](C& kp, unsigned kc, int i) {
    if constexpr (CK) kp.re = (own && i >= 0 && i < nx) ? kpre(kc) : R(0);
  };
  KX k2x_p, k2x_c;
  C k2p_c;
  unsigned kc_c = 0;
  if constexpr (CK) k2x_p = ldr(i0 - 1, 0u).re;
  else k2x_p = ldc(K2, 0, i0 - 1);
  ldxp(i0, k2x_c, k2p_c, kc_c);
  kfix(k2p_c, kc_c, i0);
  unsigned hc_c = ldhc(i0);

  double eacc_re = 0.0, eacc_im = 0.0;

  for (int i = i0; i < i1; ++i) {
    // stage plane i: the loads this step needs (own k2y, k2z and the
    // halo) go out first, since waits on loads complete in issue order
    if constexpr (CK) {
      // the pair and both table entries are loaded unconditionally, from
      // in-bounds addresses (column 0 and clamped table indices off the
      // grid), and masked afterwards, so that all three go out before one
      // wait
      const C r = ldo(reinterpret_cast<const C*>(op.k2rec), op.k2yz_off + (unsigned)i * plb + colb);
      const R ty = ldo(op.k2re[1], (unsigned)i * nz * RS + ryo);
      const R tz = ldo(op.k2re[1], rzb + (unsigned)i * ny * RS + rzo);
      sm[4][ol] = own ? C{ty, r.re} : C{R(0), R(0)};
      sm[5][ol] = own ? C{tz, r.im} : C{R(0), R(0)};
    } else {
      sm[4][ol] = ldc(op.k2, 1, i);
      sm[5][ol] = ldc(op.k2, 2, i);
    }
    if (hv) {
      // the halo cell's five loads go out together, unconditionally (a cell
      // off the grid reads node (i, 0, 0), hcolb = 0) and are masked
      // afterwards: a load under its own branch gets its own wait
      const unsigned g = (unsigned)i * plb + hcolb;
      C hf[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) hf[c] = ldo(in, c * nnb + g);
      const C hk2 = ldh(g, i, hc_c);
#pragma unroll
      for (int c = 0; c < 4; ++c) sm[c][hl] = hin ? hf[c] : C{R(0), R(0)};
      sm[3 + hc][hl] = hin ? hk2 : C{R(0), R(0)};
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) sm[c][ol] = ca[c];
    sm[6][ol] = k2p_c;
    // plane i+1 of the own column (the x-neighbours of the next step)
#pragma unroll
    for (int c = 0; c < 4; ++c) na[c] = ldc(in, c, i + 1);
    KX k2x_n;
    C k2p_n;
    unsigned kc_n = 0;
    ldxp(i + 1, k2x_n, k2p_n, kc_n);
    __syncthreads();
    kfix(k2p_n, kc_n, i + 1);

    if (own) {
      const unsigned ob = (unsigned)i * plb + colb;
      const Ax1<R> X = axis_at_k(op, 0, i);  // i is uniform over the block
      const Ax1<R> Y = {tyj[0][tj], tyj[1][tj], tyj[2][tj], tyj[3][tj],
                        tyj[4][tj], tyj[5][tj], tyj[6][tj], tyj[7][tj]};
      const Ax1<R> Z = {tzk[0][tk], tzk[1][tk], tzk[2][tk], tzk[3][tk],
                        tzk[4][tk], tzk[5][tk], tzk[6][tk], tzk[7][tk]};
      MarchAcc<R, TK, TJ> a{sm, pa, ca, na, ol, kxv(k2x_p), kxv(k2x_c), k2p_c, k2p_n};
      C rx, ry, rz, rp;
      rows4<R>(i, j, k, nx, ny, nz, X, Y, Z, a, rx, ry, rz, rp);
      sto_h<NTS>(out, ob, rx);
      sto_h<NTS>(out, nnb + ob, ry);
      sto_h<NTS>(out, 2 * nnb + ob, rz);
      sto_h<NTS>(out, 3 * nnb + ob, rp);
      if (ep.st) eps_acc<R>(ca, rx, ry, rz, rp, eacc_re, eacc_im);
      if constexpr (SEG) {
        if ((i & 3) == 3 || i == nx - 1) {
          const size_t e = 2 * ((size_t)(i >> 2) * pl + (size_t)j * nz + k);
          ep.seg[e] = eacc_re;
          ep.seg[e + 1] = eacc_im;
          eacc_re = eacc_im = 0.0;
        }
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
    hc_c = ldhc(i + 1);
  }

  if constexpr (!SEG) {
    if (ep.st) {
      double v[2] = {eacc_re, eacc_im};
      block_sum<2>(v);
      block_partial<2>(v, ep.part);
    }
  }
}

// The fixed-order sum of a reproducible apply's segment partials: block g
// of a fixed grid sums entries [g*per, (g+1)*per) of the n (re, im) pairs,
// each thread in ascending order at a stride of the block size, then the
// block's fixed tree; the one-block finalize adds the grid's partials.
// Nothing here depends on the stencil's launch shape.
__global__ void k_seg_reduce(const double* __restrict__ seg, long long n, long long per,
                             const QmrState* st, double* __restrict__ part) {
  if (st->done) return;
  double v[2] = {0.0, 0.0};
  const long long a = (long long)blockIdx.x * per, b = min(n, a + per);
  for (long long q = a + threadIdx.x; q < b; q += blockDim.x) {
    v[0] += seg[2 * q];
    v[1] += seg[2 * q + 1];
  }
  block_sum<2>(v);
  block_partial<2>(v, part);
}

// ---------------------------------------------------------------------------
// The x-march's memory traffic alone, for measurement: the same tiles, grid,
// XCD order and chunks, the same loads (own columns of the four fields and
// four k^2 grids, the halo cells' fields and one k^2 grid) and the same four
// stores per owned node (with the same non-temporal hint), but no LDS
// staging, barriers or row arithmetic.  It holds as much LDS as
// k_apply<R,TK,TJ,NTS>, so the same number of blocks share a CU.  Its time is the floor the march's access pattern sets; the loaded
// values are folded into the (meaningless) outputs so that no load is
// dropped.
template <typename R, int TK, int TJ, bool NTS, bool CK = false>
__global__ __launch_bounds__(TK* TJ) void k_apply_traffic(OpArgs<R> op,
                                                          const cplx<R>* __restrict__ in,
                                                          cplx<R>* __restrict__ out, int chunk) {
  using C = cplx<R>;
  constexpr int LK = TK + 2, LJ = TJ + 2, LS = LK * LJ;
  constexpr int NH = 2 * TK + 2 * TJ;
  constexpr unsigned CS = sizeof(C);
  // k_apply's LDS: 7 planes and the 1-D tables
  __shared__ C hold[7 * LS + (8 * (TJ + TK) * sizeof(R) + CS - 1) / CS];
  const int nx = op.nx, ny = op.ny, nz = op.nz;
  const unsigned pl = (unsigned)ny * nz;
  const unsigned nnb = (unsigned)op.nn * CS, plb = pl * CS;
  const int ntk = (nz + TK - 1) / TK, ntj = (ny + TJ - 1) / TJ;
  const unsigned t = xcd_remap(blockIdx.x, gridDim.x);
  const int bk = (int)(t % ntk), bj = (int)((t / ntk) % ntj);
  const int bi = (int)(t / ((unsigned)ntk * ntj));
  const int tk = threadIdx.x % TK, tj = threadIdx.x / TK;
  const int k = bk * TK + tk, j = bj * TJ + tj;
  const int i0 = bi * chunk, i1 = min(nx, i0 + chunk);
  const bool own = (j < ny) && (k < nz);
  const unsigned colb = own ? ((unsigned)j * nz + k) * CS : 0u;
  int hj = 0, hk = 0;
  const bool hv = threadIdx.x < NH;
  if (hv) {
    const int h = threadIdx.x;
    if (h < TK) { hj = -1; hk = h; }
    else if (h < 2 * TK) { hj = TJ; hk = h - TK; }
    else if (h < 2 * TK + TJ) { hj = h - 2 * TK; hk = -1; }
    else { hj = h - 2 * TK - TJ; hk = TK; }
  }
  const int gj = bj * TJ + hj, gk = bk * TK + hk;
  const bool hin = hv && gj >= 0 && gj < ny && gk >= 0 && gk < nz;
  const int hc = hj == -1 ? 1 : (hk == -1 ? 2 : 3);
  const unsigned hcolb = hin ? ((unsigned)gj * nz + gk) * CS : 0u;
  const C* __restrict__ K2 = op.k2;
  auto ldc = [&](const C* __restrict__ base, int c, int i) -> C {
    return (own && i >= 0 && i < nx) ? ldo(base, c * nnb + (unsigned)i * plb + colb)
                                     : C{R(0), R(0)};
  };
  // the k^2 loads of k_apply<..., CK>: whole grids, or the compact
  // record halves, k2p codes and tables
  auto ldr = [&](int i, unsigned half) -> C {
    return (own && i >= 0 && i < nx)
               ? ldo(reinterpret_cast<const C*>(op.k2rec), half + (unsigned)i * plb + colb)
               : C{R(0), R(0)};
  };
  // {k2x, k2p} of plane i
  auto ldxp = [&](int i) -> C {
    if constexpr (!CK) {
      return ldc(K2, 0, i) + ldc(K2, 3, i);
    } else {
      const C r = ldr(i, 0u);
      if (!(own && i >= 0 && i < nx)) return r;
      return C{r.re + R(op.k2pc[((unsigned)i * plb + colb) / CS]), r.im};
    }
  };
  // {k2y, k2z} of plane i
  auto ldyz = [&](int i) -> C {
    if constexpr (!CK) {
      return ldc(K2, 1, i) + ldc(K2, 2, i);
    } else {
      const C r = ldr(i, op.k2yz_off);
      if (!own) return r;
      return C{r.re + op.k2re[1][i * nz + k], r.im + op.k2re[2][i * ny + j]};
    }
  };
  auto ldh = [&](unsigned g, int i) -> C {
    if constexpr (!CK) {
      return ldo(K2, hc * nnb + g);
    } else {
      if (hc == 3) return C{ldo(op.k2rec, g + 8u), R(op.k2pc[g / CS])};
      const R im = ldo(op.k2rec, g + (hc == 1 ? op.k2yz_off : op.k2yz_off + 8u));
      const R re = hc == 1 ? op.k2re[1][i * nz + gk] : op.k2re[2][i * ny + gj];
      return C{re, im};
    }
  };
  hold[threadIdx.x] = C{R(0), R(0)};
  C acc = ldxp(i0 - 1);
  if constexpr (CK) acc.re += own ? op.k2re[0][j * nz + k] : R(0);
  C ca[4], na[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    acc += ldc(in, c, i0 - 1);
    ca[c] = ldc(in, c, i0);
  }
  C kx = ldxp(i0), kp = {R(0), R(0)};
  for (int i = i0; i < i1; ++i) {
    const C ky = ldyz(i), kz = {R(0), R(0)};
    C h = {R(0), R(0)};
    if (hin) {
      const unsigned g = (unsigned)i * plb + hcolb;
#pragma unroll
      for (int c = 0; c < 4; ++c) h += ldo(in, c * nnb + g);
      h += ldh(g, i);
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) na[c] = ldc(in, c, i + 1);
    const C kxn = ldxp(i + 1), kpn = {R(0), R(0)};
    if (own) {
      const unsigned ob = (unsigned)i * plb + colb;
      const C s = acc + h + kx + kp + ky + kz;
#pragma unroll
      for (int c = 0; c < 4; ++c) sto_h<NTS>(out, c * nnb + ob, ca[c] + s);
    }
    acc = C{acc.re * R(0.5), h.im};
#pragma unroll
    for (int c = 0; c < 4; ++c) ca[c] = na[c];
    kx = kxn;
    kp = kpn;
  }
  if (acc.re == R(1.5e300)) out[0] = hold[threadIdx.x];
}

// ---------------------------------------------------------------------------
// Variant 1: direct.  One node per thread, every neighbour read through the
// L1/L2 caches; blocks are TK x TJ tiles of one x-plane.  The linear block id
// is remapped so that the blocks one XCD receives (ids congruent mod 8) walk
// one contiguous slab of planes in order: neighbouring planes are then
// re-read from that XCD's own L2.
template <typename R>
struct DirectAcc {
  using C = cplx<R>;
  const C* __restrict__ in;
  const C* __restrict__ k2;
  unsigned nnb, plb, rowb;  // component, plane and row (j) strides in bytes
  unsigned ob;              // byte offset of the node
  int i, j, k, nx, ny, nz;
  __device__ __forceinline__ bool inside(int d) const {
    switch (d) {
      case 1: return i + 1 < nx;
      case 2: return i > 0;
      case 3: return j + 1 < ny;
      case 4: return j > 0;
      case 5: return k + 1 < nz;
      case 6: return k > 0;
      default: return true;
    }
  }
  __device__ __forceinline__ unsigned off(int d) const {
    switch (d) {
      case 1: return ob + plb;
      case 2: return ob - plb;
      case 3: return ob + rowb;
      case 4: return ob - rowb;
      case 5: return ob + (unsigned)sizeof(C);
      case 6: return ob - (unsigned)sizeof(C);
      default: return ob;
    }
  }
  __device__ __forceinline__ C fv(int c, int d) const {
    return inside(d) ? ldo(in, c * nnb + off(d)) : C{R(0), R(0)};
  }
  __device__ __forceinline__ C kv(int c, int d) const {
    return inside(d) ? ldo(k2, c * nnb + off(d)) : C{R(0), R(0)};
  }
};

template <typename R, int TK, int TJ>
__global__ __launch_bounds__(TK* TJ) void k_apply_direct(OpArgs<R> op, const cplx<R>* __restrict__ in,
                                                         cplx<R>* __restrict__ out, ApplyEpilogue ep) {
  using C = cplx<R>;
  if (ep.st && ep.st->done) return;
  const int nx = op.nx, ny = op.ny, nz = op.nz;
  const int ntk = (nz + TK - 1) / TK, ntj = (ny + TJ - 1) / TJ;
  const unsigned t = xcd_remap(blockIdx.x, gridDim.x);
  const int tkt = (int)(t % ntk);
  const int tjt = (int)((t / ntk) % ntj);
  const int i = (int)(t / ((unsigned)ntk * ntj));
  const int k = tkt * TK + (int)(threadIdx.x % TK), j = tjt * TJ + (int)(threadIdx.x / TK);
  double eacc_re = 0.0, eacc_im = 0.0;
  if (j < ny && k < nz && i < nx) {
    constexpr unsigned CS = sizeof(C);
    DirectAcc<R> a{in, op.k2, (unsigned)op.nn * CS, (unsigned)ny * nz * CS, (unsigned)nz * CS,
                   (((unsigned)i * ny + j) * nz + k) * CS, i, j, k, nx, ny, nz};
    C rx, ry, rz, rp;
    rows4<R>(i, j, k, nx, ny, nz, axis_at(op, 0, i), axis_at(op, 1, j), axis_at(op, 2, k), a, rx,
             ry, rz, rp);
    sto(out, a.ob, rx);
    sto(out, a.nnb + a.ob, ry);
    sto(out, 2 * a.nnb + a.ob, rz);
    sto(out, 3 * a.nnb + a.ob, rp);
    if (ep.st) {
      const C in4[4] = {a.fv(0, 0), a.fv(1, 0), a.fv(2, 0), a.fv(3, 0)};
      eps_acc<R>(in4, rx, ry, rz, rp, eacc_re, eacc_im);
    }
  }
  if (ep.st) {
    double v[2] = {eacc_re, eacc_im};
    block_sum<2>(v);
    block_partial<2>(v, ep.part);
  }
}

// eps = p^T K p (the partials of k_apply), beta = eps/del (src/aphid_qmr.f90:329-330)
__global__ void k_fin_eps(const double* part, int nb, QmrState* st) {
  // the state's fields are loaded with the partials (one round trip)
  const int done = st->done;
  const c128 del = st->del;
  double tot[2];
  finalize_sum<2>(part, nb, tot);
  if (threadIdx.x == 0 && !done) {
    const c128 eps = {tot[0], tot[1]};
    const c128 beta = cdiv(eps, del);
    st->eps = eps;
    st->beta = beta;
    if (is_zero(eps) || !finite_c(beta)) {
      st->status = ST_BREAK;
      st->done = 1;
    }
  }
}
