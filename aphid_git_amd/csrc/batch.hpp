// batch.hpp -- many independent A-Phi systems in one launch per phase.
//
// The reference solves one (frequency, source) system per process and farms
// them out with MPI (src/aphid_parallel.f90:75-201).  Here all systems of a
// batch -- different frequencies, sources and mesh sizes -- share every
// launch of the Jacobi-QMR iteration:
//   - all vectors of all systems are concatenated in one flat array per
//     vector kind; system s owns [off_s, off_s + n4_s) (n4_s = 4 nn_s,
//     device layout of stencil.hpp);
//   - streaming kernels run a work list of fixed-size element ranges, each
//     range inside one system (block b -> system via binary search of the
//     per-system block prefix); per-block partials are summed by one
//     finalize block per system;
//   - the stencil runs a work list of (system, tile) items over all systems'
//     planes (one node per thread, neighbours through L1/L2);
//   - systems with the same mesh and frequency share one operator (k^2 and
//     1/diag); every system carries its own QmrState, and blocks of a
//     converged system return at once (the convergence mask).
#pragma once
#include "qmr.hpp"
#include "stencil.hpp"

constexpr int kBElems = 512;  // elements per streaming work item (256 threads x 2)

struct BatchDev {
  int nsys;
  const long long* off;     // [nsys] element offset of each system's block
  const long long* n4;      // [nsys] 4*nn of each system
  const int* op;            // [nsys] operator index
  const long long* opoff;   // [nops] element offset of each operator's k2 / 1/diag block
  const int* bstart;        // [nsys+1] streaming work-item prefix
  const int* tstart;        // [nsys+1] stencil tile prefix
  QmrState* st;             // [nsys]
  const int* bsys;          // [bstart[nsys]] system of each streaming work item
  const int* tsys;          // [tstart[nsys]] system of each stencil tile
  // ||r|| per iteration, hist_stride entries per system (null: none); a
  // system's state says how many of them its solve writes (hist_cap)
  double* hist;
  int hist_stride;
};

// streaming work item of this block: system, global element range [e0, e1)
struct BItem {
  int s;
  long long e0, e1;
};
__device__ __forceinline__ BItem b_item(const BatchDev& B) {
  const int b = blockIdx.x;
  const int s = B.bsys[b];
  const long long lb = b - B.bstart[s];
  const long long e0 = lb * kBElems, e1 = min(e0 + kBElems, B.n4[s]);
  return {s, B.off[s] + e0, B.off[s] + e1};
}

#define B_LOOP(q, it) for (long long q = (it).e0 + threadIdx.x; q < (it).e1; q += blockDim.x)

// Two elements per thread and step (q, q + blockDim.x) with all their loads
// issued before either is used, then the remaining one: twice the loads in
// flight of B_LOOP, and the same per-thread order of elements (the order of
// any partial sum is unchanged).  LOAD(e) loads element e into a struct,
// BODY(e, loaded) uses it.
template <typename LOAD, typename BODY>
__device__ __forceinline__ void b_loop2(const BItem& it, LOAD load, BODY body) {
  long long q = it.e0 + threadIdx.x;
  const long long step = blockDim.x;
  for (; q + step < it.e1; q += 2 * step) {
    const auto a = load(q);
    const auto b = load(q + step);
    __builtin_amdgcn_sched_barrier(0);  // keep both elements' loads ahead of the uses
    body(q, a);
    body(q + step, b);
  }
  if (q < it.e1) body(q, load(q));
}

template <typename V>
__global__ void kb_norm_b(BatchDev B, const vec<V>* __restrict__ b, double* part) {
  const BItem it = b_item(B);
  double a[1] = {0.0};
  B_LOOP(q, it) a[0] += abs2(up(b[q]));
  block_sum<1>(a);
  block_partial<1>(a, part);
}

// finalize helper: fixed-order sum of system s's work-item partials
template <int NV>
__device__ __forceinline__ void b_fin_sum(const double* __restrict__ part, int p0, int p1,
                                          double (&tot)[NV]) {
  finalize_sum<NV>(part + (size_t)p0 * NV, p1 - p0, tot);
}

__global__ void kb_fin_norm_b(BatchDev B, const double* part, double tol) {
  const int s = blockIdx.x;
  double tot[1];
  b_fin_sum<1>(part, B.bstart[s], B.bstart[s + 1], tot);
  if (threadIdx.x == 0) {
    B.st[s].rhsnorm = sqrt(tot[0]);
    B.st[s].target = tol * B.st[s].rhsnorm;
  }
}

// r = b - t (t = K x0 when warm-started), ||r||^2 partials
template <typename V>
__global__ void kb_init_r(BatchDev B, const vec<V>* __restrict__ b, const vec<V>* __restrict__ t,
                          vec<V>* __restrict__ r, double* part) {
  const BItem it = b_item(B);
  double a[1] = {0.0};
  B_LOOP(q, it) {
    c128 e = up(b[q]);
    if (t) e = e - up(t[q]);
    const vec<V> ev = dn<V>(e);
    r[q] = ev;
    a[0] += abs2(up(ev));
  }
  block_sum<1>(a);
  block_partial<1>(a, part);
}
__global__ void kb_fin_init_r(BatchDev B, const double* part) {
  const int s = blockIdx.x;
  double tot[1];
  b_fin_sum<1>(part, B.bstart[s], B.bstart[s + 1], tot);
  if (threadIdx.x == 0) B.st[s].rho_np = sqrt(tot[0]);
}

template <typename V>
__global__ void kb_init_v(BatchDev B, const vec<V>* __restrict__ r, const vec<V>* __restrict__ invd,
                          vec<V>* __restrict__ v, vec<V>* __restrict__ p, vec<V>* __restrict__ d,
                          vec<V>* __restrict__ s_, double* part) {
  const BItem it = b_item(B);
  const double rho = B.st[it.s].rho_np;
  const long long io = B.opoff[B.op[it.s]] - B.off[it.s];  // system -> operator offset
  const vec<V> z = {V(0), V(0)};
  double a[2] = {0.0, 0.0};
  B_LOOP(q, it) {
    const vec<V> vv = dn<V>(divr(up(r[q]), rho));
    v[q] = vv;
    p[q] = z;
    d[q] = z;
    s_[q] = z;
    const c128 vd = up(vv);
    const c128 pr = vd * (vd * up(invd[q + io]));
    a[0] += pr.re;
    a[1] += pr.im;
  }
  block_sum<2>(a);
  block_partial<2>(a, part);
}
// skip (optional): systems solved elsewhere; they start done and untouched
__global__ void kb_fin_init_v(BatchDev B, const double* part, int max_it, int hist_cap,
                              const int* __restrict__ maxits, const int* __restrict__ skip) {
  const int s = blockIdx.x;
  double tot[2];
  b_fin_sum<2>(part, B.bstart[s], B.bstart[s + 1], tot);
  if (threadIdx.x == 0) {
    QmrState* st = &B.st[s];
    if (skip && skip[s]) {
      st->done = 1;
      st->status = ST_CONV;
      st->iter = 0;
      return;
    }
    const double rho = st->rho_np;
    st->max_it = maxits ? maxits[s] : max_it;
    st->hist_cap = hist_cap;
    st->c_n = 1.0;
    st->eps = {1.0, 0.0};
    st->theta_n = 0.0;
    st->eta = {-1.0, 0.0};
    st->iter = 1;
    st->done = 0;
    st->status = ST_MAXIT;
    st->rnorm = rho;
    st->c_nm = st->c_n;
    st->rho_n = rho;
    st->theta_nm = st->theta_n;
    st->del = {tot[0], tot[1]};
    st->k1 = cdiv(rho * st->del, st->eps);
    if (!(rho > 0.0) || !isfinite(rho)) {
      st->done = 1;
      st->status = (rho == 0.0) ? ST_CONV : ST_BREAK;
      st->iter = 0;
    } else if (is_zero(st->del) || !finite_c(st->k1)) {
      st->done = 1;
      st->status = ST_BREAK;
    }
  }
}

// The streaming kernels take the non-temporal hint NTS of qmr.hpp's
// ldv/stv as a template flag; the batch picks it per kernel by timing.
// 1/diag loads stay plain, as in the single-system kernels.
template <typename V, bool NTS = false>
__global__ void kb_pupdate(BatchDev B, const vec<V>* __restrict__ v, const vec<V>* __restrict__ invd,
                           vec<V>* __restrict__ p, vec<V>* __restrict__ d, vec<V>* __restrict__ x) {
  const BItem it = b_item(B);
  QmrState* st = &B.st[it.s];
  if (st->done) return;
  const c128 k1 = st->k1, eta = st->eta;
  const double k2 = st->k2s;
  const long long io = B.opoff[B.op[it.s]] - B.off[it.s];
  // the d and x updates of iteration m = iter - 1 of this system, as in
  // k_pupdate (x stored for even m only; none for m = 0)
  const int m = st->iter - 1;
  struct L {
    vec<V> v, i, p, d, x;
  };
  auto load = [&](long long q, int mode) {
    return L{ldv<NTS>(v + q), ldv<false>(invd + q + io), ldv<NTS>(p + q),
             mode != PU_P ? ldv<NTS>(d + q) : vec<V>{}, mode == PU_PDX ? ldv<NTS>(x + q) : vec<V>{}};
  };
  if (m <= 0)
    b_loop2(it, [&](long long q) { return load(q, PU_P); }, [&](long long q, const L& l) {
      pupdate_elem<V, NTS, PU_P>(p, d, x, q, l.v, l.i, l.p, l.d, l.x, k1, eta, k2);
    });
  else if (m & 1)
    b_loop2(it, [&](long long q) { return load(q, PU_PD); }, [&](long long q, const L& l) {
      pupdate_elem<V, NTS, PU_PD>(p, d, x, q, l.v, l.i, l.p, l.d, l.x, k1, eta, k2);
    });
  else
    b_loop2(it, [&](long long q) { return load(q, PU_PDX); }, [&](long long q, const L& l) {
      pupdate_elem<V, NTS, PU_PDX>(p, d, x, q, l.v, l.i, l.p, l.d, l.x, k1, eta, k2);
    });
  if ((int)blockIdx.x == B.bstart[it.s] && threadIdx.x == 0) st->dx_done = m > 0 ? m : 0;
}

// batched stencil, x-march over flat in-plane chunks.  A block owns NT
// consecutive nodes of one system's (j, k) plane (flat index q = j*nz + k,
// which fits any (ny, nz) without idle lanes) and xch consecutive x-planes:
//   - the in-plane neighbours (q +- 1, q +- nz) come from an LDS window of
//     the plane holding the chunk plus nz nodes on each side (stride
//     WS >= NT + 2 nz for every system of the batch); a halo node
//     holds the four fields and the k^2 grids its side needs (k2y and k2z
//     on the low side, k2p on the high side);
//   - the x-neighbours (i +- 1) of the own node live in a register queue;
//   - a neighbour outside the domain is read where rows4 multiplies it by
//     an exact zero (the integer-division wall masks of the coefficient
//     tables) or not at all, so a flat index that wraps to the next z-row
//     needs no test; the window's nodes outside the plane hold zeros.
// eps partial per block.
template <typename R, int NT, int WS>
struct FlatAcc {
  using C = cplx<R>;
  static constexpr int ws = WS;
  const C* sm;  // 7 arrays of stride ws: Ax Ay Az Phi k2y k2z k2p
  int w, nz;
  const C* pa;
  const C* ca;
  const C* na;
  C k2x_p, k2x_c, k2p_c, k2p_n;
  __device__ __forceinline__ C at(int a, int o) const { return sm[a * ws + w + o]; }
  __device__ __forceinline__ C fv(int c, int d) const {
    switch (d) {
      case 0: return ca[c];
      case 1: return na[c];
      case 2: return pa[c];
      case 3: return at(c, nz);
      case 4: return at(c, -nz);
      case 5: return at(c, 1);
      default: return at(c, -1);
    }
  }
  __device__ __forceinline__ C kv(int c, int d) const {
    if (c == 0) return d == 2 ? k2x_p : k2x_c;
    if (c == 3) {
      if (d == 0) return k2p_c;
      if (d == 1) return k2p_n;
      return d == 3 ? at(6, nz) : at(6, 1);
    }
    if (c == 1) return d == 4 ? at(4, -nz) : at(4, 0);
    return d == 6 ? at(5, -1) : at(5, 0);
  }
};

template <typename R, int NT, int WS>
__global__ __launch_bounds__(NT) void kb_apply(BatchDev B, const OpArgs<R>* __restrict__ ops,
                                               const cplx<R>* __restrict__ in,
                                               cplx<R>* __restrict__ out, int with_eps,
                                               double* part, int xch) {
  using C = cplx<R>;
  extern __shared__ __align__(16) unsigned char kb_smem[];
  C* sm = reinterpret_cast<C*>(kb_smem);
  constexpr unsigned CS = sizeof(C);
  const int b = (int)xcd_remap(blockIdx.x, gridDim.x);
  const int s = B.tsys[b];
  double eacc_re = 0.0, eacc_im = 0.0;
  const bool live = !(with_eps && B.st[s].done);
  if (live) {
    const OpArgs<R>& op = ops[B.op[s]];
    const int nx = op.nx, ny = op.ny, nz = op.nz;
    const int pl = ny * nz, nch = (pl + NT - 1) / NT;
    const int t = b - B.tstart[s];
    const int ch = t % nch, xc = t / nch;
    const int q0 = ch * NT, q = q0 + (int)threadIdx.x;
    const bool own = q < pl;
    const int j = own ? q / nz : 0, k = own ? q - j * nz : 0;
    const int i0 = xc * xch, i1 = min(nx, i0 + xch);
    const unsigned nnb = (unsigned)op.nn * CS, plb = (unsigned)pl * CS;
    const unsigned colb = own ? (unsigned)q * CS : 0u;
    const C* __restrict__ sin = in + B.off[s];
    const C* __restrict__ K2 = op.k2;
    C* __restrict__ sout = out + B.off[s];
    constexpr int ws = WS;  // compile-time, so window offsets fold into the LDS immediates
    const int w = nz + (int)threadIdx.x;
    auto ldc = [&](const C* __restrict__ base, int c, int i) -> C {
      return (own && i >= 0 && i < nx) ? ldg(base, c * nnb + (unsigned)i * plb + colb) : C{R(0), R(0)};
    };
    C pa[4], ca[4], na[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      pa[c] = ldc(sin, c, i0 - 1);
      ca[c] = ldc(sin, c, i0);
    }
    C k2x_p = ldc(K2, 0, i0 - 1), k2x_c = ldc(K2, 0, i0), k2p_c = ldc(K2, 3, i0);
    for (int i = i0; i < i1; ++i) {
      // stage plane i: own values, then the window's halo (nz nodes below
      // the chunk: fields, k2y, k2z; nz above: fields, k2p).  The loads go
      // out unconditionally, from in-bounds addresses (node (i, 0, 0) off
      // the system's plane), and are masked afterwards: a load under its
      // own branch gets its own wait before the next is issued
      // (own k2y and k2z are stored after the halo, so that their loads and
      // the first halo cell's share one wait)
      const unsigned go = (unsigned)i * plb + colb;
      const C ky = ldg(K2, nnb + go), kz = ldg(K2, 2 * nnb + go);
      for (int h = threadIdx.x; h < 2 * nz; h += NT) {
        const bool lo = h < nz;
        const int hq = lo ? q0 - nz + h : q0 + NT + (h - nz);
        const int hw = lo ? h : NT + h;
        const bool hin = hq >= 0 && hq < pl;
        const unsigned g = (unsigned)i * plb + (hin ? (unsigned)hq * CS : 0u);
        C hf[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) hf[c] = ldg(sin, c * nnb + g);
        // below: k2y and k2z; above: k2p (twice, the second unused)
        const C ha = ldg(K2, (lo ? nnb : 3 * nnb) + g);
        const C hb = ldg(K2, (lo ? 2 * nnb : 3 * nnb) + g);
#pragma unroll
        for (int c = 0; c < 4; ++c) sm[c * ws + hw] = hin ? hf[c] : C{R(0), R(0)};
        if (lo) {
          sm[4 * ws + hw] = hin ? ha : C{R(0), R(0)};
          sm[5 * ws + hw] = hin ? hb : C{R(0), R(0)};
        } else {
          sm[6 * ws + hw] = hin ? ha : C{R(0), R(0)};
        }
      }
      sm[4 * ws + w] = own ? ky : C{R(0), R(0)};
      sm[5 * ws + w] = own ? kz : C{R(0), R(0)};
#pragma unroll
      for (int c = 0; c < 4; ++c) sm[c * ws + w] = ca[c];
      sm[6 * ws + w] = k2p_c;
#pragma unroll
      for (int c = 0; c < 4; ++c) na[c] = ldc(sin, c, i + 1);
      const C k2x_n = ldc(K2, 0, i + 1), k2p_n = ldc(K2, 3, i + 1);
      __syncthreads();
      if (own) {
        const unsigned ob = (unsigned)i * plb + colb;
        FlatAcc<R, NT, WS> a{sm, w, nz, pa, ca, na, k2x_p, k2x_c, k2p_c, k2p_n};
        C rx, ry, rz, rp;
        rows4<R>(i, j, k, nx, ny, nz, axis_at_g(op, 0, i), axis_at_g(op, 1, j), axis_at_g(op, 2, k), a,
                 rx, ry, rz, rp);
        sto(sout, ob, rx);
        sto(sout, nnb + ob, ry);
        sto(sout, 2 * nnb + ob, rz);
        sto(sout, 3 * nnb + ob, rp);
        if (with_eps) eps_acc<R>(ca, rx, ry, rz, rp, eacc_re, eacc_im);
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
  }
  if (with_eps) {
    double v[2] = {eacc_re, eacc_im};
    block_sum<2>(v);
    if (threadIdx.x == 0) {
      part[(size_t)b * 2] = v[0];
      part[(size_t)b * 2 + 1] = v[1];
    }
  }
}

// batched stencil, one node per thread: work list of (system, x-plane,
// flat in-plane chunk of NT nodes) over all systems; eps partial per work
// item.  Used when the march's LDS window would not fit (very long z-rows).
template <typename R, int NT>
__global__ __launch_bounds__(NT) void kb_apply_direct(BatchDev B, const OpArgs<R>* __restrict__ ops,
                                                      const cplx<R>* __restrict__ in,
                                                      cplx<R>* __restrict__ out, int with_eps,
                                                      double* part) {
  using C = cplx<R>;
  const int b = (int)xcd_remap(blockIdx.x, gridDim.x);
  const int s = B.tsys[b];
  double eacc_re = 0.0, eacc_im = 0.0;
  const bool live = !(with_eps && B.st[s].done);
  if (live) {
    const OpArgs<R>& op = ops[B.op[s]];
    const int nx = op.nx, ny = op.ny, nz = op.nz;
    const int pl = ny * nz, nch = (pl + NT - 1) / NT;
    const int t = b - B.tstart[s];
    const int i = t / nch, q = (t - i * nch) * NT + (int)threadIdx.x;
    if (q < pl) {
      const int j = q / nz, k = q - j * nz;
      constexpr unsigned CS = sizeof(C);
      const C* sin = in + B.off[s];
      DirectAcc<R> a{sin, op.k2, (unsigned)op.nn * CS, (unsigned)pl * CS, (unsigned)nz * CS,
                     ((unsigned)i * pl + q) * CS, i, j, k, nx, ny, nz};
      C rx, ry, rz, rp;
      rows4<R>(i, j, k, nx, ny, nz, axis_at(op, 0, i), axis_at(op, 1, j), axis_at(op, 2, k), a,
               rx, ry, rz, rp);
      C* sout = out + B.off[s];
      sto(sout, a.ob, rx);
      sto(sout, a.nnb + a.ob, ry);
      sto(sout, 2 * a.nnb + a.ob, rz);
      sto(sout, 3 * a.nnb + a.ob, rp);
      if (with_eps) {
        const C in4[4] = {a.fv(0, 0), a.fv(1, 0), a.fv(2, 0), a.fv(3, 0)};
        eps_acc<R>(in4, rx, ry, rz, rp, eacc_re, eacc_im);
      }
    }
  }
  if (with_eps) {
    double v[2] = {eacc_re, eacc_im};
    block_sum<2>(v);
    if (threadIdx.x == 0) {
      part[(size_t)b * 2] = v[0];
      part[(size_t)b * 2 + 1] = v[1];
    }
  }
}

__global__ void kb_fin_eps(BatchDev B, const double* part) {
  const int s = blockIdx.x;
  QmrState* st = &B.st[s];
  if (st->done) return;
  double tot[2];
  b_fin_sum<2>(part, B.tstart[s], B.tstart[s + 1], tot);
  if (threadIdx.x == 0) {
    st->eps = {tot[0], tot[1]};
    st->beta = cdiv(st->eps, st->del);
    if (is_zero(st->eps) || !finite_c(st->beta)) {
      st->status = ST_BREAK;
      st->done = 1;
    }
  }
}

template <typename V, bool NTS = false>
__global__ void kb_vtnorm(BatchDev B, const vec<V>* __restrict__ t, const vec<V>* __restrict__ v,
                          double* part) {
  const BItem it = b_item(B);
  const QmrState* st = &B.st[it.s];
  if (st->done) return;
  const c128 beta = st->beta;
  double a[1] = {0.0};
  struct L {
    vec<V> t, v;
  };
  b_loop2(it, [&](long long q) { return L{ldv<NTS>(t + q), ldv<NTS>(v + q)}; },
          [&](long long, const L& l) { a[0] += abs2(up(dn<V>(up(l.t) - up(l.v) * beta))); });
  block_sum<1>(a);
  block_partial<1>(a, part);
}

__global__ void kb_fin_vtnorm(BatchDev B, const double* part) {
  const int s = blockIdx.x;
  QmrState* st = &B.st[s];
  if (st->done) return;
  double tot[1];
  b_fin_sum<1>(part, B.bstart[s], B.bstart[s + 1], tot);
  if (threadIdx.x == 0) vt_scalars(st, sqrt(tot[0]));
}

template <typename V, bool NTS = false>
__global__ void kb_update(BatchDev B, const vec<V>* __restrict__ t, vec<V>* __restrict__ s_,
                          vec<V>* __restrict__ r, vec<V>* __restrict__ v,
                          const vec<V>* __restrict__ invd, double* part) {
  const BItem it = b_item(B);
  QmrState* st = &B.st[it.s];
  if (st->done) return;
  const c128 eta = st->eta, beta = st->beta;
  const double k2 = st->k2s, rho = st->rho_np;
  const long long io = B.opoff[B.op[it.s]] - B.off[it.s];
  // r is stored on even iterations only, as in k_update
  const bool full = (st->iter & 1) == 0;
  double a[3] = {0.0, 0.0, 0.0};
  struct L {
    vec<V> t, s, r, v, i;
  };
  auto load = [&](long long q) {
    return L{ldv<NTS>(t + q), ldv<NTS>(s_ + q), ldv<NTS>(r + q), ldv<NTS>(v + q),
             ldv<false>(invd + q + io)};
  };
  if (full)
    b_loop2(it, load, [&](long long q, const L& l) {
      update_elem<V, NTS, UPD_CATCHUP>(q, l.t, l.s, l.r, l.v, l.i, eta, beta, k2, rho, s_, r, v, a);
    });
  else
    b_loop2(it, load, [&](long long q, const L& l) {
      update_elem<V, NTS, UPD_SKIP>(q, l.t, l.s, l.r, l.v, l.i, eta, beta, k2, rho, s_, r, v, a);
    });
  if ((int)blockIdx.x == B.bstart[it.s] && threadIdx.x == 0) {
    st->lag = full ? 0 : 1;
    st->upd_it = st->iter;
  }
  block_sum<3>(a);
  block_partial<3>(a, part);
}

// Completes every system of a batch as k_finish completes a single solve:
// the d and x updates its last p-update did not reach, x left one step
// behind, and r left one step behind, from the system's own flags.
template <typename V>
__global__ void kb_finish(BatchDev B, vec<V>* __restrict__ x, vec<V>* __restrict__ d,
                          const vec<V>* __restrict__ p, vec<V>* __restrict__ r,
                          const vec<V>* __restrict__ s_) {
  const BItem it = b_item(B);
  const QmrState* st = &B.st[it.s];
  const int pend = st->upd_it > st->dx_done ? st->upd_it : 0;
  const bool xlag = !pend && (st->dx_done & 1);
  const bool rlag = st->upd_it > 0 && st->lag;
  if (!pend && !xlag && !rlag) return;
  const c128 eta = st->eta;
  const double k2 = st->k2s;
  B_LOOP(q, it) {
    if (pend) {
      const vec<V> dq = d[q];
      const vec<V> ddv = dn<V>(up(p[q]) * eta + up(dq) * k2);
      d[q] = ddv;
      x[q] = (pend & 1) ? dn<V>(up(x[q]) + up(ddv))
                        : dn<V>(up(dn<V>(up(x[q]) + up(dq))) + up(ddv));
    } else if (xlag) {
      x[q] = dn<V>(up(x[q]) + up(d[q]));
    }
    if (rlag) r[q] = dn<V>(up(r[q]) - up(s_[q]));
  }
}

__global__ void kb_fin_update(BatchDev B, const double* part) {
  const int s = blockIdx.x;
  QmrState* st = &B.st[s];
  if (st->done) return;
  double tot[3];
  b_fin_sum<3>(part, B.bstart[s], B.bstart[s + 1], tot);
  if (threadIdx.x == 0) update_scalars(st, tot, B.hist ? B.hist + (size_t)s * B.hist_stride : nullptr);
}

// per-system ||b - t|| and ||b||
template <typename V>
__global__ void kb_resnorm(BatchDev B, const vec<V>* __restrict__ b, const vec<V>* __restrict__ t,
                           double* part) {
  const BItem it = b_item(B);
  double a[2] = {0.0, 0.0};
  B_LOOP(q, it) {
    const c128 bq = up(b[q]);
    a[0] += abs2(bq - up(t[q]));
    a[1] += abs2(bq);
  }
  block_sum<2>(a);
  block_partial<2>(a, part);
}
__global__ void kb_fin_resnorm(BatchDev B, const double* part, double* out) {
  const int s = blockIdx.x;
  double tot[2];
  b_fin_sum<2>(part, B.bstart[s], B.bstart[s + 1], tot);
  if (threadIdx.x == 0) {
    out[2 * s] = sqrt(tot[0]);
    out[2 * s + 1] = sqrt(tot[1]);
  }
}

// ---------------------------------------------------------------------------
// batched mixed precision: per-system FP64 refinement around a complex64
// inner batch.  After kb_resnorm/kb_fin_resnorm (res[2s] = ||b-Kx||,
// res[2s+1] = ||b||), kb_inner_setup decides per system: converged (scale 0,
// the inner right-hand side is 0 and its inner solve stops at once) or the
// inner target; the inner rhs is (b - Kx)/||b - Kx|| in complex64.
__global__ void kb_inner_setup(int nsys, const double* __restrict__ res, double tol,
                               double floor_tol, const int* __restrict__ budget,
                               double* __restrict__ scale, double* __restrict__ itol,
                               int* __restrict__ nopen) {
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < nsys; s += gridDim.x * blockDim.x) {
    const double rn = res[2 * s], bn = res[2 * s + 1];
    const bool done = !(rn >= tol * bn) || bn == 0.0;  // also NaN-safe: NaN -> not open
    const bool open = !done && isfinite(rn) && rn > 0.0 && budget[s] > 0;
    scale[s] = open ? 1.0 / rn : 0.0;
    itol[s] = open ? fmax(0.5 * tol * bn / rn, floor_tol) : 1.0;
    if (open) atomicAdd(nopen, 1);
  }
}

__global__ void kb_inner_rhs(BatchDev B, const c128* __restrict__ b, const c128* __restrict__ t,
                             const double* __restrict__ scale, c64* __restrict__ bi,
                             c64* __restrict__ e) {
  const BItem it = b_item(B);
  const double sc = scale[it.s];
  B_LOOP(q, it) {
    bi[q] = dn<float>((b[q] - t[q]) * sc);
    e[q] = c64{0.f, 0.f};
  }
}

__global__ void kb_fin_norm_b_tol(BatchDev B, const double* part, const double* __restrict__ itol) {
  const int s = blockIdx.x;
  double tot[1];
  b_fin_sum<1>(part, B.bstart[s], B.bstart[s + 1], tot);
  if (threadIdx.x == 0) {
    B.st[s].rhsnorm = sqrt(tot[0]);
    B.st[s].target = itol[s] * B.st[s].rhsnorm;
  }
}

// x += ||b - Kx|| e for the systems that were open
__global__ void kb_refine(BatchDev B, c128* __restrict__ x, const c64* __restrict__ e,
                          const double* __restrict__ res, const double* __restrict__ scale) {
  const BItem it = b_item(B);
  if (scale[it.s] == 0.0) return;
  const double rn = res[2 * it.s];
  B_LOOP(q, it) x[q] = x[q] + up(e[q]) * rn;
}
