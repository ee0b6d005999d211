// qmr.hpp -- fused vector kernels of the Jacobi-QMR iteration
// (src/aphid_qmr.f90:297-390), templated on the storage precision V of the
// Krylov vectors (double: the plain complex128 path; float: the complex64
// path).  Element arithmetic, inner products and all recurrence scalars are
// FP64.  One iteration is four launches:
//   A  p = M^-1 v - k1 p                               (:306-320)
//   B  t = K p, eps = p^T t, beta = eps/del            (stencil.hpp, :324-330)
//   C  rho' = ||t - beta v||, theta, c, eta, k2        (:332-345)
//   D  d, x, s, r, v updates, ||r||, stop test and the next del = v^T M^-1 v
//      (:347-390 and :306-313 of the next iteration)
// Every kernel returns at once when the state says the solve is done, so a
// captured graph of several iterations can be replayed without a host
// decision per iteration.
#pragma once
#include "reduce.hpp"

template <typename V>
using vec = cplx<V>;

template <typename V>
__device__ __forceinline__ c128 up(vec<V> a) { return {(double)a.re, (double)a.im}; }
template <typename V>
__device__ __forceinline__ vec<V> dn(c128 a) { return {(V)a.re, (V)a.im}; }

#define GRID_LOOP(q, n)                                                      \
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < (n); \
       q += (long long)gridDim.x * blockDim.x)

// ||b||^2 and the stop target (src/aphid_qmr.f90:169-172,386)
template <typename V>
__global__ void k_norm_b(const vec<V>* __restrict__ b, long long n, QmrState* st, double* part,
                         double tol) {
  double v[1] = {0.0}, tot[1];
  GRID_LOOP(q, n) v[0] += abs2(up(b[q]));
  block_sum<1>(v);
  if (grid_finish<1>(v, part, &st->ticket, tot) && threadIdx.x == 0) {
    st->rhsnorm = sqrt(tot[0]);
    st->target = tol * st->rhsnorm;
  }
}

// r = b - t (t = K x0, or nothing when x0 = 0), rho = ||r|| (:245-273)
template <typename V>
__global__ void k_init_r(const vec<V>* __restrict__ b, const vec<V>* __restrict__ t,
                         vec<V>* __restrict__ r, long long n, QmrState* st, double* part) {
  double v[1] = {0.0}, tot[1];
  GRID_LOOP(q, n) {
    c128 e = up(b[q]);
    if (t) e = e - up(t[q]);
    const vec<V> ev = dn<V>(e);
    r[q] = ev;
    v[0] += abs2(up(ev));
  }
  block_sum<1>(v);
  if (grid_finish<1>(v, part, &st->ticket, tot) && threadIdx.x == 0) st->rho_np = sqrt(tot[0]);
}

// v = r/rho; p = d = s = 0; del = v^T M^-1 v; the loop-entry scalars
// (:260-285) and k1 of iteration 1 (:313)
template <typename V>
__global__ void k_init_v(const vec<V>* __restrict__ r, const vec<V>* __restrict__ invd,
                         vec<V>* __restrict__ v, vec<V>* __restrict__ p, vec<V>* __restrict__ d,
                         vec<V>* __restrict__ s, long long n, QmrState* st, double* part) {
  const double rho = st->rho_np;
  double a[2] = {0.0, 0.0}, tot[2];
  const vec<V> z = {V(0), V(0)};
  GRID_LOOP(q, n) {
    const vec<V> vv = dn<V>(divr(up(r[q]), rho));
    v[q] = vv;
    p[q] = z;
    d[q] = z;
    s[q] = z;
    const c128 vd = up(vv);
    const c128 tt = vd * up(invd[q]);
    const c128 pr = vd * tt;
    a[0] += pr.re;
    a[1] += pr.im;
  }
  block_sum<2>(a);
  if (grid_finish<2>(a, part, &st->ticket, tot) && threadIdx.x == 0) {
    st->c_n = 1.0;
    st->eps = {1.0, 0.0};
    st->theta_n = 0.0;
    st->eta = {-1.0, 0.0};
    st->iter = 1;
    st->done = 0;
    st->status = ST_MAXIT;
    st->rnorm = rho;
    // start of iteration 1
    st->c_nm = st->c_n;
    st->rho_n = rho;
    st->theta_nm = st->theta_n;
    st->del = {tot[0], tot[1]};
    st->k1 = cdiv(rho * st->del, st->eps);
    if (!(rho > 0.0) || !isfinite(rho)) {  // b = 0 (or x0 exact): nothing to do
      st->done = 1;
      st->status = (rho == 0.0) ? ST_CONV : ST_BREAK;
      st->iter = 0;
    } else if (is_zero(st->del) || !finite_c(st->k1)) {
      st->done = 1;
      st->status = ST_BREAK;
    }
  }
}

// A: p = M^-1 v - p k1 (:306-320)
template <typename V>
__global__ void k_pupdate(const vec<V>* __restrict__ v, const vec<V>* __restrict__ invd,
                          vec<V>* __restrict__ p, long long n, const QmrState* st) {
  if (st->done) return;
  const c128 k1 = st->k1;
  GRID_LOOP(q, n) p[q] = dn<V>(up(v[q]) * up(invd[q]) - up(p[q]) * k1);
}

// C: rho' = ||t - beta v|| and the step-4 scalars (:332-345)
template <typename V>
__global__ void k_vtnorm(const vec<V>* __restrict__ t, const vec<V>* __restrict__ v, long long n,
                         QmrState* st, double* part) {
  if (st->done) return;
  const c128 beta = st->beta;
  double a[1] = {0.0}, tot[1];
  GRID_LOOP(q, n) a[0] += abs2(up(dn<V>(up(t[q]) - up(v[q]) * beta)));
  block_sum<1>(a);
  if (grid_finish<1>(a, part, &st->ticket, tot) && threadIdx.x == 0) {
    const double rho = sqrt(tot[0]);
    st->rho_np = rho;
    const double bb = beta.re * beta.re + beta.im * beta.im;
    st->theta_n = rho * rho / st->c_nm / bb;
    st->c_n = 1.0 / (1.0 + st->theta_n);
    const c128 num = st->c_n * (st->rho_n * (-st->eta));
    st->eta = cdiv(num, st->c_nm * beta);
    st->k2s = st->theta_nm * st->c_n;
    if (!(rho > 0.0) || !isfinite(rho) || !finite_c(st->eta)) {
      st->status = ST_BREAK;
      st->done = 1;
    }
  }
}

// D: d = eta p + k2 d; x += d; s = eta t + k2 s; r -= s; v = (t - beta v)/rho';
// ||r|| and the stop test (:347-390), then del = v^T M^-1 v and k1 for the
// next iteration (:306-313).
template <typename V>
__global__ void k_update(const vec<V>* __restrict__ p, vec<V>* __restrict__ d, vec<V>* __restrict__ x,
                         const vec<V>* __restrict__ t, vec<V>* __restrict__ s, vec<V>* __restrict__ r,
                         vec<V>* __restrict__ v, const vec<V>* __restrict__ invd, long long n,
                         QmrState* st, double* part, double* hist) {
  if (st->done) return;
  const c128 eta = st->eta, beta = st->beta;
  const double k2 = st->k2s, rho = st->rho_np;
  double a[3] = {0.0, 0.0, 0.0}, tot[3];
  GRID_LOOP(q, n) {
    const c128 tq = up(t[q]);
    const c128 dd = up(p[q]) * eta + up(d[q]) * k2;
    const vec<V> ddv = dn<V>(dd);
    d[q] = ddv;
    x[q] = dn<V>(up(x[q]) + up(ddv));
    const vec<V> ssv = dn<V>(tq * eta + up(s[q]) * k2);
    s[q] = ssv;
    const vec<V> rrv = dn<V>(up(r[q]) - up(ssv));
    r[q] = rrv;
    const vec<V> vt = dn<V>(tq - up(v[q]) * beta);
    const vec<V> vv = dn<V>(divr(up(vt), rho));
    v[q] = vv;
    a[0] += abs2(up(rrv));
    const c128 vd = up(vv);
    const c128 pr = vd * (vd * up(invd[q]));
    a[1] += pr.re;
    a[2] += pr.im;
  }
  block_sum<3>(a);
  if (grid_finish<3>(a, part, &st->ticket, tot) && threadIdx.x == 0) {
    const double rnorm = sqrt(tot[0]);
    st->rnorm = rnorm;
    const int it = st->iter;
    if (hist && it >= 1 && it <= st->hist_cap) hist[it - 1] = rnorm;
    // stop test and counter exactly as src/aphid_qmr.f90:386-390
    if (rnorm < st->target) {
      st->status = ST_CONV;
      st->done = 1;
    } else if (!isfinite(rnorm)) {
      st->status = ST_BREAK;
      st->done = 1;
    } else if (it <= st->max_it) {
      st->iter = it + 1;
      if (it + 1 > st->max_it) {
        st->status = ST_MAXIT;
        st->done = 1;
      }
    } else {
      st->status = ST_MAXIT;
      st->done = 1;
    }
    if (!st->done) {  // start of the next iteration (:300-313)
      st->c_nm = st->c_n;
      st->rho_n = st->rho_np;
      st->theta_nm = st->theta_n;
      st->del = {tot[1], tot[2]};
      st->k1 = cdiv(st->rho_n * st->del, st->eps);
      if (is_zero(st->del) || !finite_c(st->k1)) {
        st->status = ST_BREAK;
        st->done = 1;
      }
    }
  }
}

// ||b - t||^2 / ||b||^2 for the true residual (t = K x)
template <typename V>
__global__ void k_resnorm(const vec<V>* __restrict__ b, const vec<V>* __restrict__ t, long long n,
                          double* out, double* part, unsigned* ticket) {
  double a[2] = {0.0, 0.0}, tot[2];
  GRID_LOOP(q, n) {
    const c128 bq = up(b[q]);
    a[0] += abs2(bq - up(t[q]));
    a[1] += abs2(bq);
  }
  block_sum<2>(a);
  if (grid_finish<2>(a, part, ticket, tot) && threadIdx.x == 0) {
    out[0] = sqrt(tot[0]);
    out[1] = sqrt(tot[1]);
  }
}
