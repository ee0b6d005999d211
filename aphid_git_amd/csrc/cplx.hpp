// cplx.hpp -- minimal complex arithmetic for device and host code.
//
// Interleaved (re, im) storage, identical to Fortran complex(8)/complex(4),
// aligned so that one complex value is one 16-byte (complex128) or 8-byte
// (complex64) vector load.
#pragma once
#include <hip/hip_runtime.h>

template <typename R>
struct alignas(2 * sizeof(R)) cplx {
  R re, im;
};
using c128 = cplx<double>;
using c64 = cplx<float>;

#define HD __host__ __device__ __forceinline__

template <typename R> HD cplx<R> mk(R a, R b) { return cplx<R>{a, b}; }
template <typename R> HD cplx<R> operator+(cplx<R> a, cplx<R> b) { return {a.re + b.re, a.im + b.im}; }
template <typename R> HD cplx<R> operator-(cplx<R> a, cplx<R> b) { return {a.re - b.re, a.im - b.im}; }
template <typename R> HD cplx<R> operator-(cplx<R> a) { return {-a.re, -a.im}; }
// product in the Fortran rule (no NaN recovery)
template <typename R> HD cplx<R> operator*(cplx<R> a, cplx<R> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename R> HD cplx<R> operator*(R s, cplx<R> a) { return {s * a.re, s * a.im}; }
template <typename R> HD cplx<R> operator*(cplx<R> a, R s) { return {a.re * s, a.im * s}; }
template <typename R> HD cplx<R>& operator+=(cplx<R>& a, cplx<R> b) { a.re += b.re; a.im += b.im; return a; }
template <typename R> HD cplx<R> divr(cplx<R> a, R s) { return {a.re / s, a.im / s}; }
template <typename R> HD R abs2(cplx<R> a) { return a.re * a.re + a.im * a.im; }
// Smith's division
template <typename R> HD cplx<R> cdiv(cplx<R> a, cplx<R> b) {
  R ar = a.re, ai = a.im, br = b.re, bi = b.im;
  if ((br < 0 ? -br : br) >= (bi < 0 ? -bi : bi)) {
    R q = bi / br, den = br + bi * q;
    return {(ar + ai * q) / den, (ai - ar * q) / den};
  }
  R q = br / bi, den = bi + br * q;
  return {(ar * q + ai) / den, (ai * q - ar) / den};
}
template <typename To, typename From> HD cplx<To> ccast(cplx<From> a) {
  return {static_cast<To>(a.re), static_cast<To>(a.im)};
}
HD bool is_zero(c128 a) { return a.re == 0.0 && a.im == 0.0; }
