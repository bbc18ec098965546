// rtw_dual.h -- forward-mode dual numbers for the local Jacobians of the adjoint.
//
// Dual<N> carries a value and N tangents.  The geodesic right-hand side is a
// template on its scalar type, so evaluating it on Dual<7> seeded in
// (r, theta, k^t, k^r, k^theta, k^phi, spin) gives the Jacobian rows the
// reverse sweep of the RK45 adjoint needs, from the same source as the forward
// kernels.  Comparisons look at the value only (the matter test and the
// branches of the scene functions are piecewise constant).
#pragma once
#include <hip/hip_runtime.h>

namespace rtw {

template <int N> struct Dual {
  double v;
  double d[N];
  __device__ __forceinline__ Dual() : v(0) {
#pragma unroll
    for (int i = 0; i < N; ++i) d[i] = 0;
  }
  __device__ __forceinline__ Dual(double x) : v(x) {
#pragma unroll
    for (int i = 0; i < N; ++i) d[i] = 0;
  }
  __device__ __forceinline__ static Dual seed(double x, int k) {
    Dual r(x);
    r.d[k] = 1.0;
    return r;
  }
  __device__ __forceinline__ explicit operator double() const { return v; }
  __device__ __forceinline__ Dual& operator+=(const Dual& b) { return *this = *this + b; }
  __device__ __forceinline__ Dual& operator-=(const Dual& b) { return *this = *this - b; }
  __device__ __forceinline__ Dual& operator*=(const Dual& b) { return *this = *this * b; }

  // value-and-chain-rule helper: f(v) with f'(v) = g
  __device__ __forceinline__ Dual chain(double fv, double g) const {
    Dual r(fv);
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = g * d[i];
    return r;
  }

  friend __device__ __forceinline__ Dual operator+(const Dual& a, const Dual& b) {
    Dual r(a.v + b.v);
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
    return r;
  }
  friend __device__ __forceinline__ Dual operator-(const Dual& a, const Dual& b) {
    Dual r(a.v - b.v);
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
    return r;
  }
  friend __device__ __forceinline__ Dual operator-(const Dual& a) {
    Dual r(-a.v);
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
    return r;
  }
  friend __device__ __forceinline__ Dual operator*(const Dual& a, const Dual& b) {
    Dual r(a.v * b.v);
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
  }
  friend __device__ __forceinline__ Dual operator/(const Dual& a, const Dual& b) {
    const double q = a.v / b.v, ib = 1.0 / b.v;
    Dual r(q);
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - q * b.d[i]) * ib;
    return r;
  }
  friend __device__ __forceinline__ Dual operator+(const Dual& a, double b) { return a + Dual(b); }
  friend __device__ __forceinline__ Dual operator+(double a, const Dual& b) { return Dual(a) + b; }
  friend __device__ __forceinline__ Dual operator-(const Dual& a, double b) { return a - Dual(b); }
  friend __device__ __forceinline__ Dual operator-(double a, const Dual& b) { return Dual(a) - b; }
  friend __device__ __forceinline__ Dual operator*(const Dual& a, double b) {
    Dual r(a.v * b);
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b;
    return r;
  }
  friend __device__ __forceinline__ Dual operator*(double a, const Dual& b) { return b * a; }
  friend __device__ __forceinline__ Dual operator/(const Dual& a, double b) { return a / Dual(b); }
  friend __device__ __forceinline__ Dual operator/(double a, const Dual& b) { return Dual(a) / b; }

  friend __device__ __forceinline__ bool operator<(const Dual& a, const Dual& b) { return a.v < b.v; }
  friend __device__ __forceinline__ bool operator<=(const Dual& a, const Dual& b) { return a.v <= b.v; }
  friend __device__ __forceinline__ bool operator>(const Dual& a, const Dual& b) { return a.v > b.v; }
  friend __device__ __forceinline__ bool operator>=(const Dual& a, const Dual& b) { return a.v >= b.v; }
};

// The double overloads stay visible next to the Dual ones for unqualified calls
// in this namespace.
using ::fabs;
using ::sqrt;

template <int N> __device__ __forceinline__ Dual<N> sqrt(const Dual<N>& x) {
  const double s = ::sqrt(x.v);
  return x.chain(s, 0.5 / s);
}
template <int N> __device__ __forceinline__ Dual<N> fabs(const Dual<N>& x) { return x.v < 0 ? -x : x; }
template <int N> __device__ __forceinline__ Dual<N> exp_(const Dual<N>& x) {
  const double e = ::exp(x.v);
  return x.chain(e, e);
}
template <int N> __device__ __forceinline__ Dual<N> log10_(const Dual<N>& x) {
  return x.chain(::log10(x.v), 1.0 / (x.v * 2.302585092994046));
}
template <int N> __device__ __forceinline__ void sc_(const Dual<N>& x, Dual<N>& s, Dual<N>& c) {
  double sv, cv;
  ::sincos(x.v, &sv, &cv);
  s = x.chain(sv, cv);
  c = x.chain(cv, -sv);
}

}  // namespace rtw
