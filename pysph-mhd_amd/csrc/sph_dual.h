// sph_dual.h -- forward-mode dual numbers for the tangent-linear step.
//
// Dual<N> carries a value and N directional derivatives (host and device
// code: the adjoint step's self-check runs them on the host).  Arithmetic
// follows the chain rule; comparisons, min/max and branches act on the value,
// so a dual evaluation takes exactly the primal's control path and its
// derivative is the derivative of that discrete map.  Non-differentiable
// points get the one-sided choice a branch on the value makes (sqrt and cbrt
// at 0 have derivative 0; fabs at 0 passes its argument's derivative on).
#pragma once

#include <hip/hip_runtime.h>

namespace sph {

template <int N>
struct Dual {
    double v;
    double d[N];
    __host__ __device__ constexpr Dual() : v(0.0), d{} {}
    __host__ __device__ constexpr Dual(double x) : v(x), d{} {}
};

#define DUAL_LOOP for (int k_ = 0; k_ < N; ++k_)

template <int N>
__host__ __device__ __forceinline__ Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v + b.v);
    DUAL_LOOP r.d[k_] = a.d[k_] + b.d[k_];
    return r;
}
template <int N>
__host__ __device__ __forceinline__ Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v - b.v);
    DUAL_LOOP r.d[k_] = a.d[k_] - b.d[k_];
    return r;
}
template <int N>
__host__ __device__ __forceinline__ Dual<N> operator-(const Dual<N>& a) {
    Dual<N> r(-a.v);
    DUAL_LOOP r.d[k_] = -a.d[k_];
    return r;
}
template <int N>
__host__ __device__ __forceinline__ Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v * b.v);
    DUAL_LOOP r.d[k_] = a.d[k_] * b.v + a.v * b.d[k_];
    return r;
}
template <int N>
__host__ __device__ __forceinline__ Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
    const double q = a.v / b.v;
    Dual<N> r(q);
    const double ib = 1.0 / b.v;
    DUAL_LOOP r.d[k_] = (a.d[k_] - q * b.d[k_]) * ib;
    return r;
}
template <int N>
__host__ __device__ __forceinline__ Dual<N> operator+(const Dual<N>& a, double b) {
    Dual<N> r = a;
    r.v += b;
    return r;
}
template <int N>
__host__ __device__ __forceinline__ Dual<N> operator+(double a, const Dual<N>& b) { return b + a; }
template <int N>
__host__ __device__ __forceinline__ Dual<N> operator-(const Dual<N>& a, double b) {
    Dual<N> r = a;
    r.v -= b;
    return r;
}
template <int N>
__host__ __device__ __forceinline__ Dual<N> operator-(double a, const Dual<N>& b) {
    Dual<N> r(a - b.v);
    DUAL_LOOP r.d[k_] = -b.d[k_];
    return r;
}
template <int N>
__host__ __device__ __forceinline__ Dual<N> operator*(const Dual<N>& a, double b) {
    Dual<N> r(a.v * b);
    DUAL_LOOP r.d[k_] = a.d[k_] * b;
    return r;
}
template <int N>
__host__ __device__ __forceinline__ Dual<N> operator*(double a, const Dual<N>& b) { return b * a; }
template <int N>
__host__ __device__ __forceinline__ Dual<N> operator/(const Dual<N>& a, double b) {
    Dual<N> r(a.v / b);
    const double ib = 1.0 / b;
    DUAL_LOOP r.d[k_] = a.d[k_] * ib;
    return r;
}
template <int N>
__host__ __device__ __forceinline__ Dual<N> operator/(double a, const Dual<N>& b) {
    const double q = a / b.v;
    Dual<N> r(q);
    const double ib = 1.0 / b.v;
    DUAL_LOOP r.d[k_] = -q * b.d[k_] * ib;
    return r;
}
template <int N, class B>
__host__ __device__ __forceinline__ Dual<N>& operator+=(Dual<N>& a, const B& b) { return a = a + b; }
template <int N, class B>
__host__ __device__ __forceinline__ Dual<N>& operator-=(Dual<N>& a, const B& b) { return a = a - b; }
template <int N, class B>
__host__ __device__ __forceinline__ Dual<N>& operator*=(Dual<N>& a, const B& b) { return a = a * b; }
template <int N, class B>
__host__ __device__ __forceinline__ Dual<N>& operator/=(Dual<N>& a, const B& b) { return a = a / b; }

// comparisons on the value
__host__ __device__ __forceinline__ double val(double x) { return x; }
template <int N>
__host__ __device__ __forceinline__ double val(const Dual<N>& x) { return x.v; }
template <int N, class B>
__host__ __device__ __forceinline__ bool operator<(const Dual<N>& a, const B& b) { return a.v < val(b); }
template <int N>
__host__ __device__ __forceinline__ bool operator<(double a, const Dual<N>& b) { return a < b.v; }
template <int N, class B>
__host__ __device__ __forceinline__ bool operator>(const Dual<N>& a, const B& b) { return a.v > val(b); }
template <int N>
__host__ __device__ __forceinline__ bool operator>(double a, const Dual<N>& b) { return a > b.v; }
template <int N, class B>
__host__ __device__ __forceinline__ bool operator<=(const Dual<N>& a, const B& b) { return a.v <= val(b); }
template <int N>
__host__ __device__ __forceinline__ bool operator<=(double a, const Dual<N>& b) { return a <= b.v; }
template <int N, class B>
__host__ __device__ __forceinline__ bool operator>=(const Dual<N>& a, const B& b) { return a.v >= val(b); }
template <int N>
__host__ __device__ __forceinline__ bool operator>=(double a, const Dual<N>& b) { return a >= b.v; }

template <int N>
__host__ __device__ __forceinline__ Dual<N> sqrt(const Dual<N>& a) {
    const double s = ::sqrt(a.v);
    Dual<N> r(s);
    const double g = s > 0.0 ? 0.5 / s : 0.0;
    DUAL_LOOP r.d[k_] = a.d[k_] * g;
    return r;
}
template <int N>
__host__ __device__ __forceinline__ Dual<N> cbrt(const Dual<N>& a) {
    const double s = ::cbrt(a.v);
    Dual<N> r(s);
    const double g = s != 0.0 ? 1.0 / (3.0 * s * s) : 0.0;
    DUAL_LOOP r.d[k_] = a.d[k_] * g;
    return r;
}
template <int N>
__host__ __device__ __forceinline__ Dual<N> exp(const Dual<N>& a) {
    const double e = ::exp(a.v);
    Dual<N> r(e);
    DUAL_LOOP r.d[k_] = a.d[k_] * e;
    return r;
}
template <int N>
__host__ __device__ __forceinline__ Dual<N> fabs(const Dual<N>& a) { return a.v < 0.0 ? -a : a; }
// max/min pick an operand by value (ties: the first), as the primal's fmax
// does for distinct values
template <int N>
__host__ __device__ __forceinline__ Dual<N> fmax(const Dual<N>& a, const Dual<N>& b) { return a.v >= b.v ? a : b; }
template <int N>
__host__ __device__ __forceinline__ Dual<N> fmax(const Dual<N>& a, double b) { return a.v >= b ? a : Dual<N>(b); }
template <int N>
__host__ __device__ __forceinline__ Dual<N> fmax(double a, const Dual<N>& b) { return a >= b.v ? Dual<N>(a) : b; }
template <int N>
__host__ __device__ __forceinline__ Dual<N> fmin(const Dual<N>& a, const Dual<N>& b) { return a.v <= b.v ? a : b; }
template <int N>
__host__ __device__ __forceinline__ Dual<N> fmin(const Dual<N>& a, double b) { return a.v <= b ? a : Dual<N>(b); }
template <int N>
__host__ __device__ __forceinline__ Dual<N> fmin(double a, const Dual<N>& b) { return a <= b.v ? Dual<N>(a) : b; }
template <int N>
__host__ __device__ __forceinline__ bool isfinite(const Dual<N>& a) { return ::isfinite(a.v); }

#undef DUAL_LOOP

}  // namespace sph
