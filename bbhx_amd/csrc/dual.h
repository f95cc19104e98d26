// Forward-mode dual numbers for the gradient kernels.
//
// Dual<N> carries a value and N partial derivatives. The waveform and response
// device code is written once as templates over the scalar type; instantiated
// with double it is the forward path, with Dual<N> it also propagates N
// tangent directions (chain rule per operation). Comparisons and branches use
// the value only, so a Dual evaluation takes exactly the branches of the
// double evaluation at the same point.
#pragma once

#include "common.h"

namespace bbhx {

// the double versions stay visible next to the Dual overloads below
using ::atan;
using ::cbrt;
using ::cos;
using ::exp;
using ::fabs;
using ::hypot;
using ::log;
using ::pow;
using ::sin;
using ::sqrt;

template <int N>
struct Dual {
    double v;
    double d[N];
    __host__ __device__ Dual() : v(0.0)
    {
#pragma unroll
        for (int i = 0; i < N; i++) d[i] = 0.0;
    }
    __host__ __device__ Dual(double x) : v(x)
    {
#pragma unroll
        for (int i = 0; i < N; i++) d[i] = 0.0;
    }
    // value x with tangent 1 in direction k
    __host__ __device__ static Dual seed(double x, int k)
    {
        Dual r(x);
        r.d[k] = 1.0;
        return r;
    }
};

template <class S>
struct scalar_traits;
template <>
struct scalar_traits<double> {
    static constexpr int N = 0;
};
template <int N_>
struct scalar_traits<Dual<N_>> {
    static constexpr int N = N_;
};

__host__ __device__ inline double value(double x) { return x; }
template <int N>
__host__ __device__ inline double value(const Dual<N>& x)
{
    return x.v;
}

// r = a * x (scaled tangents), helper for the chain rule: f(u) -> (f, f' du)
template <int N>
__device__ inline Dual<N> chain(double f, double fp, const Dual<N>& u)
{
    Dual<N> r;
    r.v = f;
#pragma unroll
    for (int i = 0; i < N; i++) r.d[i] = fp * u.d[i];
    return r;
}

template <int N>
__device__ inline Dual<N> operator+(const Dual<N>& a, const Dual<N>& b)
{
    Dual<N> r;
    r.v = a.v + b.v;
#pragma unroll
    for (int i = 0; i < N; i++) r.d[i] = a.d[i] + b.d[i];
    return r;
}
template <int N>
__device__ inline Dual<N> operator-(const Dual<N>& a, const Dual<N>& b)
{
    Dual<N> r;
    r.v = a.v - b.v;
#pragma unroll
    for (int i = 0; i < N; i++) r.d[i] = a.d[i] - b.d[i];
    return r;
}
template <int N>
__device__ inline Dual<N> operator-(const Dual<N>& a)
{
    Dual<N> r;
    r.v = -a.v;
#pragma unroll
    for (int i = 0; i < N; i++) r.d[i] = -a.d[i];
    return r;
}
template <int N>
__device__ inline Dual<N> operator*(const Dual<N>& a, const Dual<N>& b)
{
    Dual<N> r;
    r.v = a.v * b.v;
#pragma unroll
    for (int i = 0; i < N; i++) r.d[i] = fma(a.d[i], b.v, a.v * b.d[i]);
    return r;
}
template <int N>
__device__ inline Dual<N> operator/(const Dual<N>& a, const Dual<N>& b)
{
    // (a / b)' = (a' - q b') / b
    const double ib = 1.0 / b.v;
    Dual<N> r;
    r.v = a.v / b.v;
#pragma unroll
    for (int i = 0; i < N; i++) r.d[i] = fma(-r.v, b.d[i], a.d[i]) * ib;
    return r;
}

// mixed with plain doubles
template <int N>
__device__ inline Dual<N> operator+(const Dual<N>& a, double b)
{
    Dual<N> r = a;
    r.v += b;
    return r;
}
template <int N>
__device__ inline Dual<N> operator+(double a, const Dual<N>& b)
{
    return b + a;
}
template <int N>
__device__ inline Dual<N> operator-(const Dual<N>& a, double b)
{
    Dual<N> r = a;
    r.v -= b;
    return r;
}
template <int N>
__device__ inline Dual<N> operator-(double a, const Dual<N>& b)
{
    Dual<N> r;
    r.v = a - b.v;
#pragma unroll
    for (int i = 0; i < N; i++) r.d[i] = -b.d[i];
    return r;
}
template <int N>
__device__ inline Dual<N> operator*(const Dual<N>& a, double b)
{
    Dual<N> r;
    r.v = a.v * b;
#pragma unroll
    for (int i = 0; i < N; i++) r.d[i] = a.d[i] * b;
    return r;
}
template <int N>
__device__ inline Dual<N> operator*(double a, const Dual<N>& b)
{
    return b * a;
}
template <int N>
__device__ inline Dual<N> operator/(const Dual<N>& a, double b)
{
    Dual<N> r;
    r.v = a.v / b;
    const double ib = 1.0 / b;
#pragma unroll
    for (int i = 0; i < N; i++) r.d[i] = a.d[i] * ib;
    return r;
}
template <int N>
__device__ inline Dual<N> operator/(double a, const Dual<N>& b)
{
    const double q = a / b.v;
    return chain(q, -q / b.v, b);
}

template <int N, class T>
__device__ inline Dual<N>& operator+=(Dual<N>& a, const T& b)
{
    a = a + b;
    return a;
}
template <int N, class T>
__device__ inline Dual<N>& operator-=(Dual<N>& a, const T& b)
{
    a = a - b;
    return a;
}
template <int N, class T>
__device__ inline Dual<N>& operator*=(Dual<N>& a, const T& b)
{
    a = a * b;
    return a;
}
template <int N, class T>
__device__ inline Dual<N>& operator/=(Dual<N>& a, const T& b)
{
    a = a / b;
    return a;
}

#define BBHX_DUAL_CMP(op)                                                    \
    template <int N>                                                         \
    __device__ inline bool operator op(const Dual<N>& a, const Dual<N>& b)   \
    {                                                                        \
        return a.v op b.v;                                                   \
    }                                                                        \
    template <int N>                                                         \
    __device__ inline bool operator op(const Dual<N>& a, double b)           \
    {                                                                        \
        return a.v op b;                                                     \
    }                                                                        \
    template <int N>                                                         \
    __device__ inline bool operator op(double a, const Dual<N>& b)           \
    {                                                                        \
        return a op b.v;                                                     \
    }
BBHX_DUAL_CMP(<)
BBHX_DUAL_CMP(>)
BBHX_DUAL_CMP(<=)
BBHX_DUAL_CMP(>=)
BBHX_DUAL_CMP(==)
BBHX_DUAL_CMP(!=)
#undef BBHX_DUAL_CMP

// elementary functions: value from the double routine, derivative by the chain rule
template <int N>
__device__ inline Dual<N> sqrt(const Dual<N>& x)
{
    const double s = ::sqrt(x.v);
    return chain(s, 0.5 / s, x);
}
template <int N>
__device__ inline Dual<N> cbrt(const Dual<N>& x)
{
    const double c = ::cbrt(x.v);
    return chain(c, c / (3.0 * x.v), x);
}
template <int N>
__device__ inline Dual<N> cbrt_pos(const Dual<N>& x)
{
    const double c = cbrt_pos(x.v);
    return chain(c, c * rcp_pos(3.0 * x.v), x);
}
template <int N>
__device__ inline Dual<N> root6_pos(const Dual<N>& x)
{
    const double c = root6_pos(x.v);
    return chain(c, c * rcp_pos(6.0 * x.v), x);
}

template <int N>
__device__ inline Dual<N> rcp_pos(const Dual<N>& x)
{
    const double r = rcp_pos(x.v);
    return chain(r, -r * r, x);
}
template <int N>
__device__ inline Dual<N> log(const Dual<N>& x)
{
    return chain(::log(x.v), rcp_pos(x.v), x);
}
template <int N>
__device__ inline Dual<N> log_pos(const Dual<N>& x)
{
    return chain(log_pos(x.v), rcp_pos(x.v), x);
}
template <int N>
__device__ inline Dual<N> exp_fast(const Dual<N>& x)
{
    const double e = exp_fast(x.v);
    return chain(e, e, x);
}
template <int N>
__device__ inline Dual<N> atan_fast(const Dual<N>& x)
{
    return chain(atan_fast(x.v), rcp_pos(1.0 + x.v * x.v), x);
}
template <int N>
__device__ inline Dual<N> exp(const Dual<N>& x)
{
    const double e = ::exp(x.v);
    return chain(e, e, x);
}
template <int N>
__device__ inline Dual<N> atan(const Dual<N>& x)
{
    return chain(::atan(x.v), 1.0 / (1.0 + x.v * x.v), x);
}
template <int N>
__device__ inline Dual<N> fabs(const Dual<N>& x)
{
    return x.v < 0.0 ? -x : x;
}
template <int N>
__device__ inline Dual<N> pow(const Dual<N>& x, double p)
{
    const double y = ::pow(x.v, p);
    return chain(y, p * y / x.v, x);
}
template <int N>
__device__ inline Dual<N> hypot(const Dual<N>& a, const Dual<N>& b)
{
    const double h = ::hypot(a.v, b.v);
    Dual<N> r;
    r.v = h;
    const double ih = h > 0.0 ? 1.0 / h : 0.0;
#pragma unroll
    for (int i = 0; i < N; i++) r.d[i] = (a.v * a.d[i] + b.v * b.d[i]) * ih;
    return r;
}
template <int N>
__device__ inline Dual<N> sin(const Dual<N>& x)
{
    double s, c;
    ::sincos(x.v, &s, &c);
    return chain(s, c, x);
}
template <int N>
__device__ inline Dual<N> cos(const Dual<N>& x)
{
    double s, c;
    ::sincos(x.v, &s, &c);
    return chain(c, -s, x);
}
template <int N>
__device__ inline void sincos_red(const Dual<N>& x, Dual<N>* s, Dual<N>* c)
{
    double sv, cv;
    sincos_red(x.v, &sv, &cv);
    *s = chain(sv, cv, x);
    *c = chain(cv, -sv, x);
}

}  // namespace bbhx
