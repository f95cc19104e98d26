// Operation-counting stand-in for double, used to count the algorithmic
// floating-point work of the oracle's scalar code (TEST INFRASTRUCTURE ONLY).
//
// opcount.cpp includes the oracle sources with `double` defined to
// `opc::Real`; every arithmetic operator and libm call on a Real increments a
// counter and then does the FP64 operation itself, so the counted run also
// returns the oracle's ordinary results. Counting rules:
//   add, sub, mul, div, sqrt ... 1 flop each (negation and comparisons: 0)
//   each libm call (sin, cos, exp, log, pow, cbrt, atan, atan2, hypot, ...)
//   ... tallied per function; the flop total adds a fixed weight for each
//   (1 by default, see weighted_total in opcount.cpp).
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace opc {

enum Op { ADD, MUL, DIV, SQRT, SIN, COS, EXP, LOG, POW, CBRT, ATAN, ATAN2, HYPOT, FLOOR, FABS, NOPS };

struct Counts {
    uint64_t n[NOPS];
};
extern Counts g;

inline void tick(Op o) { g.n[o]++; }

struct Real {
    double v;
    constexpr Real() : v(0.0) {}
    constexpr Real(double x) : v(x) {}
    constexpr Real(int x) : v(x) {}
    constexpr Real(long x) : v(double(x)) {}
    constexpr Real(unsigned x) : v(x) {}
    constexpr Real(unsigned long x) : v(double(x)) {}
    explicit constexpr operator double() const { return v; }
    explicit constexpr operator int() const { return int(v); }
    explicit constexpr operator long() const { return long(v); }
    explicit constexpr operator bool() const { return v != 0.0; }
    Real& operator+=(Real b) { tick(ADD); v += b.v; return *this; }
    Real& operator-=(Real b) { tick(ADD); v -= b.v; return *this; }
    Real& operator*=(Real b) { tick(MUL); v *= b.v; return *this; }
    Real& operator/=(Real b) { tick(DIV); v /= b.v; return *this; }
};
static_assert(sizeof(Real) == sizeof(double), "Real must alias double arrays");

inline Real operator+(Real a, Real b) { tick(ADD); return Real(a.v + b.v); }
inline Real operator-(Real a, Real b) { tick(ADD); return Real(a.v - b.v); }
inline Real operator*(Real a, Real b) { tick(MUL); return Real(a.v * b.v); }
inline Real operator/(Real a, Real b) { tick(DIV); return Real(a.v / b.v); }
inline Real operator-(Real a) { return Real(-a.v); }
inline Real operator+(Real a) { return a; }
inline bool operator<(Real a, Real b) { return a.v < b.v; }
inline bool operator>(Real a, Real b) { return a.v > b.v; }
inline bool operator<=(Real a, Real b) { return a.v <= b.v; }
inline bool operator>=(Real a, Real b) { return a.v >= b.v; }
inline bool operator==(Real a, Real b) { return a.v == b.v; }
inline bool operator!=(Real a, Real b) { return a.v != b.v; }
inline bool operator!(Real a) { return a.v == 0.0; }

#define OPC_UNARY(name, op)                      \
    inline Real name(Real a)                     \
    {                                            \
        tick(op);                                \
        return Real(std::name(a.v));             \
    }
OPC_UNARY(sqrt, SQRT)
OPC_UNARY(sin, SIN)
OPC_UNARY(cos, COS)
OPC_UNARY(exp, EXP)
OPC_UNARY(log, LOG)
OPC_UNARY(cbrt, CBRT)
OPC_UNARY(atan, ATAN)
OPC_UNARY(floor, FLOOR)
OPC_UNARY(fabs, FABS)
#undef OPC_UNARY
inline Real abs(Real a) { tick(FABS); return Real(std::fabs(a.v)); }
inline Real pow(Real a, Real b) { tick(POW); return Real(std::pow(a.v, b.v)); }
inline Real pow(Real a, double b) { tick(POW); return Real(std::pow(a.v, b)); }
inline Real pow(Real a, int b) { tick(POW); return Real(std::pow(a.v, b)); }
inline Real atan2(Real a, Real b) { tick(ATAN2); return Real(std::atan2(a.v, b.v)); }
inline Real hypot(Real a, Real b) { tick(HYPOT); return Real(std::hypot(a.v, b.v)); }
inline Real log10(Real a) { tick(LOG); return Real(std::log10(a.v)); }
inline bool isnan(Real a) { return std::isnan(a.v); }
inline bool isinf(Real a) { return std::isinf(a.v); }
inline bool isfinite(Real a) { return std::isfinite(a.v); }
inline Real copysign(Real a, Real b) { return Real(std::copysign(a.v, b.v)); }
inline Real fmax(Real a, Real b) { return Real(std::fmax(a.v, b.v)); }
inline Real fmin(Real a, Real b) { return Real(std::fmin(a.v, b.v)); }

}  // namespace opc

// Complex numbers of Reals. The oracle's FP64 build uses std::complex<double>
// (GCC: products 4 mul + 2 add, quotients by Smith's method, abs by hypot,
// exp as exp(re) (cos im, sin im)); this specialisation counts those same
// operations, and like complex<double> it converts implicitly from a real.
namespace std {
template <>
class complex<opc::Real> {
public:
    typedef opc::Real value_type;
    constexpr complex(opc::Real re = opc::Real(), opc::Real im = opc::Real()) : re_(re), im_(im) {}
    constexpr complex(double re) : re_(re), im_(0.0) {}
    constexpr complex(int re) : re_(re), im_(0.0) {}
    constexpr opc::Real real() const { return re_; }
    constexpr opc::Real imag() const { return im_; }
    void real(opc::Real v) { re_ = v; }
    void imag(opc::Real v) { im_ = v; }
    complex& operator+=(const complex& b) { re_ += b.re_; im_ += b.im_; return *this; }
    complex& operator-=(const complex& b) { re_ -= b.re_; im_ -= b.im_; return *this; }
    complex& operator*=(const complex& b)
    {
        const opc::Real r = re_ * b.re_ - im_ * b.im_;
        im_ = re_ * b.im_ + im_ * b.re_;
        re_ = r;
        return *this;
    }
    complex& operator/=(const complex& b)
    {
        // Smith's algorithm
        using opc::Real;
        if (std::fabs(b.re_.v) >= std::fabs(b.im_.v)) {
            const Real r = b.im_ / b.re_, den = b.re_ + b.im_ * r;
            const Real x = (re_ + im_ * r) / den, y = (im_ - re_ * r) / den;
            re_ = x;
            im_ = y;
        } else {
            const Real r = b.re_ / b.im_, den = b.re_ * r + b.im_;
            const Real x = (re_ * r + im_) / den, y = (im_ * r - re_) / den;
            re_ = x;
            im_ = y;
        }
        return *this;
    }
    complex& operator*=(opc::Real s) { re_ *= s; im_ *= s; return *this; }
    complex& operator/=(opc::Real s) { re_ /= s; im_ /= s; return *this; }
    complex& operator+=(opc::Real s) { re_ += s; return *this; }
    complex& operator-=(opc::Real s) { re_ -= s; return *this; }

private:
    opc::Real re_, im_;
};
typedef complex<opc::Real> opc_cplx;
inline opc_cplx operator+(opc_cplx a, const opc_cplx& b) { return a += b; }
inline opc_cplx operator-(opc_cplx a, const opc_cplx& b) { return a -= b; }
inline opc_cplx operator*(opc_cplx a, const opc_cplx& b) { return a *= b; }
inline opc_cplx operator/(opc_cplx a, const opc_cplx& b) { return a /= b; }
inline opc_cplx operator*(opc_cplx a, opc::Real s) { return a *= s; }
inline opc_cplx operator*(opc::Real s, opc_cplx a) { return a *= s; }
inline opc_cplx operator*(opc_cplx a, double s) { return a *= opc::Real(s); }
inline opc_cplx operator*(double s, opc_cplx a) { return a *= opc::Real(s); }
inline opc_cplx operator/(opc_cplx a, opc::Real s) { return a /= s; }
inline opc_cplx operator/(opc_cplx a, double s) { return a /= opc::Real(s); }
inline opc_cplx operator+(opc_cplx a, opc::Real s) { return a += s; }
inline opc_cplx operator+(opc::Real s, opc_cplx a) { return a += s; }
inline opc_cplx operator+(opc_cplx a, double s) { return a += opc::Real(s); }
inline opc_cplx operator+(double s, opc_cplx a) { return a += opc::Real(s); }
inline opc_cplx operator-(opc_cplx a, opc::Real s) { return a -= s; }
inline opc_cplx operator-(opc_cplx a, double s) { return a -= opc::Real(s); }
inline opc_cplx operator-(opc::Real s, const opc_cplx& a) { return opc_cplx(s - a.real(), -a.imag()); }
inline opc_cplx operator-(double s, const opc_cplx& a) { return opc_cplx(opc::Real(s) - a.real(), -a.imag()); }
inline opc_cplx operator/(double s, const opc_cplx& a) { return opc_cplx(opc::Real(s)) / a; }
inline opc_cplx operator/(opc::Real s, const opc_cplx& a) { return opc_cplx(s) / a; }
inline opc_cplx operator-(const opc_cplx& a) { return opc_cplx(-a.real(), -a.imag()); }
inline bool operator==(const opc_cplx& a, const opc_cplx& b) { return a.real() == b.real() && a.imag() == b.imag(); }
inline bool operator!=(const opc_cplx& a, const opc_cplx& b) { return !(a == b); }
inline opc_cplx conj(const opc_cplx& a) { return opc_cplx(a.real(), -a.imag()); }
inline opc::Real real(const opc_cplx& a) { return a.real(); }
inline opc::Real imag(const opc_cplx& a) { return a.imag(); }
inline opc::Real norm(const opc_cplx& a) { return a.real() * a.real() + a.imag() * a.imag(); }
inline opc::Real abs(const opc_cplx& a) { return opc::hypot(a.real(), a.imag()); }
inline opc::Real arg(const opc_cplx& a) { return opc::atan2(a.imag(), a.real()); }
inline opc_cplx polar(opc::Real rho, opc::Real theta) { return opc_cplx(rho * opc::cos(theta), rho * opc::sin(theta)); }
inline opc_cplx exp(const opc_cplx& a) { return polar(opc::exp(a.real()), a.imag()); }
inline opc_cplx sqrt(const opc_cplx& a)
{
    // principal root, (|a| + |re|)/2 form
    using opc::Real;
    if (a.real() == Real(0.0) && a.imag() == Real(0.0)) return opc_cplx(0.0);
    const Real t = opc::sqrt((opc::fabs(a.real()) + abs(a)) * Real(0.5));
    if (a.real() >= Real(0.0)) return opc_cplx(t, a.imag() / (Real(2.0) * t));
    return opc_cplx(opc::fabs(a.imag()) / (Real(2.0) * t), std::signbit(a.imag().v) ? -t : t);
}
inline opc_cplx pow(const opc_cplx& a, double e)
{
    // through exp(e log a), as the generic complex power
    const opc::Real r = abs(a), th = arg(a);
    return polar(opc::exp(opc::Real(e) * opc::log(r)), opc::Real(e) * th);
}
}  // namespace std

// the oracle calls std::sqrt etc. explicitly
namespace std {
using opc::abs;
using opc::atan;
using opc::atan2;
using opc::cbrt;
using opc::cos;
using opc::exp;
using opc::fabs;
using opc::floor;
using opc::hypot;
using opc::isfinite;
using opc::isinf;
using opc::isnan;
using opc::log;
using opc::log10;
using opc::pow;
using opc::sin;
using opc::sqrt;
}  // namespace std
