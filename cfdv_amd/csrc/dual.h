// Forward-mode dual numbers for the element-build tangents (discrete
// adjoint).  Dual<N> carries a value and N directional derivatives; the
// element physics in fdv_device.h and build.hip is templated on its scalar
// type, so the same code yields R, rhs (S = double) or R, rhs and their
// derivatives along N seeded input directions (S = Dual<N>).  Branches
// follow the value, so derivatives are the one-sided ones of the branch
// taken, exactly as a tangent of the double code.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#ifndef CFDV_HD
#define CFDV_HD __host__ __device__ __forceinline__
#endif

namespace cfdv {

// the double overloads stay visible next to the Dual ones inside cfdv
using ::fabs;
using ::pow;
using ::sqrt;

template <int N>
struct Dual {
    double v;
    double d[N];
    CFDV_HD Dual() : v(0.0) {
#pragma unroll
        for (int i = 0; i < N; i++) d[i] = 0.0;
    }
    CFDV_HD Dual(double x) : v(x) {
#pragma unroll
        for (int i = 0; i < N; i++) d[i] = 0.0;
    }
    CFDV_HD Dual& operator+=(const Dual& o) {
        v += o.v;
#pragma unroll
        for (int i = 0; i < N; i++) d[i] += o.d[i];
        return *this;
    }
    CFDV_HD Dual& operator-=(const Dual& o) {
        v -= o.v;
#pragma unroll
        for (int i = 0; i < N; i++) d[i] -= o.d[i];
        return *this;
    }
    CFDV_HD Dual& operator*=(const Dual& o) {
#pragma unroll
        for (int i = 0; i < N; i++) d[i] = d[i] * o.v + v * o.d[i];
        v *= o.v;
        return *this;
    }
    CFDV_HD Dual& operator/=(const Dual& o) {
        const double inv = 1.0 / o.v;
        const double q = v * inv;
#pragma unroll
        for (int i = 0; i < N; i++) d[i] = (d[i] - q * o.d[i]) * inv;
        v = q;
        return *this;
    }
};

template <int N> CFDV_HD Dual<N> operator+(Dual<N> a, const Dual<N>& b) { return a += b; }
template <int N> CFDV_HD Dual<N> operator-(Dual<N> a, const Dual<N>& b) { return a -= b; }
template <int N> CFDV_HD Dual<N> operator*(Dual<N> a, const Dual<N>& b) { return a *= b; }
template <int N> CFDV_HD Dual<N> operator/(Dual<N> a, const Dual<N>& b) { return a /= b; }
template <int N> CFDV_HD Dual<N> operator+(Dual<N> a, double b) { a.v += b; return a; }
template <int N> CFDV_HD Dual<N> operator+(double b, Dual<N> a) { a.v += b; return a; }
template <int N> CFDV_HD Dual<N> operator-(Dual<N> a, double b) { a.v -= b; return a; }
template <int N> CFDV_HD Dual<N> operator-(double b, const Dual<N>& a) {
    Dual<N> r;
    r.v = b - a.v;
#pragma unroll
    for (int i = 0; i < N; i++) r.d[i] = -a.d[i];
    return r;
}
template <int N> CFDV_HD Dual<N> operator-(const Dual<N>& a) { return 0.0 - a; }
template <int N> CFDV_HD Dual<N> operator*(Dual<N> a, double b) {
    a.v *= b;
#pragma unroll
    for (int i = 0; i < N; i++) a.d[i] *= b;
    return a;
}
template <int N> CFDV_HD Dual<N> operator*(double b, Dual<N> a) { return a * b; }
template <int N> CFDV_HD Dual<N> operator/(Dual<N> a, double b) {
    a.v /= b;
#pragma unroll
    for (int i = 0; i < N; i++) a.d[i] /= b;
    return a;
}
template <int N> CFDV_HD Dual<N> operator/(double b, const Dual<N>& a) { return Dual<N>(b) / a; }
template <int N> CFDV_HD Dual<N>& operator+=(Dual<N>& a, double b) { a.v += b; return a; }
template <int N> CFDV_HD Dual<N>& operator-=(Dual<N>& a, double b) { a.v -= b; return a; }
template <int N> CFDV_HD Dual<N>& operator*=(Dual<N>& a, double b) { a = a * b; return a; }
template <int N> CFDV_HD Dual<N>& operator/=(Dual<N>& a, double b) { a = a / b; return a; }

template <int N> CFDV_HD bool operator<(const Dual<N>& a, double b) { return a.v < b; }
template <int N> CFDV_HD bool operator>(const Dual<N>& a, double b) { return a.v > b; }
template <int N> CFDV_HD bool operator<=(const Dual<N>& a, double b) { return a.v <= b; }
template <int N> CFDV_HD bool operator>=(const Dual<N>& a, double b) { return a.v >= b; }
template <int N> CFDV_HD bool operator==(const Dual<N>& a, double b) { return a.v == b; }
template <int N> CFDV_HD bool operator!=(const Dual<N>& a, double b) { return a.v != b; }
template <int N> CFDV_HD bool operator<(double a, const Dual<N>& b) { return a < b.v; }
template <int N> CFDV_HD bool operator>(double a, const Dual<N>& b) { return a > b.v; }
template <int N> CFDV_HD bool operator<(const Dual<N>& a, const Dual<N>& b) { return a.v < b.v; }
template <int N> CFDV_HD bool operator>(const Dual<N>& a, const Dual<N>& b) { return a.v > b.v; }

template <int N> CFDV_HD Dual<N> sqrt(const Dual<N>& a) {
    Dual<N> r;
    r.v = ::sqrt(a.v);
    const double g = (r.v > 0.0) ? 0.5 / r.v : 0.0;
#pragma unroll
    for (int i = 0; i < N; i++) r.d[i] = g * a.d[i];
    return r;
}
// x^p; at x = 0 the derivative is taken as 0 (the element code only reaches
// pow(0, p) on a branch whose argument is a constant)
template <int N> CFDV_HD Dual<N> pow(const Dual<N>& a, double p) {
    Dual<N> r;
    r.v = ::pow(a.v, p);
    const double g = (a.v != 0.0) ? p * ::pow(a.v, p - 1.0) : 0.0;
#pragma unroll
    for (int i = 0; i < N; i++) r.d[i] = g * a.d[i];
    return r;
}
template <int N> CFDV_HD Dual<N> fabs(const Dual<N>& a) { return (a.v < 0.0) ? -a : a; }

// value of a scalar of either kind
CFDV_HD double value_of(double x) { return x; }
template <int N> CFDV_HD double value_of(const Dual<N>& x) { return x.v; }

}  // namespace cfdv
