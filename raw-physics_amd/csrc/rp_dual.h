// rp_dual.h — forward-mode dual numbers for the adjoint's local Jacobians.
//
// Dual<N> carries a value and N directional derivatives. Arithmetic follows
// the chain rule; the value part is computed with exactly the double
// operations of the primal expression, so a dual evaluation reproduces the
// primal results bit for bit. Comparisons look at values only (branches are
// those of the primal run). The functions live in rp::ad so that argument-
// dependent lookup finds them for duals while doubles keep the global ones.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#ifndef RP_HD
#define RP_HD __host__ __device__ __forceinline__
#endif

namespace rp {
namespace ad {

template <int N>
struct Dual {
    double v;
    double d[N];
    RP_HD Dual() : v(0.0) {
        for (int k = 0; k < N; ++k) d[k] = 0.0;
    }
    RP_HD Dual(double x) : v(x) {
        for (int k = 0; k < N; ++k) d[k] = 0.0;
    }
    explicit RP_HD operator double() const { return v; }
};

template <int N> RP_HD Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v + b.v);
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}
template <int N> RP_HD Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v - b.v);
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
}
template <int N> RP_HD Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v * b.v);
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
template <int N> RP_HD Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v / b.v);
    for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) / b.v;
    return r;
}
template <int N> RP_HD Dual<N> operator-(const Dual<N>& a) {
    Dual<N> r(-a.v);
    for (int k = 0; k < N; ++k) r.d[k] = -a.d[k];
    return r;
}
// mixed with double (the double is a constant)
template <int N> RP_HD Dual<N> operator+(const Dual<N>& a, double b) { return a + Dual<N>(b); }
template <int N> RP_HD Dual<N> operator+(double a, const Dual<N>& b) { return Dual<N>(a) + b; }
template <int N> RP_HD Dual<N> operator-(const Dual<N>& a, double b) { return a - Dual<N>(b); }
template <int N> RP_HD Dual<N> operator-(double a, const Dual<N>& b) { return Dual<N>(a) - b; }
template <int N> RP_HD Dual<N> operator*(const Dual<N>& a, double b) {
    Dual<N> r(a.v * b);
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b;
    return r;
}
template <int N> RP_HD Dual<N> operator*(double a, const Dual<N>& b) {
    Dual<N> r(a * b.v);
    for (int k = 0; k < N; ++k) r.d[k] = a * b.d[k];
    return r;
}
template <int N> RP_HD Dual<N> operator/(const Dual<N>& a, double b) {
    Dual<N> r(a.v / b);
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] / b;
    return r;
}
template <int N> RP_HD Dual<N> operator/(double a, const Dual<N>& b) { return Dual<N>(a) / b; }
template <int N> RP_HD Dual<N>& operator+=(Dual<N>& a, const Dual<N>& b) { return a = a + b; }
template <int N> RP_HD Dual<N>& operator-=(Dual<N>& a, const Dual<N>& b) { return a = a - b; }

#define RP_DUAL_CMP(op)                                                                         \
    template <int N> RP_HD bool operator op(const Dual<N>& a, const Dual<N>& b) { return a.v op b.v; } \
    template <int N> RP_HD bool operator op(const Dual<N>& a, double b) { return a.v op b; }           \
    template <int N> RP_HD bool operator op(double a, const Dual<N>& b) { return a op b.v; }
RP_DUAL_CMP(<)
RP_DUAL_CMP(>)
RP_DUAL_CMP(<=)
RP_DUAL_CMP(>=)
RP_DUAL_CMP(==)
RP_DUAL_CMP(!=)
#undef RP_DUAL_CMP

template <int N> RP_HD Dual<N> sqrt(const Dual<N>& a) {
    Dual<N> r(::sqrt(a.v));
    double g = (r.v > 0.0) ? 0.5 / r.v : 0.0;
    for (int k = 0; k < N; ++k) r.d[k] = g * a.d[k];
    return r;
}
template <int N> RP_HD Dual<N> fabs(const Dual<N>& a) {
    Dual<N> r(::fabs(a.v));
    double g = (a.v < 0.0) ? -1.0 : 1.0;
    for (int k = 0; k < N; ++k) r.d[k] = g * a.d[k];
    return r;
}

}  // namespace ad
}  // namespace rp
