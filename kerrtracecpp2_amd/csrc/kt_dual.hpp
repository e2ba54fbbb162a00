// ============================================================================
//  kt_dual.hpp -- forward-mode dual numbers with N tangent directions.
//
//  Used by the adjoint to form exact local Jacobians of the per-ray maps (camera
//  ray, right-hand side, redshift and intensity) from the same templated source
//  the binary64 kernels run; val() reads the value of either a double or a dual
//  so branches in templated code test the same quantities in both.
// ============================================================================
#pragma once
#include "kt_device.hpp"

namespace kt {

#if defined(__HIPCC__)
// the dual overloads below would hide the scalar functions inside this namespace
using ::cbrt;
using ::fabs;
using ::fma;
using ::sqrt;
#endif

template <int N> struct DN {
    double v;
    double d[N];
    KT_HD DN() : v(0.0) { for (int i = 0; i < N; ++i) d[i] = 0.0; }
    KT_HD DN(double x) : v(x) { for (int i = 0; i < N; ++i) d[i] = 0.0; }
    // x with unit tangent in direction k (k < 0: constant)
    KT_HD static DN var(double x, int k) { DN r(x); if (k >= 0 && k < N) r.d[k] = 1.0; return r; }
};

KT_HD double val(double x) { return x; }
template <int N> KT_HD double val(const DN<N>& x) { return x.v; }

template <int N> KT_HD DN<N> operator+(const DN<N>& x, const DN<N>& y) {
    DN<N> r; r.v = x.v + y.v; for (int i = 0; i < N; ++i) r.d[i] = x.d[i] + y.d[i]; return r;
}
template <int N> KT_HD DN<N> operator+(const DN<N>& x, double y) { DN<N> r = x; r.v = x.v + y; return r; }
template <int N> KT_HD DN<N> operator+(double x, const DN<N>& y) { DN<N> r = y; r.v = x + y.v; return r; }
template <int N> KT_HD DN<N> operator-(const DN<N>& x, const DN<N>& y) {
    DN<N> r; r.v = x.v - y.v; for (int i = 0; i < N; ++i) r.d[i] = x.d[i] - y.d[i]; return r;
}
template <int N> KT_HD DN<N> operator-(const DN<N>& x, double y) { DN<N> r = x; r.v = x.v - y; return r; }
template <int N> KT_HD DN<N> operator-(double x, const DN<N>& y) {
    DN<N> r; r.v = x - y.v; for (int i = 0; i < N; ++i) r.d[i] = -y.d[i]; return r;
}
template <int N> KT_HD DN<N> operator-(const DN<N>& x) {
    DN<N> r; r.v = -x.v; for (int i = 0; i < N; ++i) r.d[i] = -x.d[i]; return r;
}
template <int N> KT_HD DN<N> operator*(const DN<N>& x, const DN<N>& y) {
    DN<N> r; r.v = x.v*y.v; for (int i = 0; i < N; ++i) r.d[i] = fma(x.d[i], y.v, x.v*y.d[i]); return r;
}
template <int N> KT_HD DN<N> operator*(const DN<N>& x, double y) {
    DN<N> r; r.v = x.v*y; for (int i = 0; i < N; ++i) r.d[i] = x.d[i]*y; return r;
}
template <int N> KT_HD DN<N> operator*(double x, const DN<N>& y) { return y*x; }
template <int N> KT_HD DN<N> kt_rcp(const DN<N>& x) {
    const double q = 1.0 / x.v, m = -q*q;
    DN<N> r; r.v = q; for (int i = 0; i < N; ++i) r.d[i] = x.d[i]*m; return r;
}
template <int N> KT_HD DN<N> operator/(const DN<N>& x, const DN<N>& y) {
    const double q = x.v / y.v, iy = 1.0 / y.v;
    DN<N> r; r.v = q; for (int i = 0; i < N; ++i) r.d[i] = (x.d[i] - q*y.d[i])*iy; return r;
}
template <int N> KT_HD DN<N> operator/(const DN<N>& x, double y) {
    DN<N> r; r.v = x.v / y; for (int i = 0; i < N; ++i) r.d[i] = x.d[i] / y; return r;
}
template <int N> KT_HD DN<N> operator/(double x, const DN<N>& y) { return DN<N>(x) / y; }
template <int N> KT_HD DN<N> fma(const DN<N>& x, const DN<N>& y, const DN<N>& z) {
    DN<N> r; r.v = fma(x.v, y.v, z.v);
    for (int i = 0; i < N; ++i) r.d[i] = fma(x.d[i], y.v, fma(x.v, y.d[i], z.d[i]));
    return r;
}
template <int N> KT_HD DN<N> fma(double x, const DN<N>& y, const DN<N>& z) { return fma(DN<N>(x), y, z); }
template <int N> KT_HD DN<N> fma(const DN<N>& x, double y, const DN<N>& z) { return fma(x, DN<N>(y), z); }
template <int N> KT_HD DN<N> fma(const DN<N>& x, const DN<N>& y, double z) { return fma(x, y, DN<N>(z)); }
template <int N> KT_HD DN<N> sqrt(const DN<N>& x) {
    const double s = sqrt(x.v), h = 0.5 / s;
    DN<N> r; r.v = s; for (int i = 0; i < N; ++i) r.d[i] = x.d[i]*h; return r;
}
template <int N> KT_HD DN<N> cbrt(const DN<N>& x) {
    const double c = cbrt(x.v), h = 1.0 / (3.0*c*c);
    DN<N> r; r.v = c; for (int i = 0; i < N; ++i) r.d[i] = x.d[i]*h; return r;
}
template <int N> KT_HD DN<N> fabs(const DN<N>& x) { return x.v < 0.0 ? -x : x; }
template <int N> KT_HD void kt_sincos(const DN<N>& x, DN<N>* s, DN<N>* c) {
    double sv, cv;
    kt_sincos(x.v, &sv, &cv);
    s->v = sv; c->v = cv;
    for (int i = 0; i < N; ++i) { s->d[i] = cv*x.d[i]; c->d[i] = -sv*x.d[i]; }
}
// select by value (the branch structure of the templated code is decided on values)
template <class T> KT_HD T vmin(const T& x, const T& y) { return val(y) < val(x) ? y : x; }
template <class T> KT_HD T vmax(const T& x, const T& y) { return val(y) > val(x) ? y : x; }

}  // namespace kt
