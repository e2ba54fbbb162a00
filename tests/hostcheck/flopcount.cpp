// Counts the floating-point operations of the product's analytic right-hand side
// (kerrtracecpp2_amd/csrc/kt_device.hpp, rhs<CHART,T>) and of the adjoint's reverse-mode
// vector-Jacobian product of it (kt_adjoint.hpp, rhs_vjp_sc<CHART,T>) by instantiating them
// with a counting number type.  add/sub/mul = 1, fma = 2, reciprocal/div/sqrt = 1 (and
// tallied as "special").  Prints "BL|KS <flops> <special>" and "VJP_BL|VJP_KS <flops> <special>".
#include <cstdio>

struct Flop;
static long g_flops = 0, g_special = 0;
struct Flop {
    double v;
    Flop(double x = 0.0) : v(x) {}
};
static inline Flop operator+(Flop a, Flop b) { ++g_flops; return Flop(a.v + b.v); }
static inline Flop operator-(Flop a, Flop b) { ++g_flops; return Flop(a.v - b.v); }
static inline Flop operator*(Flop a, Flop b) { ++g_flops; return Flop(a.v * b.v); }
static inline Flop operator/(Flop a, Flop b) { ++g_flops; ++g_special; return Flop(a.v / b.v); }
static inline Flop operator-(Flop a) { return Flop(-a.v); }
static inline Flop& operator+=(Flop& a, Flop b) { ++g_flops; a.v += b.v; return a; }
static inline Flop& operator-=(Flop& a, Flop b) { ++g_flops; a.v -= b.v; return a; }
namespace kt {
static inline Flop fma(Flop a, Flop b, Flop c) { g_flops += 2; return Flop(a.v * b.v + c.v); }
static inline Flop kt_rcp(Flop a) { ++g_flops; ++g_special; return Flop(1.0 / a.v); }
}
#include "../../kerrtracecpp2_amd/csrc/kt_adjoint.hpp"

// counting mirror of kt_bundle.hpp's dual number D2 (value + two tangents), with the same
// operation formulas, to count the bundle kernel's dual-number right-hand side
struct DF { Flop v, a, b; DF() {} DF(Flop v_, Flop a_, Flop b_) : v(v_), a(a_), b(b_) {} };
static DF operator+(const DF& x, const DF& y) { return DF(x.v + y.v, x.a + y.a, x.b + y.b); }
static DF operator+(const DF& x, Flop y) { return DF(x.v + y, x.a, x.b); }
static DF operator+(Flop x, const DF& y) { return DF(x + y.v, y.a, y.b); }
static DF operator-(const DF& x, const DF& y) { return DF(x.v - y.v, x.a - y.a, x.b - y.b); }
static DF operator-(const DF& x) { return DF(-x.v, -x.a, -x.b); }
static DF operator*(const DF& x, const DF& y) {
    return DF(x.v*y.v, kt::fma(x.a, y.v, x.v*y.a), kt::fma(x.b, y.v, x.v*y.b));
}
static DF operator*(const DF& x, Flop y) { return DF(x.v*y, x.a*y, x.b*y); }
static DF operator*(Flop x, const DF& y) { return DF(x*y.v, x*y.a, x*y.b); }
// found by argument-dependent lookup from the templates in namespace kt
using kt::fma;
static DF fma(const DF& x, const DF& y, const DF& z) {
    return DF(fma(x.v, y.v, z.v), fma(x.a, y.v, fma(x.v, y.a, z.a)), fma(x.b, y.v, fma(x.v, y.b, z.b)));
}
static DF fma(Flop x, const DF& y, const DF& z) { return DF(fma(x, y.v, z.v), fma(x, y.a, z.a), fma(x, y.b, z.b)); }
static DF fma(const DF& x, Flop y, const DF& z) { return DF(fma(x.v, y, z.v), fma(x.a, y, z.a), fma(x.b, y, z.b)); }
static DF fma(const DF& x, const DF& y, Flop z) {
    return DF(fma(x.v, y.v, z), fma(x.a, y.v, x.v*y.a), fma(x.b, y.v, x.v*y.b));
}
static DF fma(Flop x, const DF& y, Flop z) { return DF(fma(x, y.v, z), x*y.a, x*y.b); }
static DF fma(const DF& x, Flop y, Flop z) { return DF(fma(x.v, y, z), x.a*y, x.b*y); }
static DF kt_rcp(const DF& x) { const Flop r = kt::kt_rcp(x.v), m = -(r*r); return DF(r, x.a*m, x.b*m); }

int main() {
    kt::Knds<Flop> P;
    P.a = 0.9; P.a2 = 0.81; P.M = 1; P.M2 = 2; P.Q2 = 0.09; P.lam3 = 1e-5 / 3; P.lam3a2 = 0.81e-5 / 3;
    P.xi = 1.0; P.xi2 = 1.0;
    kt::SinCos<Flop> sc{Flop(0.99), Flop(0.1)};
    Flop out[5];
    g_flops = g_special = 0;
    kt::rhs<kt::CHART_BL>(P, Flop(10.0), sc, Flop(-0.5), Flop(0.3), Flop(-1.0), Flop(2.5), out);
    std::printf("BL %ld %ld\n", g_flops, g_special);
    g_flops = g_special = 0;
    kt::rhs<kt::CHART_KS>(P, Flop(10.0), sc, Flop(-0.5), Flop(0.3), Flop(-1.0), Flop(2.5), out);
    std::printf("KS %ld %ld\n", g_flops, g_special);
    const Flop y[4] = {Flop(10.0), Flop(1.4), Flop(-0.5), Flop(0.3)}, mu[4] = {Flop(0.1), Flop(0.2), Flop(0.3), Flop(0.4)};
    Flop g[9];
    g_flops = g_special = 0;
    kt::rhs_vjp_sc<kt::CHART_BL, Flop>(P, Flop(0.3), Flop(1e-5), y, sc, Flop(-1.0), Flop(2.5), mu, g);
    std::printf("VJP_BL %ld %ld\n", g_flops, g_special);
    g_flops = g_special = 0;
    kt::rhs_vjp_sc<kt::CHART_KS, Flop>(P, Flop(0.3), Flop(1e-5), y, sc, Flop(-1.0), Flop(2.5), mu, g);
    std::printf("VJP_KS %ld %ld\n", g_flops, g_special);
    kt::SinCos<DF> scd{DF(Flop(0.99), Flop(0.1), Flop(0.2)), DF(Flop(0.1), Flop(-0.3), Flop(0.4))};
    DF od[5];
    g_flops = g_special = 0;
    kt::rhs<kt::CHART_BL, DF, Flop>(P, DF(Flop(10.0), Flop(1), Flop(2)), scd, DF(Flop(-0.5), Flop(1), Flop(2)),
                                    DF(Flop(0.3), Flop(1), Flop(2)), Flop(-1.0), Flop(2.5), od);
    std::printf("DUAL_BL %ld %ld\n", g_flops, g_special);
    return 0;
}
