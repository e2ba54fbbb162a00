// Counts the floating-point operations of the product's analytic right-hand side
// (kerrtracecpp2_amd/csrc/kt_device.hpp, rhs<CHART,T>) by instantiating it with a
// counting number type.  add/sub/mul = 1, fma = 2, reciprocal/div/sqrt = 1 (and tallied
// as "special").  Prints "BL <flops> <special>" and "KS <flops> <special>".
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
namespace kt {
static inline Flop fma(Flop a, Flop b, Flop c) { g_flops += 2; return Flop(a.v * b.v + c.v); }
static inline Flop kt_rcp(Flop a) { ++g_flops; ++g_special; return Flop(1.0 / a.v); }
}
#include "../../kerrtracecpp2_amd/csrc/kt_device.hpp"

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
    return 0;
}
