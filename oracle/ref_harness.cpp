// Function-level known answers from the reference headers, compiled where they lie
// (make -C oracle ref).  Reads "name args..." lines on stdin, prints results with
// %.17g.  Test infrastructure only: used by tests/golden/make_golden.py in the
// build container to write fixtures; never shipped or run on the GPU box.
#include "camera.hpp"
#include "geodesic.hpp"
#include "knds_metric.hpp"
#include "ray_bundle.hpp"

#include <cstdio>
#include <cstring>
#include <string>

static void pr(double v) { std::printf(" %.17g", v); }

int main() {
    char cmd[64];
    while (std::scanf("%63s", cmd) == 1) {
        std::string c(cmd);
        double M, a, Q, L;
        if (std::scanf("%lf %lf %lf %lf", &M, &a, &Q, &L) != 4) return 1;
        KNdSMetric g(M, a, Q, L);
        std::printf("%s", cmd);
        if (c == "upper" || c == "lower") {
            double r, th, G[4][4];
            std::scanf("%lf %lf", &r, &th);
            if (c == "upper") g.contravariant_BL(r, th, G); else g.covariant_BL(r, th, G);
            for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) pr(G[i][j]);
        } else if (c == "upper_ks") {
            double X, Y, Z, G[4][4];
            std::scanf("%lf %lf %lf", &X, &Y, &Z);
            g.contravariant_KS(0.0, X, Y, Z, true, G);
            for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) pr(G[i][j]);
        } else if (c == "ham") {
            double s[6];
            for (double& v : s) std::scanf("%lf", &v);
            pr(g.hamiltonian(s[0], s[1], s[2], s[3], s[4], s[5]));
        } else if (c == "rhs") {
            double s[6], o[4];
            for (double& v : s) std::scanf("%lf", &v);
            geodesic_rhs(g, s[0], s[1], s[2], s[3], s[4], s[5], o[0], o[1], o[2], o[3]);
            for (double v : o) pr(v);
            pr(dphi_vel(g, s[0], s[1], s[4], s[5]));
        } else if (c == "horizon") {
            pr(g.r_horizon());
        } else if (c == "isco") {
            pr(g.r_isco());
        } else if (c == "omega") {
            double r; std::scanf("%lf", &r);
            pr(g.keplerian_omega(r));
        } else if (c == "camera") {
            double cam[6], al, be;
            for (double& v : cam) std::scanf("%lf", &v);
            std::scanf("%lf %lf", &al, &be);
            Camera C(cam[0], cam[1], cam[2], cam[3], (int)cam[4], (int)cam[5]);
            GeodesicState s = C.angle_ray(al, be, g);
            pr(s.r); pr(s.theta); pr(s.phi); pr(s.pr); pr(s.ptheta); pr(s.pt); pr(s.pphi);
        } else if (c == "step") {
            int method; double s[7], h, tol;
            std::scanf("%d", &method);
            for (double& v : s) std::scanf("%lf", &v);
            std::scanf("%lf %lf", &h, &tol);
            GeodesicState st{s[0], s[1], s[2], s[3], s[4], s[5], s[6]};
            Vec4d f = Vec4d::nan_init();
            const bool ok = adaptive_step(g, st, h, method ? Integrator::DOPRI5 : Integrator::RK4_DOUBLING, f, tol);
            pr(ok ? 1.0 : 0.0); pr(st.r); pr(st.theta); pr(st.phi); pr(st.pr); pr(st.ptheta); pr(h);
        } else if (c == "bl2ks") {
            double r, th, ph, X, Y, Z;
            std::scanf("%lf %lf %lf", &r, &th, &ph);
            KNdSMetric::BL_to_KS_spatial(r, th, ph, a, X, Y, Z);
            pr(X); pr(Y); pr(Z);
        } else if (c == "ks2bl") {
            double X, Y, Z, r, th, ph;
            std::scanf("%lf %lf %lf", &X, &Y, &Z);
            KNdSMetric::KS_to_BL_spatial(X, Y, Z, a, r, th, ph);
            pr(r); pr(th); pr(ph);
        } else if (c == "hermite_root") {
            double y0, y1, d0, d1, h, tgt, al = -1.0;
            std::scanf("%lf %lf %lf %lf %lf %lf", &y0, &y1, &d0, &d1, &h, &tgt);
            const bool ok = first_event_alpha_hermite(y0, y1, d0, d1, h, tgt, al, 8, 8);
            pr(ok ? 1.0 : 0.0); pr(al);
        } else {
            return 2;
        }
        std::printf("\n");
    }
    return 0;
}
