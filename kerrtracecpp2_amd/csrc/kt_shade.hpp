// ============================================================================
//  kt_shade.hpp -- Phase 2 shading: GeoPixel -> RGB8 (blackbody disk palette,
//  Page-Thorne emissivity, Doppler boost, Reinhard tone map, equirect sky).
//
//  Reference anchors (KerrTraceCpp2 tree):
//    blackbody ........ main.cpp:1663-1674
//    page_thorne ...... main.cpp:1676-1680
//    tonemap .......... main.cpp:1683-1687
//    disk_colour ...... main.cpp:1799-1814
//    colorize_buffer .. main.cpp:1817-1866 (outcome 1 disk, 0 sky, 2 black)
//    sky sample ....... main.cpp:150-170 (bilinear, u wraps, v clamps)
// ============================================================================
#pragma once
#include "kt_device.hpp"

namespace kt {

struct Rgb8 { uint8_t r, g, b; };

KT_HD double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

KT_HD Rgb8 blackbody(double T) {
    T = clampd(T, 800.0, 4e4);
    const double t = log10(T/800.0) / log10(4e4/800.0);
    double R, G, B;
    if (t < 0.25)      { R = 1; G = t/0.25*0.4; B = 0; }
    else if (t < 0.5)  { const double f = (t-0.25)/0.25; R = 1; G = 0.4+f*0.4; B = f*0.3; }
    else if (t < 0.75) { const double f = (t-0.5)/0.25;  R = 1; G = 0.8+f*0.2; B = 0.3+f*0.5; }
    else               { const double f = (t-0.75)/0.25; R = 1-f*0.2; G = 1; B = 0.8+f*0.2; }
    Rgb8 c;
    c.r = (uint8_t)(clampd(R, 0.0, 1.0)*255);
    c.g = (uint8_t)(clampd(G, 0.0, 1.0)*255);
    c.b = (uint8_t)(clampd(B, 0.0, 1.0)*255);
    return c;
}

KT_HD double page_thorne(double r, double risco) {
    if (r <= risco) return 0.0;
    const double x = sqrt(risco/r);
    return (1.0-x)/(r*r*r);
}

KT_HD double tonemap(double x, double exposure, double gamma) {
    x = x * exposure;
    x = x / (1.0 + x);
    return pow(clampd(x, 0.0, 1.0), 1.0/gamma);
}

struct ShadeParams { double M, r_isco, exposure, gamma, temp_scale; };

KT_HD Rgb8 disk_colour(double r, double red, double magnif, const ShadeParams& p) {
    const double T = 6500.0 * p.temp_scale * sqrt(6.0*p.M/r) * clampd(red, 0.2, 5.0);
    const double pn = page_thorne(r, p.r_isco);
    const double pk = page_thorne(3.0*p.r_isco, p.r_isco);
    double I = (pk > 0.0) ? pn/pk : 0.0;
    const double rc = clampd(red, 0.1, 10.0);
    const double lift = 1.0 + 0.85 * clampd(1.0 - rc, 0.0, 1.0);
    I *= pow(rc, 4.0) * lift;
    I *= clampd(1.0/magnif, 0.05, 5.0);
    const Rgb8 c = blackbody(T);
    Rgb8 o;
    o.r = (uint8_t)(tonemap(c.r/255.0*I, p.exposure, p.gamma)*255);
    o.g = (uint8_t)(tonemap(c.g/255.0*I, p.exposure, p.gamma)*255);
    o.b = (uint8_t)(tonemap(c.b/255.0*I, p.exposure, p.gamma)*255);
    return o;
}

KT_HD Rgb8 sky_sample(const uint8_t* px, int w, int h, double theta, double phi) {
    Rgb8 o = {0, 0, 0};
    if (!px || w <= 0 || h <= 0) return o;
    double pw = fmod(phi, 2.0*kPi);
    if (pw < 0) pw += 2.0*kPi;
    const double uf = pw / (2.0*kPi) * (w-1);
    const double vf = theta / kPi * (h-1);
    const int u0 = (int)uf, u1 = (w > 1) ? ((u0 + 1) % w) : 0;
    const int v0 = (int)vf, v1 = (v0 + 1 < h - 1) ? v0 + 1 : h - 1;
    const double fu = uf - u0, fv = vf - v0;
    auto at = [&](int u, int v, int ch) -> double {
        const int vv = v < 0 ? 0 : (v > h-1 ? h-1 : v);
        const int uu = u < 0 ? 0 : (u > w-1 ? w-1 : u);
        return (double)px[3*(vv*w + uu) + ch];
    };
    auto lp = [](double a, double b, double t) { return a + t*(b-a); };
    double c[3];
    for (int ch = 0; ch < 3; ++ch)
        c[ch] = lp(lp(at(u0, v0, ch), at(u1, v0, ch), fu), lp(at(u0, v1, ch), at(u1, v1, ch), fu), fv);
    o.r = (uint8_t)c[0]; o.g = (uint8_t)c[1]; o.b = (uint8_t)c[2];
    return o;
}

// outcome 1 disk, 0 (and 3) sky, anything else black; packed ABGR with alpha 0xFF
KT_HD uint32_t shade_pixel(uint8_t outcome, float r, float red, float magnif, float th_esc, float ph_esc,
                           const ShadeParams& p, const uint8_t* bg, int bw, int bh) {
    Rgb8 c = {0, 0, 0};
    if (outcome == 1) c = disk_colour(r, red, magnif, p);
    else if (outcome == 0 || outcome == 3) c = sky_sample(bg, bw, bh, th_esc, ph_esc);
    return (uint32_t)c.r | ((uint32_t)c.g << 8) | ((uint32_t)c.b << 16) | 0xFF000000u;
}

}  // namespace kt
