// ============================================================================
//  kt_shade.hpp -- Phase 2 shading: GeoPixel -> RGB8 (blackbody disk palette,
//  Page-Thorne emissivity, Doppler boost, Reinhard tone map, equirect sky).
//
//  Reference anchors (KerrTraceCpp2 tree):
//    blackbody ........ main.cpp:1663-1674
//    page_thorne ...... main.cpp:1676-1680
//    tonemap .......... main.cpp:1683-1687
//    disk_colour ...... main.cpp:1799-1814
//    Interstellar ..... main.cpp:1689-1797 (hsv_to_rgb, cell_hash, disk_colour_interstellar)
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

struct ShadeParams {
    double M, r_isco, exposure, gamma, temp_scale;
    int palette;                 // 0 blackbody, 1 Interstellar cells
    int rings, sectors;          // Interstellar cell grid
    double sigma, hue_offset;    // Gaussian blend width (cells), hue offset
    double r_in, r_out;          // disk extent (cell rows span r_in..r_out)
};

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

// HSV -> RGB8 with round-to-nearest (main.cpp:1689-1709)
KT_HD Rgb8 hsv_to_rgb(double h, double s, double v) {
    h = h - floor(h);
    const double hh = h * 6.0;
    const int hi = (int)hh;
    const double f = hh - hi;
    const double p = v * (1.0 - s), q = v * (1.0 - s * f), t = v * (1.0 - s * (1.0 - f));
    double rr, gg, bb;
    switch (hi % 6) {
        case 0: rr = v; gg = t; bb = p; break;
        case 1: rr = q; gg = v; bb = p; break;
        case 2: rr = p; gg = v; bb = t; break;
        case 3: rr = p; gg = q; bb = v; break;
        case 4: rr = t; gg = p; bb = v; break;
        default: rr = v; gg = p; bb = q; break;
    }
    Rgb8 o;
    o.r = (uint8_t)(rr*255 + 0.5); o.g = (uint8_t)(gg*255 + 0.5); o.b = (uint8_t)(bb*255 + 0.5);
    return o;
}

// per-cell value in [0,1): fractional part of 43758.5453 sin(id a + b) (main.cpp:1711-1715)
KT_HD double cell_hash(int cell_id, double salt_a, double salt_b) {
    double sn, cs;
    kt_sincos((double)cell_id * salt_a + salt_b, &sn, &cs);
    const double v = sn * 43758.5453;
    return v - floor(v);
}

// Interstellar palette (main.cpp:1717-1797): rings x sectors warm cells, ~30 % transparent,
// Gaussian-blended across cell edges, modulated by the same emissivity and redshift as the
// blackbody palette.
KT_HD Rgb8 disk_colour_interstellar(double r, double phi, double red, double magnif, const ShadeParams& p) {
    const int NR = p.rings > 0 ? p.rings : 1;
    const int NS = p.sectors > 0 ? p.sectors : 1;
    const double sigma = p.sigma > 0 ? p.sigma : 0.5;
    const double r_norm = clampd((r - p.r_in) / (p.r_out - p.r_in), 0.0, 1.0);
    const double phi_norm = (phi - floor(phi / (2.0*kPi)) * (2.0*kPi)) / (2.0*kPi);
    const double ring_pos = r_norm * NR, sec_pos = phi_norm * NS;
    double acc_r = 0.0, acc_g = 0.0, acc_b = 0.0, w_total = 0.0;
    const double inv2sig2 = 1.0 / (2.0 * sigma * sigma);
    const int search = (int)ceil(3.0 * sigma) + 1;
    for (int ri = -search; ri <= NR + search; ++ri) {
        const double dr = ring_pos - (double)ri;
        if (dr * dr * inv2sig2 > 16.0) continue;
        const int ric = ri < 0 ? 0 : (ri >= NR ? NR - 1 : ri);
        for (int si = -search; si <= NS + search; ++si) {
            double ds = sec_pos - (double)si;
            ds = ds - floor((ds + NS * 0.5) / NS) * NS;
            const double w = exp(-(dr * dr + ds * ds) * inv2sig2);
            if (w < 1e-4) continue;
            const int cell = ric * NS + ((si % NS) + NS) % NS;
            const double h1 = cell_hash(cell, 127.1, 311.7);
            const double h2 = cell_hash(cell, 269.5, 183.3);
            const double h3 = cell_hash(cell, 419.2, 371.9);
            if (h3 < 0.30) continue;
            const Rgb8 c = hsv_to_rgb(fmod(h1 * 0.14 + p.hue_offset, 1.0), 0.70 + h2 * 0.30, 0.50 + h1 * 0.50);
            acc_r += w * c.r; acc_g += w * c.g; acc_b += w * c.b;
            w_total += w;
        }
    }
    Rgb8 o = {0, 0, 0};
    if (w_total < 1e-6) return o;
    const double pn = page_thorne(r, p.r_isco);
    const double pk = page_thorne(3.0*p.r_isco, p.r_isco);
    double I = (pk > 0.0) ? pn/pk : 0.0;
    const double rc = clampd(red, 0.1, 10.0);
    const double lift = 1.0 + 0.85 * clampd(1.0 - rc, 0.0, 1.0);
    I *= pow(rc, 4.0) * lift;
    I *= clampd(1.0/magnif, 0.05, 5.0);
    const double norm = w_total * 255.0;
    o.r = (uint8_t)(tonemap(acc_r / norm * I, p.exposure, p.gamma)*255);
    o.g = (uint8_t)(tonemap(acc_g / norm * I, p.exposure, p.gamma)*255);
    o.b = (uint8_t)(tonemap(acc_b / norm * I, p.exposure, p.gamma)*255);
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
KT_HD uint32_t shade_pixel(uint8_t outcome, float r, float red, float magnif, float ph_disk, float th_esc,
                           float ph_esc, const ShadeParams& p, const uint8_t* bg, int bw, int bh) {
    Rgb8 c = {0, 0, 0};
    if (outcome == 1) c = p.palette == 1 ? disk_colour_interstellar(r, ph_disk, red, magnif, p) : disk_colour(r, red, magnif, p);
    else if (outcome == 0 || outcome == 3) c = sky_sample(bg, bw, bh, th_esc, ph_esc);
    return (uint32_t)c.r | ((uint32_t)c.g << 8) | ((uint32_t)c.b << 16) | 0xFF000000u;
}

}  // namespace kt
