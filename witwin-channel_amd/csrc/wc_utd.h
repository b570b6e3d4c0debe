// First-order uniform theory of diffraction (Kouyoumjian & Pathak, Proc. IEEE
// 62(11), 1974) for a straight wedge edge with exterior angle n*pi, with
// Luebbers' (IEEE TAP 32(1), 1984) face reflection coefficients for lossy
// faces and the finite-edge Fresnel truncation factor
//   T = (0.5 + 0.5j) * conj(Fr(sigma*(t_max - t*)) - Fr(sigma*(t_min - t*))),
//   sigma = sqrt(k*kappa/pi), kappa = rho'^2/R'^3 + rho^2/R^3
// (reference docs/dev/audit/utd-continuity-fix-design.md:91-140).
// The transition function is evaluated in double precision: near a shadow
// boundary its argument kL*a goes to zero and the float cot*F product loses
// the digits that carry the geometrical-optics compensation.
#pragma once

#include "wc_em.h"

namespace wc {

// Fresnel integrals C(x), S(x) (definition C + jS = int_0^x e^{j pi t^2 / 2} dt)
// by the power series for |x| <= 1.5 and the continued fraction of the
// complementary error function (modified Lentz) above
// (Numerical Recipes, 2nd ed., section 6.9).
// For x > 1.5 the continued-fraction value h is also returned so callers can
// form the complementary integral without cancellation.
WC_HD void fresnel_cs(double x, double& c, double& s, bool& use_cf, double& h_re, double& h_im) {
    const double ax = fabs(x);
    use_cf = false;
    h_re = 0.0;
    h_im = 0.0;
    if (ax < 1.0e-150) {
        c = x;
        s = 0.0;
        return;
    }
    if (ax <= 1.5) {
        double sum = 0.0, sums = 0.0, sumc = ax;
        double sign = 1.0;
        const double fact = 0.5 * kPi * ax * ax;
        bool odd = true;
        double term = ax;
        double n = 3.0;
        for (int k = 1; k <= 200; ++k) {
            term *= fact / k;
            sum += sign * term / n;
            const double test = fabs(sum) * 1.0e-16;
            if (odd) {
                sign = -sign;
                sums = sum;
                sum = sumc;
            } else {
                sumc = sum;
                sum = sums;
            }
            if (term < test) break;
            odd = !odd;
            n += 2.0;
        }
        c = sumc;
        s = sums;
    } else {
        const double pix2 = kPi * ax * ax;
        // complex continued fraction, b = 1 - j*pix2
        double b_re = 1.0, b_im = -pix2;
        double cc_re = 1.0e300, cc_im = 0.0;
        // d = h = 1 / b
        double den = b_re * b_re + b_im * b_im;
        double d_re = b_re / den, d_im = -b_im / den;
        double hr = d_re, hi = d_im;
        double n = -1.0;
        for (int k = 2; k <= 400; ++k) {
            n += 2.0;
            const double a = -n * (n + 1.0);
            b_re += 4.0;
            // d = 1 / (a*d + b)
            double t_re = a * d_re + b_re, t_im = a * d_im + b_im;
            den = t_re * t_re + t_im * t_im;
            d_re = t_re / den;
            d_im = -t_im / den;
            // cc = b + a / cc
            den = cc_re * cc_re + cc_im * cc_im;
            const double q_re = a * cc_re / den, q_im = -a * cc_im / den;
            cc_re = b_re + q_re;
            cc_im = b_im + q_im;
            // del = cc * d
            const double del_re = cc_re * d_re - cc_im * d_im;
            const double del_im = cc_re * d_im + cc_im * d_re;
            const double nh_re = hr * del_re - hi * del_im;
            const double nh_im = hr * del_im + hi * del_re;
            hr = nh_re;
            hi = nh_im;
            if (fabs(del_re - 1.0) + fabs(del_im) < 1.0e-15) break;
        }
        // h *= (ax - j ax)
        const double h2_re = ax * hr + ax * hi;
        const double h2_im = ax * hi - ax * hr;
        // cs = (0.5 + 0.5j) * (1 - e^{j pix2/2} h)
        const double e_re = cos(0.5 * pix2), e_im = sin(0.5 * pix2);
        const double eh_re = e_re * h2_re - e_im * h2_im;
        const double eh_im = e_re * h2_im + e_im * h2_re;
        const double one_re = 1.0 - eh_re, one_im = -eh_im;
        c = 0.5 * one_re - 0.5 * one_im;
        s = 0.5 * one_re + 0.5 * one_im;
        use_cf = true;
        h_re = h2_re;
        h_im = h2_im;
    }
    if (x < 0.0) {
        c = -c;
        s = -s;
    }
}

// Kouyoumjian-Pathak transition function
//   F(X) = 2j sqrt(X) e^{jX} int_{sqrt X}^inf e^{-j tau^2} dtau,  X >= 0,
// with the substitution tau = u sqrt(pi/2):
//   int_{sqrt X}^inf e^{-j tau^2} dtau = sqrt(pi/2) * conj(int_{u0}^inf e^{j pi u^2/2} du),
//   u0 = sqrt(2X/pi).
WC_HD void utd_transition(double X, double& f_re, double& f_im) {
    if (X <= 0.0) {
        f_re = 0.0;
        f_im = 0.0;
        return;
    }
    const double u0 = sqrt(2.0 * X / kPi);
    double c, s, h_re, h_im;
    bool use_cf;
    fresnel_cs(u0, c, s, use_cf, h_re, h_im);
    const double sx = sqrt(X);
    const double pref = 2.0 * sx * sqrt(0.5 * kPi);
    if (use_cf) {
        // int_{u0}^inf e^{j pi u^2/2} = (0.5+0.5j) e^{jX} h, so
        // F = 2j sqrt(X) sqrt(pi/2) (0.5-0.5j) conj(h)   (the e^{+-jX} cancel)
        const double ch_re = h_re, ch_im = -h_im;
        const double m_re = 0.5 * ch_re + 0.5 * ch_im;   // (0.5-0.5j)*conj(h)
        const double m_im = 0.5 * ch_im - 0.5 * ch_re;
        // multiply by j * pref
        f_re = -pref * m_im;
        f_im = pref * m_re;
        return;
    }
    // G = conj((0.5 - C) + j(0.5 - S))
    const double g_re = 0.5 - c;
    const double g_im = -(0.5 - s);
    // F = j * pref * e^{jX} * G
    const double ex_re = cos(X), ex_im = sin(X);
    const double t_re = ex_re * g_re - ex_im * g_im;
    const double t_im = ex_re * g_im + ex_im * g_re;
    f_re = -pref * t_im;
    f_im = pref * t_re;
}

// Fr(x) = C(x) + j S(x)
WC_HD void fresnel_complex(double x, double& re, double& im) {
    double c, s, hr, hi;
    bool cf;
    fresnel_cs(x, c, s, cf, hr, hi);
    re = c;
    im = s;
}

// One cot * F term of the wedge coefficient. beta = phi -+ phi'; sgn selects
// a+ (+1) or a- (-1): N is the integer nearest to (beta + sgn*pi)/(2 n pi)
// and a(beta) = 2 cos^2((2 n pi N - beta)/2).
WC_HD void utd_term(double n, double beta, int sgn, double kL, double& re, double& im) {
    const double arg = (kPi + sgn * beta) / (2.0 * n);
    const double N = nearbyint((beta + sgn * kPi) / (2.0 * n * kPi));
    const double c = cos((2.0 * n * kPi * N - beta) * 0.5);
    const double a = 2.0 * c * c;
    const double X = kL * a;
    const double sa = sin(arg);
    if (fabs(sa) < 1.0e-9) {
        // On a shadow boundary cot -> inf and F -> 0; the product has the finite
        // limit n * sqrt(2 pi kL) * e^{j pi/4} * sign (Kouyoumjian-Pathak eq. 29).
        // signed offset from the boundary: eps = pi +- beta -+ 2 pi n N
        const double eps = (kPi + sgn * beta) - sgn * 2.0 * kPi * n * N;
        const double sg = eps >= 0.0 ? 1.0 : -1.0;
        const double mag = n * sqrt(2.0 * kPi * kL) * sg;
        re = mag * cos(0.25 * kPi);
        im = mag * sin(0.25 * kPi);
        (void)X;
        return;
    }
    const double cot = cos(arg) / sa;
    double f_re, f_im;
    utd_transition(X, f_re, f_im);
    re = cot * f_re;
    im = cot * f_im;
}

struct UtdCoefficients {
    float ds_re, ds_im, dh_re, dh_im;
};

// D_s, D_h of a wedge with exterior angle n*pi at observation angle phi and
// incidence angle phi_p (both measured from the 0-face into the exterior),
// skew angle beta0 (sin_beta0 > 0), spherical-wave distance parameter L.
// r0_s/r0_h, rn_s/rn_h are the 0-face and n-face reflection coefficients for
// the soft (E parallel to the edge) and hard polarizations (PEC: -1 / +1).
WC_HD UtdCoefficients utd_wedge(double n, double phi, double phi_p, double sin_beta0, double k, double L,
                                Cx<float> r0_s, Cx<float> rn_s, Cx<float> r0_h, Cx<float> rn_h) {
    const double kL = k * L;
    double t1r, t1i, t2r, t2i, t3r, t3i, t4r, t4i;
    utd_term(n, phi - phi_p, +1, kL, t1r, t1i);   // cot((pi + (phi-phi'))/2n) F(kL a+(phi-phi'))
    utd_term(n, phi - phi_p, -1, kL, t2r, t2i);   // cot((pi - (phi-phi'))/2n) F(kL a-(phi-phi'))
    utd_term(n, phi + phi_p, -1, kL, t3r, t3i);   // 0-face reflection term
    utd_term(n, phi + phi_p, +1, kL, t4r, t4i);   // n-face reflection term
    // prefactor -e^{-j pi/4} / (2 n sqrt(2 pi k) sin beta0)
    const double scale = 1.0 / (2.0 * n * sqrt(2.0 * kPi * k) * fmax(sin_beta0, 1.0e-6));
    const double p_re = -cos(-0.25 * kPi) * scale;
    const double p_im = -sin(-0.25 * kPi) * scale;
    auto assemble = [&](Cx<float> r0, Cx<float> rn, double& o_re, double& o_im) {
        const double s_re = t1r + t2r + ((double)r0.re * t3r - (double)r0.im * t3i) + ((double)rn.re * t4r - (double)rn.im * t4i);
        const double s_im = t1i + t2i + ((double)r0.re * t3i + (double)r0.im * t3r) + ((double)rn.re * t4i + (double)rn.im * t4r);
        o_re = p_re * s_re - p_im * s_im;
        o_im = p_re * s_im + p_im * s_re;
    };
    double a, b, c, d;
    assemble(r0_s, rn_s, a, b);
    assemble(r0_h, rn_h, c, d);
    return {(float)a, (float)b, (float)c, (float)d};
}

}  // namespace wc
