// First-order uniform theory of diffraction (Kouyoumjian & Pathak, Proc. IEEE
// 62(11), 1974) for a straight wedge edge with exterior angle n*pi, with
// Luebbers' (IEEE TAP 32(1), 1984) face reflection coefficients for lossy
// faces and the finite-edge Fresnel truncation factor
//   T = (0.5 + 0.5j) * conj(Fr(sigma*(t_max - t*)) - Fr(sigma*(t_min - t*))),
//   sigma = sqrt(k*kappa/pi), kappa = rho'^2/R'^3 + rho^2/R^3
// (reference docs/dev/audit/utd-continuity-fix-design.md:91-140), applied
// inside the four-term assembly with the odd-part corner mend (E1d, :160-205).
// The transition function is evaluated in double precision: near a shadow
// boundary its argument kL*a goes to zero and the float cot*F product loses
// the digits that carry the geometrical-optics compensation.
#pragma once

#include "wc_em.h"
#include "wc_fresnel_aux_table.h"

namespace wc {

// Fresnel auxiliary functions f(x), g(x), x >= 0 (Abramowitz & Stegun
// 7.3.5-7.3.6), through which
//   C(x) = 1/2 + f sin(pi x^2/2) - g cos(pi x^2/2),
//   S(x) = 1/2 - f cos(pi x^2/2) - g sin(pi x^2/2),
//   int_x^inf e^{j pi t^2/2} dt = (g + j f) e^{j pi x^2/2}.
// Both are smooth and positive, so they carry full relative precision where
// C and S themselves cancel. Below x = 8: 12-term Chebyshev expansions on 32
// intervals (wc_fresnel_aux_table.h, generated in extended precision by
// gen_fresnel_aux.py; max relative error 1.6e-15); above: the asymptotic
// series A&S 7.3.27-7.3.28 to nine terms (truncation below 1e-19 at x = 8).
__device__ inline void fresnel_aux(double x, double& f, double& g) {
    if (x < kFresnelAuxMax) {
        const double scaled = x * (kFresnelAuxIntervals / kFresnelAuxMax);
        int i = (int)scaled;
        if (i > kFresnelAuxIntervals - 1) i = kFresnelAuxIntervals - 1;
        const double t = 2.0 * (scaled - (double)i) - 1.0;
        const double* cf = kFresnelAuxF[i];
        const double* cg = kFresnelAuxG[i];
        double f1 = 0.0, f2 = 0.0, g1 = 0.0, g2 = 0.0;
#pragma unroll
        for (int j = kFresnelAuxTerms - 1; j >= 1; --j) {
            const double nf = fma(2.0 * t, f1, cf[j] - f2);
            const double ng = fma(2.0 * t, g1, cg[j] - g2);
            f2 = f1;
            f1 = nf;
            g2 = g1;
            g1 = ng;
        }
        f = fma(t, f1, cf[0] - f2);
        g = fma(t, g1, cg[0] - g2);
        return;
    }
    const double z = kPi * x * x;
    const double inv_z2 = 1.0 / (z * z);
    double tf = 1.0, tg = 1.0, sf = 1.0, sg = 1.0;
#pragma unroll
    for (int m = 1; m <= 9; ++m) {
        tf *= -(double)((4 * m - 3) * (4 * m - 1)) * inv_z2;
        tg *= -(double)((4 * m - 1) * (4 * m + 1)) * inv_z2;
        sf += tf;
        sg += tg;
    }
    f = sf / (kPi * x);
    g = sg / (kPi * kPi * x * x * x);
}

// Kouyoumjian-Pathak transition function
//   F(X) = 2j sqrt(X) e^{jX} int_{sqrt X}^inf e^{-j tau^2} dtau,  X >= 0.
// With tau = u sqrt(pi/2) and u0 = sqrt(2X/pi) (so pi u0^2/2 = X) the
// integral is sqrt(pi/2) conj((g + j f)(u0) e^{jX}), the e^{+-jX} cancel and
//   F(X) = sqrt(2 pi X) (f(u0) + j g(u0)),
// which keeps full relative precision as X -> 0 (F ~ sqrt(pi X) e^{j pi/4},
// the shadow-boundary regime) and tends to 1 + j/(2X) for large X.
__device__ inline void utd_transition(double X, double& f_re, double& f_im) {
    if (X <= 0.0) {
        f_re = 0.0;
        f_im = 0.0;
        return;
    }
    double f, g;
    fresnel_aux(sqrt(2.0 * X / kPi), f, g);
    const double r = sqrt(2.0 * kPi * X);
    f_re = r * f;
    f_im = r * g;
}

// Fr(x) = C(x) + j S(x) (absolute accuracy ~1e-16); optionally also the
// phasor e^{j pi x^2/2} = dFr/dx. The phase is reduced as (pi/2) (x^2 mod 4)
// with x^2 split exactly into hi + lo (fma), so large x lose no phase digits.
__device__ inline void fresnel_complex(double x, double& re, double& im, double* d_re = nullptr,
                                       double* d_im = nullptr) {
    const double ax = fabs(x);
    double f, g;
    fresnel_aux(ax, f, g);
    const double hi = ax * ax;
    const double lo = fma(ax, ax, -hi);
    double sn, cs;
    sincos(0.5 * kPi * (fmod(hi, 4.0) + lo), &sn, &cs);
    double c = 0.5 + f * sn - g * cs;
    double s = 0.5 - f * cs - g * sn;
    if (x < 0.0) {
        c = -c;
        s = -s;
    }
    re = c;
    im = s;
    if (d_re) {
        *d_re = cs;
        *d_im = sn;
    }
}

// One cot * F term of the wedge coefficient. beta = phi -+ phi'; sgn selects
// a+ (+1) or a- (-1): N is the integer nearest to (beta + sgn*pi)/(2 n pi)
// and a(beta) = 2 cos^2((2 n pi N - beta)/2).
__device__ inline void utd_term(double n, double beta, int sgn, double kL, double& re, double& im) {
    const double arg = (kPi + sgn * beta) / (2.0 * n);
    const double N = nearbyint((beta + sgn * kPi) / (2.0 * n * kPi));
    // a = 2 cos^2 of the half angle (keeps relative precision where a -> 0)
    const double c = cos((2.0 * n * kPi * N - beta) * 0.5);
    const double a = 2.0 * c * c;
    const double X = kL * a;
    double sa, ca;
    sincos(arg, &sa, &ca);
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
    const double cot = ca / sa;
    double f_re, f_im;
    utd_transition(X, f_re, f_im);
    re = cot * f_re;
    im = cot * f_im;
}

struct UtdCoefficients {
    float ds_re, ds_im, dh_re, dh_im;
};

// Finite-edge parameters of the E1d assembly: the truncation factor
// T = (0.5 + 0.5j) conj(Fr(sigma (l - t*)) - Fr(sigma (0 - t*))) and the corner
// blend gamma = 1 / (1 + exp(u / 0.1)) of the stationary point's exit
// coordinate u = sigma max(-t*, t* - l) (u < 0 inside the edge).
struct FiniteEdge {
    double t_re, t_im, gamma;
};

// Nearest geometrical-optics boundaries of the two cot*F groups: the incident
// group beta = phi - phi' has boundaries -pi (term 1) and +pi (term 2), the
// reflection group beta = phi + phi' has pi (term 3, face 0) and (2n - 1) pi
// (term 4, face n). Only the term singular at the nearer boundary of its
// group is mended.
struct UtdBoundaries {
    bool incident_up;  // term 2 (+pi) active, else term 1 (-pi)
    bool reflect_0;    // term 3 (pi) active, else term 4 ((2n-1) pi)
};
__device__ inline UtdBoundaries utd_boundaries(double n, double phi, double phi_p) {
    const double br = phi + phi_p;
    return {phi - phi_p >= 0.0, fabs(br - kPi) <= fabs(br - (2.0 * n - 1.0) * kPi)};
}

// One cot*F term in the E1d assembly (reference
// docs/dev/audit/utd-continuity-fix-design.md:160-205): the active term is
// split about its boundary b* into odd_w = (t(beta) - t(2 b* - beta)) / 2 * w,
// w = exp(-((beta - b*) / delta_w)^2), delta_w = 4 sqrt(2 pi / kL), and enters
// as (t - odd_w) T + odd_w gamma, so the step-carrying odd part keeps factor
// gamma (1 inside the edge) and only the smooth background is truncated; an
// inactive term enters as t T. Where w < 1e-20 the mirrored evaluation is
// skipped (its contribution is below double rounding).
__device__ inline void utd_term_e1d(double n, double beta, int sgn, double kL, bool active, double bstar,
                                    const FiniteEdge& fe, double& re, double& im) {
    double tr, ti;
    utd_term(n, beta, sgn, kL, tr, ti);
    double orr = 0.0, oi = 0.0;
    if (active) {
        const double dw = 4.0 * sqrt(2.0 * kPi / kL);
        const double d = (beta - bstar) / dw;
        const double w = exp(-d * d);
        if (w > 1.0e-20) {
            double mr, mi;
            utd_term(n, 2.0 * bstar - beta, sgn, kL, mr, mi);
            orr = 0.5 * (tr - mr) * w;
            oi = 0.5 * (ti - mi) * w;
        }
    }
    const double ar = tr - orr, ai = ti - oi;
    re = ar * fe.t_re - ai * fe.t_im + orr * fe.gamma;
    im = ar * fe.t_im + ai * fe.t_re + oi * fe.gamma;
}

// D_s, D_h of a wedge with exterior angle n*pi at observation angle phi and
// incidence angle phi_p (both measured from the 0-face into the exterior),
// skew angle beta0 (sin_beta0 > 0), spherical-wave distance parameter L, with
// the finite-edge factors folded into the four terms (utd_term_e1d).
// r0_s/r0_h, rn_s/rn_h are the 0-face and n-face reflection coefficients for
// the soft (E parallel to the edge) and hard polarizations (PEC: -1 / +1).
__device__ inline UtdCoefficients utd_wedge(double n, double phi, double phi_p, double sin_beta0, double k, double L,
                                Cx<float> r0_s, Cx<float> rn_s, Cx<float> r0_h, Cx<float> rn_h,
                                const FiniteEdge& fe) {
    const double kL = k * L;
    const UtdBoundaries b = utd_boundaries(n, phi, phi_p);
    double t1r, t1i, t2r, t2i, t3r, t3i, t4r, t4i;
    utd_term_e1d(n, phi - phi_p, +1, kL, !b.incident_up, -kPi, fe, t1r, t1i);  // cot((pi + b-)/2n) F(kL a+(b-))
    utd_term_e1d(n, phi - phi_p, -1, kL, b.incident_up, kPi, fe, t2r, t2i);    // cot((pi - b-)/2n) F(kL a-(b-))
    utd_term_e1d(n, phi + phi_p, -1, kL, b.reflect_0, kPi, fe, t3r, t3i);      // 0-face reflection term
    utd_term_e1d(n, phi + phi_p, +1, kL, !b.reflect_0, (2.0 * n - 1.0) * kPi, fe, t4r, t4i);  // n-face term
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
    double a, bb, c, d;
    assemble(r0_s, rn_s, a, bb);
    assemble(r0_h, rn_h, c, d);
    return {(float)a, (float)bb, (float)c, (float)d};
}

}  // namespace wc
