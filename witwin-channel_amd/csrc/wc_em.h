// Electromagnetic transport operators shared by every path kernel: free-space
// propagation, thin-slab Fresnel reflection, Jones re-projection at a bounce,
// layered (Rouard) transmission and the receiver projection. Templated on the
// scalar so the tangent pass differentiates exactly the forward arithmetic.
#pragma once

#include "wc_math.h"

namespace wc {

// Replace the value of a tangent-carrying phase with the double-precision
// reduction while keeping d(-k*L) as its derivative.
template <typename R> struct PhaseOf {
    WC_HD static R make(const R& k, const R& L, double k_d, double L_d) {
        (void)k; (void)L;
        return (R)neg_kd_phase(k_d, L_d);
    }
};
template <int N> struct PhaseOf<Dual<N>> {
    WC_HD static Dual<N> make(const Dual<N>& k, const Dual<N>& L, double k_d, double L_d) {
        Dual<N> r = -(k * L);
        r.v = neg_kd_phase(k_d, L_d);
        return r;
    }
};

template <typename R>
WC_HD R wave_number(const R& frequency_hz) {
    return frequency_hz * (float)(kTwoPi / kC0);
}

// Thin slab of thickness t between vacuum half-spaces: single-interface
// Fresnel coefficients with the interior Fabry-Perot series summed in closed
// form, r = g * r0 (1 - e^{-2jq}) / (1 - r0^2 e^{-2jq}), q = k t sqrt(eta mu - sin^2).
// Complex permittivity eta = eps_r - j sigma / (omega eps0)
// (reference tests/ad/_reference_fields.py:85-115; tests/kernels/test_field_transport.py:13-24).
template <typename R>
WC_HD void slab_fresnel(const R& cos_theta, const R& eps_r, const R& sigma_e, const R& mu_r, const R& gain,
                        const R& thickness, const R& frequency_hz, Cx<R>& r_te, Cx<R>& r_tm) {
    const R omega = w_max(frequency_hz * (float)kTwoPi, kSmallEpsF);
    const R wavelength = lift<R>((float)kC0) / frequency_hz;
    const R ct = w_min(w_max(w_abs(cos_theta), kSmallEpsF), 1.0f);
    const R sin2 = w_max(lift<R>(1.0f) - ct * ct, 0.0f);
    const R eps = w_max(eps_r, kSmallEpsF);
    const R sig = w_max(sigma_e, 0.0f);
    const R mu = w_max(mu_r, kSmallEpsF);
    // sigma / (omega eps0) evaluated as (sigma / omega) * (1 / eps0) to stay in float range
    const Cx<R> eta = {eps, -(sig / omega) * (float)(1.0 / kEps0)};
    const Cx<R> root = csqrt(Cx<R>{eta.re * mu - sin2, eta.im * mu});
    const Cx<R> mu_ct = {mu * ct, lift<R>(0.f)};
    const Cx<R> eta_ct = scale(eta, ct);
    const Cx<R> i_te = (mu_ct - root) / (mu_ct + root);
    const Cx<R> i_tm = (eta_ct - root) / (eta_ct + root);
    const R q_scale = w_max(thickness, 0.0f) * (float)kTwoPi / w_max(wavelength, kSmallEpsF);
    const Cx<R> q = scale(root, q_scale);
    // e^{-2jq} = e^{2 Im q} * e^{-2j Re q}
    const R mag = w_exp(q.im * 2.0f);
    const Cx<R> ph = scale(cexpj(-(q.re * 2.0f)), mag);
    const Cx<R> one = cxc<R>(1.f, 0.f);
    const Cx<R> num_te = i_te * (one - ph);
    const Cx<R> den_te = one - i_te * i_te * ph;
    const Cx<R> num_tm = i_tm * (one - ph);
    const Cx<R> den_tm = one - i_tm * i_tm * ph;
    r_te = scale(num_te / den_te, gain);
    r_tm = scale(num_tm / den_tm, gain);
}

// Reflect a transverse complex field 3-vector at one planar interface: orient
// the normal against the incident ray, build the s/p bases, apply the slab
// coefficients and recompose on the reflected p axis.
// (reference native/channel/kernels/fields.cu:113-204 via RayD reflect_complex3;
// restated from tests/ad/_reference_fields.py:118-186).
template <typename R>
WC_HD C3<R> reflect_c3(const C3<R>& value, const V3<R>& incident, const V3<R>& normal_in, const R& eps_r,
                       const R& sigma_e, const R& mu_r, const R& gain, const R& thickness,
                       const R& frequency_hz, V3<R>& outgoing) {
    const V3<R> ez = v3c<R>(0.f, 0.f, 1.f);
    V3<R> n = safe_normalize(normal_in, ez);
    if (val(dot(incident, n)) > 0.0f) n = -n;
    const R dot_in = dot(incident, n);
    const V3<R> reflected = safe_normalize(incident - n * (dot_in * 2.0f), -incident);
    const V3<R> s = safe_normalize(cross(n, incident), stable_perp_basis(incident, n));
    const V3<R> p_in = safe_normalize(cross(s, incident), stable_perp_basis(incident, s));
    const V3<R> p_out = safe_normalize(cross(s, reflected), stable_perp_basis(reflected, s));
    Cx<R> r_te, r_tm;
    slab_fresnel(w_abs(dot_in), eps_r, sigma_e, mu_r, gain, thickness, frequency_hz, r_te, r_tm);
    const Cx<R> e_s = c3_dot_real(value, s);
    const Cx<R> e_p = c3_dot_real(value, p_in);
    outgoing = reflected;
    return c3_axis(s, r_te * e_s) + c3_axis(p_out, r_tm * e_p);
}

// Layered wall, backward Rouard recursion over layers entry->exit between
// vacuum half-spaces; returns the amplitude transmission coefficient for one
// polarization (pol 0 = TE, 1 = TM). Layer arrays are float; admittances are
// normalized by the vacuum values so everything stays in float range.
// (reference tests/ad/_reference_fields.py:196-266, tests/reference/em_oracle.py:204).
// stack_rt also returns the stack's reflection coefficient seen from the
// entry side (r_out); thickness_override, like the eps / sigma overrides,
// replaces the float layer values with R values (tangent-carrying).
template <typename R>
WC_HD Cx<R> stack_rt(const R& cos_theta, const float* thickness, const float* eps_r, const float* sigma_e,
                     const float* mu_r, int count, const R& frequency_hz, int pol, const R* eps_override,
                     const R* sigma_override, const R* thickness_override, Cx<R>& r_out) {
    if (count <= 0) {
        r_out = cxc<R>(0.f, 0.f);
        return cxc<R>(1.f, 0.f);
    }
    const R omega = w_max(frequency_hz * (float)kTwoPi, kSmallEpsF);
    const R ct = w_min(w_max(w_abs(cos_theta), kSmallEpsF), 1.0f);
    const R sin2 = w_max(lift<R>(1.0f) - ct * ct, 0.0f);
    const R k0 = omega * (float)(1.0 / kC0);
    // Relative quantities: eps_c = eps_r - j sigma/(omega eps0), kz = k0 * sqrt(eps_c mu - sin2).
    // Normalized admittance: TE y = kz / (mu k0), TM y = eps_c k0 / kz.
    auto layer = [&](int i, Cx<R>& kz_rel, Cx<R>& y) {
        const R e = w_max(eps_override ? eps_override[i] : lift<R>(eps_r[i]), kSmallEpsF);
        const R s = w_max(sigma_override ? sigma_override[i] : lift<R>(sigma_e[i]), 0.0f);
        const R m = lift<R>(fmaxf(mu_r[i], kSmallEpsF));
        const Cx<R> eps_c = {e, -(s / omega) * (float)(1.0 / kEps0)};
        kz_rel = csqrt(Cx<R>{eps_c.re * m - sin2, eps_c.im * m});
        if (pol == 0) {
            y = scale(kz_rel, lift<R>(1.0f) / m);
        } else {
            y = eps_c / kz_rel;
        }
    };
    const Cx<R> kz_vac = {ct, lift<R>(0.f)};
    const Cx<R> y_vac = pol == 0 ? kz_vac : Cx<R>{lift<R>(1.0f) / ct, lift<R>(0.f)};
    const Cx<R> one = cxc<R>(1.f, 0.f);
    Cx<R> kz_below, y_below;
    layer(count - 1, kz_below, y_below);
    Cx<R> r_total = (y_below - y_vac) / (y_below + y_vac);
    Cx<R> t_total = scale(y_below, 2.0f) / (y_below + y_vac);
    for (int l = count - 1; l >= 0; --l) {
        // phase = exp(-j k0 kz_rel d)
        const R kd = k0 * (thickness_override ? w_max(thickness_override[l], 0.0f) : lift<R>(fmaxf(thickness[l], 0.0f)));
        const Cx<R> arg = scale(kz_below, kd);  // k0 kz d
        const Cx<R> phase = scale(cexpj(-arg.re), w_exp(arg.im));
        const Cx<R> phase2 = phase * phase;
        Cx<R> kz_above, y_above;
        if (l > 0) {
            layer(l - 1, kz_above, y_above);
        } else {
            kz_above = kz_vac;
            y_above = y_vac;
        }
        const Cx<R> r_top = (y_above - y_below) / (y_above + y_below);
        const Cx<R> t_top = scale(y_above, 2.0f) / (y_above + y_below);
        const Cx<R> loop = phase2 * r_total;
        const Cx<R> den = one + r_top * loop;
        r_total = (r_top + loop) / den;
        t_total = t_top * phase * t_total / den;
        kz_below = kz_above;
        y_below = y_above;
    }
    r_out = r_total;
    return t_total;
}

template <typename R>
WC_HD Cx<R> stack_t(const R& cos_theta, const float* thickness, const float* eps_r, const float* sigma_e,
                    const float* mu_r, int count, const R& frequency_hz, int pol, const R* eps_override,
                    const R* sigma_override) {
    Cx<R> r;
    return stack_rt<R>(cos_theta, thickness, eps_r, sigma_e, mu_r, count, frequency_hz, pol, eps_override,
                       sigma_override, nullptr, r);
}

// Complex coefficient seen by the receiver: p_rx,perp . E
// (reference fields.cu:84-105 project_complex3_kernel / RayD project_receiver).
template <typename R>
WC_HD Cx<R> project_receiver(const C3<R>& value, const V3<R>& direction, const V3<R>& rx_pol) {
    return c3_dot_real(value, transverse(direction, rx_pol));
}

}  // namespace wc
