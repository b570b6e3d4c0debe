// FP64 CPU oracle for the deterministic / path propagation hot path.
//
// TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load this library, and only as the
// checker or the timed CPU baseline; the product path never calls it.
//
// Every quantity is computed in float64 / complex<double>; visibility is a
// brute-force test of the segment against every triangle (no acceleration
// structure), so nothing here shares code or traversal order with the GPU
// kernels. Each function cites the reference source it restates:
//   free space ............ native/channel/kernels/fields.cu:36-82,
//                           tests/ad/_reference_fields.py:64-82
//   slab Fresnel .......... tests/ad/_reference_fields.py:85-115,
//                           tests/kernels/test_field_transport.py:13-24
//   reflection sequence ... native/channel/kernels/fields.cu:113-204,
//                           tests/ad/_reference_fields.py:118-186
//   Rouard layer stack .... tests/ad/_reference_fields.py:196-266,
//                           tests/reference/em_oracle.py:204
//   transmission sequence . tests/ad/_reference_fields.py:269-336
//   UTD ................... Kouyoumjian & Pathak 1974; Luebbers 1984;
//                           finite-edge factor and E1d assembly of
//                           docs/dev/audit/utd-continuity-fix-design.md:91-205
//   accumulation .......... native/channel/kernels/deterministic.cu:829-970
//   canonical order ....... native/channel/kernels/evaluated_paths.cu:126-310,654-700
//
// Parity status per function is recorded in the header comment of
// oracle/__init__.py.

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

using cd = std::complex<double>;
constexpr double C0 = 299792458.0;
constexpr double EPS0 = 8.8541878128e-12;
constexpr double PI = 3.14159265358979323846;
constexpr double EPS = 1.0e-10;
constexpr double SMALL = 1.0e-6;
constexpr double BIAS = 1.0e-4;   // segment end clearance (m)
constexpr double FLOOR = 1.0e-30; // diffraction existence floor on path gain
constexpr int MAXD = 5;

struct V {
    double x, y, z;
};
inline V operator+(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V operator-(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V operator-(V a) { return {-a.x, -a.y, -a.z}; }
inline V operator*(V a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V cross(V a, V b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double len(V a) { return std::sqrt(dot(a, a)); }

struct CV {
    cd x, y, z;
};
inline CV cv_axis(V a, cd s) { return {a.x * s, a.y * s, a.z * s}; }
inline CV operator+(CV a, CV b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline CV operator*(CV a, cd s) { return {a.x * s, a.y * s, a.z * s}; }
inline cd cdot(CV a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

V safe_normalize(V v, V fallback) {
    const double n = len(v);
    if (n > SMALL) return v * (1.0 / (n + EPS));
    return fallback * (1.0 / (len(fallback) + EPS));
}

V stable_perp(V ray, V preferred) {
    const V proj = preferred - ray * dot(preferred, ray);
    const V alt = std::fabs(ray.z) < 0.9 ? V{0, 0, 1} : V{0, 1, 0};
    const V alt_proj = alt - ray * dot(alt, ray);
    return safe_normalize(proj, alt_proj);
}

V transverse(V ray, V pol) { return pol - ray * dot(pol, ray); }

// ---------------------------------------------------------------- EM
void slab_fresnel(double cos_theta, double eps_r, double sigma, double mu_r, double gain, double thickness,
                  double f, cd& r_te, cd& r_tm) {
    const double omega = std::max(2.0 * PI * f, SMALL);
    const double wavelength = C0 / f;
    const double ct = std::min(std::max(std::fabs(cos_theta), SMALL), 1.0);
    const double sin2 = std::max(1.0 - ct * ct, 0.0);
    const cd eta(std::max(eps_r, SMALL), -std::max(sigma, 0.0) / (omega * EPS0));
    const double mu = std::max(mu_r, SMALL);
    const cd root = std::sqrt(eta * mu - sin2);
    const cd mu_ct(mu * ct, 0.0);
    const cd eta_ct = eta * ct;
    const cd ite = (mu_ct - root) / (mu_ct + root);
    const cd itm = (eta_ct - root) / (eta_ct + root);
    const cd q = root * (2.0 * PI * std::max(thickness, 0.0) / std::max(wavelength, SMALL));
    const cd ph = std::exp(cd(0.0, -2.0) * q);
    r_te = gain * (ite * (1.0 - ph)) / (1.0 - ite * ite * ph);
    r_tm = gain * (itm * (1.0 - ph)) / (1.0 - itm * itm * ph);
}

CV reflect(CV value, V incident, V normal, double eps, double sig, double mu, double gain, double thick, double f,
           V& outgoing) {
    V n = safe_normalize(normal, {0, 0, 1});
    if (dot(incident, n) > 0.0) n = -n;
    const double din = dot(incident, n);
    const V refl = safe_normalize(incident - n * (2.0 * din), -incident);
    const V s = safe_normalize(cross(n, incident), stable_perp(incident, n));
    const V p_in = safe_normalize(cross(s, incident), stable_perp(incident, s));
    const V p_out = safe_normalize(cross(s, refl), stable_perp(refl, s));
    cd rte, rtm;
    slab_fresnel(std::fabs(din), eps, sig, mu, gain, thick, f, rte, rtm);
    const cd es = cdot(value, s), ep = cdot(value, p_in);
    outgoing = refl;
    return cv_axis(s, rte * es) + cv_axis(p_out, rtm * ep);
}

// Rouard recursion over layers entry->exit, vacuum both sides; returns t.
cd stack_t(double cos_theta, const float* th, const float* er, const float* sg, const float* mr, int count, double f,
           int pol) {
    if (count <= 0) return cd(1.0, 0.0);
    const double omega = std::max(2.0 * PI * f, SMALL);
    const double ct = std::min(std::max(std::fabs(cos_theta), SMALL), 1.0);
    const double sin2 = std::max(1.0 - ct * ct, 0.0);
    const double k0 = omega / C0;
    auto medium = [&](int i, cd& kz, cd& y) {
        const cd ec(std::max((double)er[i], SMALL), -std::max((double)sg[i], 0.0) / (omega * EPS0));
        const double m = std::max((double)mr[i], SMALL);
        kz = std::sqrt(ec * m - sin2);
        y = pol == 0 ? kz / m : ec / kz;
    };
    const cd kz_vac(ct, 0.0);
    const cd y_vac = pol == 0 ? kz_vac : cd(1.0 / ct, 0.0);
    cd kzb, yb;
    medium(count - 1, kzb, yb);
    cd r = (yb - y_vac) / (yb + y_vac);
    cd t = 2.0 * yb / (yb + y_vac);
    for (int l = count - 1; l >= 0; --l) {
        const cd phase = std::exp(cd(0.0, -1.0) * kzb * (k0 * std::max((double)th[l], 0.0)));
        const cd phase2 = phase * phase;
        cd kza, ya;
        if (l > 0) medium(l - 1, kza, ya); else { kza = kz_vac; ya = y_vac; }
        const cd rtop = (ya - yb) / (ya + yb);
        const cd ttop = 2.0 * ya / (ya + yb);
        const cd loop = phase2 * r;
        const cd den = 1.0 + rtop * loop;
        r = (rtop + loop) / den;
        t = ttop * phase * t / den;
        kzb = kza;
        yb = ya;
    }
    return t;
}

// Fresnel integrals by direct numerical quadrature-free series / asymptotic
// forms in long double: C + jS = int_0^x e^{j pi t^2/2} dt.
void fresnel(double x, double& c, double& s) {
    const double ax = std::fabs(x);
    if (ax < 4.0) {
        // power series in long double (converges; |terms| <= e^{pi x^2/2} ~ 1e11 at x=4,
        // leaving ~1e-8 absolute error at worst in 64-bit mantissa)
        long double sum_c = 0.0L, sum_s = 0.0L;
        const long double z = (long double)PI * ax * ax / 2.0L;
        long double term = 1.0L;  // z^n / n!
        for (int n = 0; n < 200; ++n) {
            if (n > 0) term *= z / n;
            const long double sign = ((n / 2) % 2 == 0) ? 1.0L : -1.0L;
            const long double contrib = sign * term / (2 * n + 1);
            if (n % 2 == 0) sum_c += contrib; else sum_s += contrib;
            if (n > 10 && term < 1e-22L) break;
        }
        c = (double)(sum_c * ax);
        s = (double)(sum_s * ax);
    } else {
        // auxiliary functions f, g asymptotic series (Abramowitz & Stegun 7.3.27-28 with 7.3.9-10)
        const double z = PI * ax * ax;
        double f = 0.0, g = 0.0;
        {
            // f(x) ~ 1/(pi x) sum (-1)^m (1*3*...*(4m-1)) / (pi x^2)^{2m}
            double term = 1.0;
            double sum = 1.0;
            for (int m = 1; m < 12; ++m) {
                term *= -((4.0 * m - 3.0) * (4.0 * m - 1.0)) / (z * z);
                if (std::fabs(term) < 1e-17) break;
                sum += term;
            }
            f = sum / (PI * ax);
        }
        {
            double term = 1.0;
            double sum = 1.0;
            for (int m = 1; m < 12; ++m) {
                term *= -((4.0 * m - 1.0) * (4.0 * m + 1.0)) / (z * z);
                if (std::fabs(term) < 1e-17) break;
                sum += term;
            }
            g = sum / (PI * PI * ax * ax * ax);
        }
        const double arg = 0.5 * PI * ax * ax;
        c = 0.5 + f * std::sin(arg) - g * std::cos(arg);
        s = 0.5 - f * std::cos(arg) - g * std::sin(arg);
    }
    if (x < 0.0) {
        c = -c;
        s = -s;
    }
}

// K-P transition function F(X) = 2j sqrt(X) e^{jX} int_{sqrt X}^inf e^{-j t^2} dt.
cd transition(double X) {
    if (X <= 0.0) return cd(0.0, 0.0);
    const double u0 = std::sqrt(2.0 * X / PI);
    double c, s;
    fresnel(u0, c, s);
    const cd g(0.5 - c, -(0.5 - s));
    return cd(0.0, 2.0 * std::sqrt(X) * std::sqrt(PI / 2.0)) * std::exp(cd(0.0, X)) * g;
}

cd utd_term(double n, double beta, int sgn, double kL) {
    const double arg = (PI + sgn * beta) / (2.0 * n);
    const double N = std::nearbyint((beta + sgn * PI) / (2.0 * n * PI));
    const double cc = std::cos((2.0 * n * PI * N - beta) * 0.5);
    const double a = 2.0 * cc * cc;
    const double sa = std::sin(arg);
    if (std::fabs(sa) < 1.0e-9) {
        const double e = (PI + sgn * beta) - sgn * 2.0 * PI * n * N;
        return n * std::sqrt(2.0 * PI * kL) * (e >= 0.0 ? 1.0 : -1.0) * std::exp(cd(0.0, PI / 4.0));
    }
    return (std::cos(arg) / sa) * transition(kL * a);
}

// Wedge coefficients with the E1d finite-edge assembly
// (reference docs/dev/audit/utd-continuity-fix-design.md:160-205). Each cot*F
// group (incident: beta = phi - phi', boundaries -pi / +pi; reflection:
// beta = phi + phi', boundaries pi / (2n-1) pi) splits the term that is
// singular at the boundary b* nearest to beta into an odd part
//   odd_w = (t(beta) - t(2 b* - beta)) / 2 * exp(-(delta / delta_w)^2),
//   delta = beta - b*, delta_w = 4 sqrt(2 pi / kL),
// and enters as (t - odd_w) T + odd_w gamma; every other term enters as t T.
// T is the finite-edge truncation factor, gamma = 1 / (1 + exp(u / 0.1)) the
// corner blend of the stationary point's exit coordinate u (u < 0 inside the
// edge). With T = gamma = 1 this is the plain Kouyoumjian-Pathak wedge.
void utd_wedge(double n, double phi, double phip, double sinb, double k, double L, cd r0s, cd rns, cd r0h, cd rnh,
               cd trunc, double gamma, cd& ds, cd& dh) {
    const double kL = k * L;
    const double dw = 4.0 * std::sqrt(2.0 * PI / kL);
    auto term = [&](double beta, int sgn, bool active, double bstar) {
        const cd t = utd_term(n, beta, sgn, kL);
        if (!active) return t * trunc;
        const double delta = (beta - bstar) / dw;
        const cd odd = 0.5 * (t - utd_term(n, 2.0 * bstar - beta, sgn, kL)) * std::exp(-delta * delta);
        return (t - odd) * trunc + odd * gamma;
    };
    const double bi = phi - phip, br = phi + phip;
    const bool up = bi >= 0.0;                                             // nearest incident boundary +pi
    const bool face0 = std::fabs(br - PI) <= std::fabs(br - (2.0 * n - 1.0) * PI);  // nearest reflection boundary pi
    const cd t1 = term(bi, +1, !up, -PI);
    const cd t2 = term(bi, -1, up, PI);
    const cd t3 = term(br, -1, face0, PI);
    const cd t4 = term(br, +1, !face0, (2.0 * n - 1.0) * PI);
    const cd pre = -std::exp(cd(0.0, -PI / 4.0)) / (2.0 * n * std::sqrt(2.0 * PI * k) * std::max(sinb, 1.0e-6));
    ds = pre * (t1 + t2 + r0s * t3 + rns * t4);
    dh = pre * (t1 + t2 + r0h * t3 + rnh * t4);
}

// ---------------------------------------------------------------- scene
struct Scene {
    int face_count;
    const float* tri;        // [F, 9] vertices of each face, original order
    const int* face_group;   // [F]
    const int* face_material;
    int group_count;
    const double* plane;     // [G,4]
    const int* group_off;    // [G+1]
    const int* group_faces;
    const float* mat_eps;
    const float* mat_sigma;
    const float* mat_mu;
    const float* mat_gain;
    const float* mat_thick;
    const int* layer_off;
    const int* layer_cnt;
    const float* layer_thick;
    const float* layer_eps;
    const float* layer_sigma;
    const float* layer_mu;
    int edge_count;
    const float* edge_f;     // [E,14] floats of wc_edge
    const int* edge_i;       // [E,6] ints of wc_edge
};

struct Solve {
    double tx[3], tx_pol[3], tx_power;
    double rx[3], rx_pol[3];
    double frequency;
    int components;
    int max_depth;
    int chain_count;
    const int* chain_groups;  // [K,5]
    const int* chain_depth;   // [K]
};

V tri_v(const Scene& s, int f, int k) {
    const float* p = s.tri + 9 * f + 3 * k;
    return {(double)p[0], (double)p[1], (double)p[2]};
}

// Moller-Trumbore in double; hit when t in (tmin, tmax) and barycentrics inside.
bool seg_hits_tri(const Scene& s, int f, V o, V d, double tmin, double tmax, double* t_out) {
    const V a = tri_v(s, f, 0), b = tri_v(s, f, 1), c = tri_v(s, f, 2);
    const V e1 = b - a, e2 = c - a;
    const V p = cross(d, e2);
    const double det = dot(e1, p);
    if (std::fabs(det) < 1e-300) return false;
    const double inv = 1.0 / det;
    const V tv = o - a;
    const double u = dot(tv, p) * inv;
    if (u < 0.0 || u > 1.0) return false;
    const V q = cross(tv, e1);
    const double v = dot(d, q) * inv;
    if (v < 0.0 || u + v > 1.0) return false;
    const double t = dot(e2, q) * inv;
    if (!(t > tmin && t < tmax)) return false;
    if (t_out) *t_out = t;
    return true;
}

bool occluded(const Scene& s, V a, V b, int ig0, int ig1) {
    const V d = b - a;
    const double l = len(d);
    if (l <= 2.0 * BIAS) return false;
    const double eps = BIAS / l;
    for (int f = 0; f < s.face_count; ++f) {
        const int g = s.face_group[f];
        if (g == ig0 || g == ig1) continue;
        if (seg_hits_tri(s, f, a, d, eps, 1.0 - eps, nullptr)) return true;
    }
    return false;
}

V plane_n(const Scene& s, int g, double& off) {
    const double* p = s.plane + 4 * g;
    off = p[3];
    return {p[0], p[1], p[2]};
}

int face_containing(const Scene& s, int g, V p, V n) {
    for (int i = s.group_off[g]; i < s.group_off[g + 1]; ++i) {
        const int f = s.group_faces[i];
        const V a = tri_v(s, f, 0), b = tri_v(s, f, 1), c = tri_v(s, f, 2);
        const double e0 = dot(cross(b - a, p - a), n), e1 = dot(cross(c - b, p - b), n),
                     e2 = dot(cross(a - c, p - c), n);
        if ((e0 >= 0 && e1 >= 0 && e2 >= 0) || (e0 <= 0 && e1 <= 0 && e2 <= 0)) return f;
    }
    return -1;
}

// ---------------------------------------------------------------- coupled paths
// Restatement of the coupled reflection-diffraction candidates (reference
// interactions/coupled.py:93-230, native/channel/kernels/coupled.cu:2279-2710)
// in float64 with brute-force visibility: R->D via the image of the
// transmitter, D->R via the image of the receiver, D->D by bisection of the
// path length's derivative along edge 2 (edge-1 Keller point in closed form).
struct EdgeRec {
    V A, d, t0, n0;
    double length, n;
    int group0, group1, mat0, mat1;
};

EdgeRec edge_rec(const Scene& s, int e) {
    const float* f = s.edge_f + 14 * e;
    const int* i = s.edge_i + 6 * e;
    return {{f[0], f[1], f[2]}, {f[3], f[4], f[5]}, {f[8], f[9], f[10]}, {f[11], f[12], f[13]}, f[6], f[7],
            i[2], i[3], i[4], i[5]};
}

struct Keller {
    V q;
    double t, sp, so, rp, rr, phip, phi;
};

// Keller point of end points a (source side) and b on the edge line; false
// when either end is on the line or the source is inside the wedge
bool keller(const EdgeRec& e, V a, V b, Keller& k) {
    const double ta = dot(a - e.A, e.d), tb = dot(b - e.A, e.d);
    const V pa = a - e.A - e.d * ta, pb = b - e.A - e.d * tb;
    k.rp = len(pa);
    k.rr = len(pb);
    if (k.rp < 1e-6 || k.rr < 1e-6) return false;
    k.t = ta + (tb - ta) * k.rp / (k.rp + k.rr);
    k.q = e.A + e.d * k.t;
    k.sp = len(a - k.q);
    k.so = len(b - k.q);
    auto ang = [&](V u) { double x = std::atan2(dot(u, e.n0), dot(u, e.t0)); if (x < 0) x += 2 * PI; return x; };
    const double npi = e.n * PI;
    k.phip = ang(pa);
    if (k.phip >= npi) return false;
    double phi = ang(pb);
    if (phi > npi) phi = (phi - npi < 2 * PI - phi) ? npi : 0.0;
    k.phi = phi;
    return true;
}

// K-P wedge diffraction of the incident vector einc at k.q (spreading from
// its source applied) observed along src -> q -> obs:
// [-Ds b0 b0' - Dh phi phi'] einc sqrt(sp/(so(so+sp))) e^{-jk so}
CV wedge(const Scene& s, const EdgeRec& e, const Keller& k, V src, V obs, CV einc, double f, double kw, V& so_dir) {
    const V si = (k.q - src) * (1.0 / k.sp);
    const V so = (obs - k.q) * (1.0 / k.so);
    so_dir = so;
    const double cb = dot(si, e.d);
    const double sinb = std::sqrt(std::max(1.0 - cb * cb, 0.0));
    V phi_i = -cross(e.d, si); phi_i = phi_i * (1.0 / std::max(len(phi_i), 1e-30));
    const V b0_i = cross(phi_i, si);
    V phi_o = cross(e.d, so); phi_o = phi_o * (1.0 / std::max(len(phi_o), 1e-30));
    const V b0_o = cross(phi_o, so);
    const cd eb = cdot(einc, b0_i), ep = cdot(einc, phi_i);
    const double L = k.so * k.sp * sinb * sinb / (k.so + k.sp);
    const double npi = e.n * PI;
    const int m0 = e.mat0, mn = e.mat1 >= 0 ? e.mat1 : e.mat0;
    cd r0s, r0h, rns, rnh;
    slab_fresnel(std::fabs(std::sin(k.phip)), s.mat_eps[m0], s.mat_sigma[m0], s.mat_mu[m0], s.mat_gain[m0],
                 s.mat_thick[m0], f, r0s, r0h);
    slab_fresnel(std::fabs(std::sin(npi - k.phi)), s.mat_eps[mn], s.mat_sigma[mn], s.mat_mu[mn], s.mat_gain[mn],
                 s.mat_thick[mn], f, rns, rnh);
    const double kappa = k.rp * k.rp / (k.sp * k.sp * k.sp) + k.rr * k.rr / (k.so * k.so * k.so);
    const double sg = std::sqrt(kw * kappa / PI);
    double c1, s1, c0, s0;
    fresnel(sg * (e.length - k.t), c1, s1);
    fresnel(sg * (0.0 - k.t), c0, s0);
    const cd trunc = cd(0.5, 0.5) * std::conj(cd(c1 - c0, s1 - s0));
    const double u = sg * std::max(-k.t, k.t - e.length);
    const double gamma = 1.0 / (1.0 + std::exp(u / 0.1));
    cd ds, dh;
    utd_wedge(e.n, k.phi, k.phip, sinb, kw, L, r0s, rns, r0h, rnh, trunc, gamma, ds, dh);
    const cd prop = std::sqrt(k.sp / (k.so * (k.so + k.sp))) * std::exp(cd(0.0, -kw * k.so));
    return cv_axis(b0_o, -ds * eb * prop) + cv_axis(phi_o, -dh * ep * prop);
}

bool occluded4(const Scene& s, V a, V b, int g0, int g1, int g2, int g3) {
    const V d = b - a;
    const double l = len(d);
    if (l <= 2.0 * BIAS) return false;
    const double eps = BIAS / l;
    for (int f = 0; f < s.face_count; ++f) {
        const int g = s.face_group[f];
        if (g == g0 || g == g1 || g == g2 || g == g3) continue;
        if (seg_hits_tri(s, f, a, d, eps, 1.0 - eps, nullptr)) return true;
    }
    return false;
}

}  // namespace

extern "C" {

struct OraclePath {
    int component;
    int depth;
    int prim[MAXD];
    int mat[MAXD];
    int edge;
    double length;
    double pos[MAXD][3];
    double field[6];       // complex 3-vector, unit excitation
    double coefficient[2];
    double path_field[2];
    double gain;
    double arrival[3];
    double departure[3];
};

// ------------------------------------------------------------ unit operators
void oracle_slab_fresnel(double cos_theta, double eps, double sigma, double mu, double gain, double thick, double f,
                         double* out4) {
    cd te, tm;
    slab_fresnel(cos_theta, eps, sigma, mu, gain, thick, f, te, tm);
    out4[0] = te.real(); out4[1] = te.imag(); out4[2] = tm.real(); out4[3] = tm.imag();
}

void oracle_stack_t(double cos_theta, const float* th, const float* er, const float* sg, const float* mr, int count,
                    double f, int pol, double* out2) {
    const cd t = stack_t(cos_theta, th, er, sg, mr, count, f, pol);
    out2[0] = t.real();
    out2[1] = t.imag();
}

void oracle_transition(double X, double* out2) {
    const cd v = transition(X);
    out2[0] = v.real();
    out2[1] = v.imag();
}

void oracle_fresnel(double x, double* out2) { fresnel(x, out2[0], out2[1]); }

void oracle_utd(double n, double phi, double phip, double sinb, double k, double L, const double* r8, double* out4) {
    cd ds, dh;
    utd_wedge(n, phi, phip, sinb, k, L, cd(r8[0], r8[1]), cd(r8[2], r8[3]), cd(r8[4], r8[5]), cd(r8[6], r8[7]),
              cd(1.0, 0.0), 1.0, ds, dh);
    out4[0] = ds.real(); out4[1] = ds.imag(); out4[2] = dh.real(); out4[3] = dh.imag();
}

// Wedge coefficients with finite-edge factors T = t2[0] + j t2[1] and gamma.
void oracle_utd_finite(double n, double phi, double phip, double sinb, double k, double L, const double* r8,
                       const double* t2, double gamma, double* out4) {
    cd ds, dh;
    utd_wedge(n, phi, phip, sinb, k, L, cd(r8[0], r8[1]), cd(r8[2], r8[3]), cd(r8[4], r8[5]), cd(r8[6], r8[7]),
              cd(t2[0], t2[1]), gamma, ds, dh);
    out4[0] = ds.real(); out4[1] = ds.imag(); out4[2] = dh.real(); out4[3] = dh.imag();
}

// Free-space row: out = [field(6), coefficient(2), path_field(2), gain, length]
void oracle_free_space(const double* src, const double* tgt, double power, const double* txp, const double* rxp,
                       double f, double* out) {
    const V a{src[0], src[1], src[2]}, b{tgt[0], tgt[1], tgt[2]};
    const V off = b - a;
    const double d = len(off);
    const V dir = safe_normalize(off, {0, 0, 1});
    const double k = 2.0 * PI * f / C0;
    const cd carrier = (1.0 / (2.0 * std::max(k, SMALL) * std::max(d, EPS))) * std::exp(cd(0.0, -k * d));
    const CV v = cv_axis(transverse(dir, {txp[0], txp[1], txp[2]}), carrier);
    const cd coef = cdot(v, transverse(dir, {rxp[0], rxp[1], rxp[2]}));
    const cd pf = coef * std::sqrt(std::max(power, 0.0));
    const cd xs[3] = {v.x, v.y, v.z};
    for (int i = 0; i < 3; ++i) { out[2 * i] = xs[i].real(); out[2 * i + 1] = xs[i].imag(); }
    out[6] = coef.real(); out[7] = coef.imag(); out[8] = pf.real(); out[9] = pf.imag();
    out[10] = std::norm(pf);
    out[11] = d;
}

// Reflection-sequence row with explicit interaction points: same output layout.
void oracle_reflection_sequence(int depth, const double* src, const double* tgt, const double* pos,
                                const double* nrm, double power, const double* txp, const double* rxp,
                                const double* eps, const double* sig, const double* mu, const double* gain,
                                const double* thick, double f, double* out) {
    V prev{src[0], src[1], src[2]};
    V inc = safe_normalize(V{pos[0], pos[1], pos[2]} - prev, {0, 0, 1});
    CV value = cv_axis(transverse(inc, {txp[0], txp[1], txp[2]}), cd(1.0, 0.0));
    V outg = inc;
    double total = 0.0;
    for (int b = 0; b < depth; ++b) {
        const V h{pos[3 * b], pos[3 * b + 1], pos[3 * b + 2]};
        inc = safe_normalize(h - prev, outg);
        total += len(h - prev);
        value = reflect(value, inc, {nrm[3 * b], nrm[3 * b + 1], nrm[3 * b + 2]}, eps[b], sig[b], mu[b], gain[b],
                        thick[b], f, outg);
        prev = h;
    }
    const V t{tgt[0], tgt[1], tgt[2]};
    const V fd = safe_normalize(t - prev, outg);
    total += len(t - prev);
    const double k = 2.0 * PI * f / C0;
    const cd prop = (1.0 / (2.0 * k * std::max(total, EPS))) * std::exp(cd(0.0, -k * total));
    value = value * prop;
    const cd coef = cdot(value, transverse(fd, {rxp[0], rxp[1], rxp[2]}));
    const cd pf = coef * std::sqrt(std::max(power, 0.0));
    const cd xs[3] = {value.x, value.y, value.z};
    for (int i = 0; i < 3; ++i) { out[2 * i] = xs[i].real(); out[2 * i + 1] = xs[i].imag(); }
    out[6] = coef.real(); out[7] = coef.imag(); out[8] = pf.real(); out[9] = pf.imag();
    out[10] = std::norm(pf);
    out[11] = total;
}

// ------------------------------------------------------------ path enumeration
// All valid paths of one (transmitter, receiver) pair in canonical order
// (depth, component, primitive sequence). Returns the count (> cap: truncated).
int oracle_pair_paths(const Scene* sp, const Solve* qp, OraclePath* out, int cap) {
    const Scene& s = *sp;
    const Solve& q = *qp;
    const V T{q.tx[0], q.tx[1], q.tx[2]}, R{q.rx[0], q.rx[1], q.rx[2]};
    const V tpol{q.tx_pol[0], q.tx_pol[1], q.tx_pol[2]}, rpol{q.rx_pol[0], q.rx_pol[1], q.rx_pol[2]};
    const double f = q.frequency;
    const double k = 2.0 * PI * f / C0;
    const double sqp = std::sqrt(std::max(q.tx_power, 0.0));
    std::vector<OraclePath> paths;
    auto finish = [&](OraclePath& p, const CV& v, V arrival, V departure) {
        const cd xs[3] = {v.x, v.y, v.z};
        for (int i = 0; i < 3; ++i) { p.field[2 * i] = xs[i].real(); p.field[2 * i + 1] = xs[i].imag(); }
        const cd coef = cdot(v, transverse(arrival, rpol));
        const cd pf = coef * sqp;
        p.coefficient[0] = coef.real(); p.coefficient[1] = coef.imag();
        p.path_field[0] = pf.real(); p.path_field[1] = pf.imag();
        p.gain = std::norm(pf);
        p.arrival[0] = arrival.x; p.arrival[1] = arrival.y; p.arrival[2] = arrival.z;
        p.departure[0] = departure.x; p.departure[1] = departure.y; p.departure[2] = departure.z;
    };
    auto blank = [&](int comp, int depth) {
        OraclePath p;
        std::memset(&p, 0, sizeof(p));
        p.component = comp;
        p.depth = depth;
        p.edge = -1;
        for (int i = 0; i < MAXD; ++i) { p.prim[i] = -1; p.mat[i] = -1; }
        return p;
    };

    // straight segment: ordered wall hits (one per group, lowest face id)
    const V seg = R - T;
    const double dist = len(seg);
    const V dir = safe_normalize(seg, {0, 0, 1});
    struct H { double t; int face; int group; };
    std::vector<H> hits;
    if ((q.components & 9) && dist > 2.0 * BIAS) {
        const double eps = BIAS / dist;
        for (int fc = 0; fc < s.face_count; ++fc) {
            double t;
            if (!seg_hits_tri(s, fc, T, seg, eps, 1.0 - eps, &t)) continue;
            const int g = s.face_group[fc];
            bool dup = false;
            for (auto& h : hits) if (h.group == g) { dup = true; if (fc < h.face) { h.face = fc; h.t = t; } }
            if (!dup) hits.push_back({t, fc, g});
        }
        std::sort(hits.begin(), hits.end(), [](const H& a, const H& b) { return a.t < b.t; });
    }
    const int nh = (int)hits.size();
    if ((q.components & 1) && nh == 0) {
        OraclePath p = blank(0, 0);
        p.length = dist;
        const cd carrier = (1.0 / (2.0 * std::max(k, SMALL) * std::max(dist, EPS))) * std::exp(cd(0.0, -k * dist));
        finish(p, cv_axis(transverse(dir, tpol), carrier), dir, dir);
        paths.push_back(p);
    }
    for (int d = 1; d <= q.max_depth; ++d) {
        if (q.components & 2) {
            std::vector<OraclePath> level;
            for (int c = 0; c < q.chain_count; ++c) {
                if (q.chain_depth[c] != d) continue;
                const int* gs = q.chain_groups + MAXD * c;
                // images
                V img[MAXD];
                V p = T;
                for (int j = 0; j < d; ++j) {
                    double off;
                    const V n = plane_n(s, gs[j], off);
                    p = p - n * (2.0 * (dot(p, n) - off));
                    img[j] = p;
                }
                V hit[MAXD];
                int face[MAXD];
                V next = R;
                bool ok = true;
                for (int j = d - 1; j >= 0 && ok; --j) {
                    double off;
                    const V n = plane_n(s, gs[j], off);
                    const double si = dot(n, img[j]) - off, sn = dot(n, next) - off;
                    if (!(si * sn < 0.0)) { ok = false; break; }
                    const double t = si / (si - sn);
                    hit[j] = img[j] + (next - img[j]) * t;
                    face[j] = face_containing(s, gs[j], hit[j], n);
                    if (face[j] < 0) ok = false;
                    next = hit[j];
                }
                if (!ok) continue;
                V a = T;
                int ga = -1;
                for (int j = 0; j <= d && ok; ++j) {
                    const V b = j < d ? hit[j] : R;
                    const int gb = j < d ? gs[j] : -1;
                    if (occluded(s, a, b, ga, gb)) ok = false;
                    a = b;
                    ga = gb;
                }
                if (!ok) continue;
                OraclePath pth = blank(1, d);
                for (int j = 0; j < d; ++j) {
                    pth.prim[j] = face[j];
                    pth.mat[j] = s.face_material[face[j]];
                    pth.pos[j][0] = hit[j].x; pth.pos[j][1] = hit[j].y; pth.pos[j][2] = hit[j].z;
                }
                pth.length = len(R - img[d - 1]);
                // field
                V prev = T;
                V inc = safe_normalize(hit[0] - prev, {0, 0, 1});
                const V departure = inc;
                CV value = cv_axis(transverse(inc, tpol), cd(1.0, 0.0));
                V outg = inc;
                for (int j = 0; j < d; ++j) {
                    inc = safe_normalize(hit[j] - prev, outg);
                    double off;
                    const V n = plane_n(s, gs[j], off);
                    const int m = pth.mat[j];
                    value = reflect(value, inc, n, s.mat_eps[m], s.mat_sigma[m], s.mat_mu[m], s.mat_gain[m],
                                    s.mat_thick[m], f, outg);
                    prev = hit[j];
                }
                const V arrival = safe_normalize(R - prev, outg);
                const cd prop = (1.0 / (2.0 * k * std::max(pth.length, EPS))) * std::exp(cd(0.0, -k * pth.length));
                finish(pth, value * prop, arrival, departure);
                level.push_back(pth);
            }
            std::stable_sort(level.begin(), level.end(), [d](const OraclePath& a, const OraclePath& b) {
                for (int j = 0; j < d; ++j) if (a.prim[j] != b.prim[j]) return a.prim[j] < b.prim[j];
                return false;
            });
            for (auto& p : level) paths.push_back(p);
        }
        if (d == 1 && (q.components & 4)) {
            for (int e = 0; e < s.edge_count; ++e) {
                const float* ef = s.edge_f + 14 * e;
                const int* ei = s.edge_i + 6 * e;
                const V A{ef[0], ef[1], ef[2]}, ed{ef[3], ef[4], ef[5]};
                const double elen = ef[6], n = ef[7];
                const V t0{ef[8], ef[9], ef[10]}, n0{ef[11], ef[12], ef[13]};
                const double a = dot(T - A, ed), b = dot(R - A, ed);
                const V pt = T - A - ed * a, pr = R - A - ed * b;
                const double rp = len(pt), rr = len(pr);
                if (rp < 1e-6 || rr < 1e-6) continue;
                const double ts = a + (b - a) * rp / (rp + rr);
                const V Q = A + ed * ts;
                auto ang = [&](V u) { double x = std::atan2(dot(u, n0), dot(u, t0)); if (x < 0) x += 2 * PI; return x; };
                const double npi = n * PI;
                const double phip = ang(pt);
                if (phip >= npi) continue;
                double phi = ang(pr);
                if (phi > npi) phi = (phi - npi < 2 * PI - phi) ? npi : 0.0;
                const double tc = std::min(std::max(ts, 0.0), elen);
                const V Qc = A + ed * tc;
                if (occluded(s, T, Qc, ei[2], ei[3]) || occluded(s, Qc, R, ei[2], ei[3])) continue;
                const double sp_ = len(T - Q), so_ = len(R - Q);
                const V si = (Q - T) * (1.0 / sp_), so = (R - Q) * (1.0 / so_);
                const double cb = dot(si, ed);
                const double sinb = std::sqrt(std::max(1.0 - cb * cb, 0.0));
                V phi_i = -cross(ed, si); phi_i = phi_i * (1.0 / std::max(len(phi_i), 1e-30));
                const V b0_i = cross(phi_i, si);
                V phi_o = cross(ed, so); phi_o = phi_o * (1.0 / std::max(len(phi_o), 1e-30));
                const V b0_o = cross(phi_o, so);
                const V tp = transverse(si, tpol);
                const double eb = dot(tp, b0_i), epp = dot(tp, phi_i);
                const double L = so_ * sp_ * sinb * sinb / (so_ + sp_);
                const int m0 = ei[4], mn = ei[5] >= 0 ? ei[5] : ei[4];
                cd r0s, r0h, rns, rnh;
                slab_fresnel(std::fabs(std::sin(phip)), s.mat_eps[m0], s.mat_sigma[m0], s.mat_mu[m0], s.mat_gain[m0],
                             s.mat_thick[m0], f, r0s, r0h);
                slab_fresnel(std::fabs(std::sin(npi - phi)), s.mat_eps[mn], s.mat_sigma[mn], s.mat_mu[mn],
                             s.mat_gain[mn], s.mat_thick[mn], f, rns, rnh);
                const double amp = std::sqrt(sp_ / (so_ * (so_ + sp_))) / (2.0 * k * sp_);
                const double kappa = rp * rp / (sp_ * sp_ * sp_) + rr * rr / (so_ * so_ * so_);
                const double sg = std::sqrt(k * kappa / PI);
                double c1, s1, c0, s0;
                fresnel(sg * (elen - ts), c1, s1);
                fresnel(sg * (0.0 - ts), c0, s0);
                const cd trunc = cd(0.5, 0.5) * std::conj(cd(c1 - c0, s1 - s0));
                const double u = sg * std::max(-ts, ts - elen);
                const double gamma = 1.0 / (1.0 + std::exp(u / 0.1));
                cd ds, dh;
                utd_wedge(n, phi, phip, sinb, k, L, r0s, rns, r0h, rnh, trunc, gamma, ds, dh);
                const cd prop = amp * std::exp(cd(0.0, -k * (so_ + sp_)));
                const CV v = cv_axis(b0_o, -ds * eb * prop) + cv_axis(phi_o, -dh * epp * prop);
                OraclePath p = blank(2, 1);
                p.edge = e;
                p.prim[0] = e;
                p.mat[0] = m0;
                p.pos[0][0] = Q.x; p.pos[0][1] = Q.y; p.pos[0][2] = Q.z;
                p.length = so_ + sp_;
                finish(p, v, so, si);
                if (!(p.gain >= FLOOR)) continue;
                paths.push_back(p);
            }
        }
        if ((q.components & 8) && nh == d) {
            OraclePath p = blank(5, d);
            CV value = cv_axis(transverse(dir, tpol), cd(1.0, 0.0));
            double carrier = dist;
            for (int w = 0; w < d; ++w) {
                double off;
                V n = safe_normalize(plane_n(s, hits[w].group, off), {0, 0, 1});
                if (dot(dir, n) > 0) n = -n;
                const double ct = std::min(std::max(std::fabs(dot(dir, n)), SMALL), 1.0);
                const int m = s.face_material[hits[w].face];
                const int lo = s.layer_off[m], lc = s.layer_cnt[m];
                const cd tte = stack_t(ct, s.layer_thick + lo, s.layer_eps + lo, s.layer_sigma + lo, s.layer_mu + lo, lc, f, 0);
                const cd ttm = stack_t(ct, s.layer_thick + lo, s.layer_eps + lo, s.layer_sigma + lo, s.layer_mu + lo, lc, f, 1);
                const V sa = safe_normalize(cross(n, dir), stable_perp(dir, n));
                const V pa = safe_normalize(cross(sa, dir), stable_perp(dir, sa));
                const cd es = cdot(value, sa), ep = cdot(value, pa);
                value = cv_axis(sa, tte * es) + cv_axis(pa, ttm * ep);
                double wall = 0.0;
                for (int l = 0; l < lc; ++l) wall += std::max((double)s.layer_thick[lo + l], 0.0);
                carrier -= wall * ct;
                p.prim[w] = hits[w].face;
                p.mat[w] = m;
                const V hp = T + seg * hits[w].t;
                p.pos[w][0] = hp.x; p.pos[w][1] = hp.y; p.pos[w][2] = hp.z;
            }
            p.length = dist;
            const cd prop = (1.0 / (2.0 * k * std::max(dist, EPS))) * std::exp(cd(0.0, -k * carrier));
            finish(p, value * prop, dir, dir);
            paths.push_back(p);
        }
    }
    const int count = (int)paths.size();
    for (int i = 0; i < count && i < cap; ++i) out[i] = paths[i];
    return count;
}

// Coupled paths of one (transmitter, receiver) pair in candidate order:
// (group, edge) R->D (component 3), (group, edge) D->R (4), ordered edge
// pairs D->D (7). kinds: 1 R->D, 2 D->R, 4 D->D. Returns the path count.
int oracle_coupled_paths(const Scene* sp, const Solve* qp, int kinds, OraclePath* out, int cap) {
    const Scene& s = *sp;
    const Solve& q = *qp;
    const V T{q.tx[0], q.tx[1], q.tx[2]}, X{q.rx[0], q.rx[1], q.rx[2]};
    const V tpol{q.tx_pol[0], q.tx_pol[1], q.tx_pol[2]}, rpol{q.rx_pol[0], q.rx_pol[1], q.rx_pol[2]};
    const double f = q.frequency, kw = 2.0 * PI * f / C0;
    const double sqp = std::sqrt(std::max(q.tx_power, 0.0));
    const double eps = 1e-6;
    std::vector<OraclePath> paths;
    auto emit = [&](int comp, CV v, V arrival, double length, V p0, V p1, int prim0, int edge) {
        OraclePath p;
        std::memset(&p, 0, sizeof(p));
        p.component = comp;
        p.depth = 2;
        for (int i = 0; i < MAXD; ++i) { p.prim[i] = -1; p.mat[i] = -1; }
        p.prim[0] = prim0;
        p.edge = edge;
        p.length = length;
        p.pos[0][0] = p0.x; p.pos[0][1] = p0.y; p.pos[0][2] = p0.z;
        p.pos[1][0] = p1.x; p.pos[1][1] = p1.y; p.pos[1][2] = p1.z;
        const cd xs[3] = {v.x, v.y, v.z};
        for (int i = 0; i < 3; ++i) { p.field[2 * i] = xs[i].real(); p.field[2 * i + 1] = xs[i].imag(); }
        const cd coef = cdot(v, transverse(arrival, rpol));
        const cd pf = coef * sqp;
        p.coefficient[0] = coef.real(); p.coefficient[1] = coef.imag();
        p.path_field[0] = pf.real(); p.path_field[1] = pf.imag();
        p.gain = std::norm(pf);
        p.arrival[0] = arrival.x; p.arrival[1] = arrival.y; p.arrival[2] = arrival.z;
        if (p.gain >= FLOOR) paths.push_back(p);
    };
    auto inside = [&](const EdgeRec& e, double t) { return t > eps && t < e.length - eps; };
    const int G = s.group_count, E = s.edge_count;
    for (int dir = 0; dir < 2; ++dir) {
        if (!(kinds & (1 << dir))) continue;
        for (int g = 0; g < G; ++g) {
            double off;
            const V n = plane_n(s, g, off);
            for (int ei = 0; ei < E; ++ei) {
                const EdgeRec e = edge_rec(s, ei);
                Keller k;
                if (dir == 0) {  // R->D: image transmitter
                    const V I = T - n * (2.0 * (dot(n, T) - off));
                    if (!keller(e, I, X, k) || !inside(e, k.t)) continue;
                    const double den = dot(n, k.q - I);
                    if (std::fabs(den) < 1e-12) continue;
                    const double t = (off - dot(n, I)) / den;
                    if (!(t > eps && t < 1.0 - eps)) continue;
                    const V R = I + (k.q - I) * t;
                    const int face = face_containing(s, g, R, n);
                    if (face < 0) continue;
                    if (occluded4(s, T, R, g, -1, -1, -1) || occluded4(s, R, k.q, g, e.group0, e.group1, -1) ||
                        occluded4(s, k.q, X, e.group0, e.group1, -1, -1))
                        continue;
                    const int m = s.face_material[face];
                    const V d1 = safe_normalize(R - T, {0, 0, 1});
                    V outg;
                    CV v = cv_axis(transverse(d1, tpol), cd(1.0, 0.0));
                    v = reflect(v, d1, n, s.mat_eps[m], s.mat_sigma[m], s.mat_mu[m], s.mat_gain[m], s.mat_thick[m], f,
                                outg);
                    const cd c = (1.0 / (2.0 * kw * k.sp)) * std::exp(cd(0.0, -kw * k.sp));
                    V so;
                    const CV fld = wedge(s, e, k, I, X, v * c, f, kw, so);
                    emit(3, fld, so, k.sp + k.so, R, k.q, face, ei);
                } else {  // D->R: image receiver
                    const V Xi = X - n * (2.0 * (dot(n, X) - off));
                    if (!keller(e, T, Xi, k) || !inside(e, k.t)) continue;
                    const double den = dot(n, Xi - k.q);
                    if (std::fabs(den) < 1e-12) continue;
                    const double t = (off - dot(n, k.q)) / den;
                    if (!(t > eps && t < 1.0 - eps)) continue;
                    const V R = k.q + (Xi - k.q) * t;
                    const int face = face_containing(s, g, R, n);
                    if (face < 0) continue;
                    if (occluded4(s, T, k.q, e.group0, e.group1, -1, -1) ||
                        occluded4(s, k.q, R, e.group0, e.group1, g, -1) || occluded4(s, R, X, g, -1, -1, -1))
                        continue;
                    const V si = (k.q - T) * (1.0 / k.sp);
                    const cd c = (1.0 / (2.0 * kw * k.sp)) * std::exp(cd(0.0, -kw * k.sp));
                    V so;
                    const CV at_r = wedge(s, e, k, T, Xi, cv_axis(transverse(si, tpol), c), f, kw, so);
                    const int m = s.face_material[face];
                    const V inc = safe_normalize(R - k.q, {0, 0, 1});
                    V outg;
                    const CV fld = reflect(at_r, inc, n, s.mat_eps[m], s.mat_sigma[m], s.mat_mu[m], s.mat_gain[m],
                                           s.mat_thick[m], f, outg);
                    emit(4, fld, safe_normalize(X - R, outg), k.sp + k.so, k.q, R, face, ei);
                }
            }
        }
    }
    if (kinds & 4) {
        for (int a = 0; a < E; ++a) {
            for (int b = 0; b < E; ++b) {
                if (a == b) continue;
                const EdgeRec e1 = edge_rec(s, a), e2 = edge_rec(s, b);
                if (len(cross(e1.d, e2.d)) < 1e-9 && len(cross(e2.A - e1.A, e1.d)) < 1e-6) continue;
                // edges of one face group: the middle leg would graze that face
                if ((e1.group0 >= 0 && (e1.group0 == e2.group0 || e1.group0 == e2.group1)) ||
                    (e1.group1 >= 0 && (e1.group1 == e2.group0 || e1.group1 == e2.group1)))
                    continue;
                auto q1_of = [&](double t2, V& q1, V& q2) {
                    q2 = e2.A + e2.d * t2;
                    const double ta = dot(T - e1.A, e1.d), tb = dot(q2 - e1.A, e1.d);
                    const double ra = len(T - e1.A - e1.d * ta), rb = len(q2 - e1.A - e1.d * tb);
                    const double t1 = (ra + rb) > 0 ? ta + (tb - ta) * ra / (ra + rb) : ta;
                    q1 = e1.A + e1.d * t1;
                    return t1;
                };
                auto slope = [&](double t2) {
                    V q1, q2;
                    q1_of(t2, q1, q2);
                    const V mid = q2 - q1, o = q2 - X;
                    const double lm = len(mid), lo = len(o);
                    if (!(lm > eps) || !(lo > eps)) return std::nan("");
                    return dot(e2.d, mid * (1.0 / lm) + o * (1.0 / lo));
                };
                double lo = eps, hi = e2.length - eps;
                if (!(hi > lo)) continue;
                if (!(slope(lo) < 0.0 && slope(hi) > 0.0)) continue;
                bool bad = false;
                for (int it = 0; it < 200 && hi - lo > 1e-13; ++it) {
                    const double mid = 0.5 * (lo + hi);
                    const double sm = slope(mid);
                    if (std::isnan(sm)) { bad = true; break; }
                    if (sm > 0.0) hi = mid; else lo = mid;
                }
                if (bad) continue;
                const double t2 = 0.5 * (lo + hi);
                V Q1, Q2;
                const double t1 = q1_of(t2, Q1, Q2);
                if (!inside(e1, t1)) continue;
                Keller k1, k2;
                if (!keller(e1, T, Q2, k1) || !keller(e2, Q1, X, k2)) continue;
                k2.q = Q2;
                k2.t = t2;
                k2.sp = len(Q2 - Q1);
                k2.so = len(X - Q2);
                if (occluded4(s, T, Q1, e1.group0, e1.group1, -1, -1) ||
                    occluded4(s, Q1, Q2, e1.group0, e1.group1, e2.group0, e2.group1) ||
                    occluded4(s, Q2, X, e2.group0, e2.group1, -1, -1))
                    continue;
                const V si = (k1.q - T) * (1.0 / k1.sp);
                const cd c = (1.0 / (2.0 * kw * k1.sp)) * std::exp(cd(0.0, -kw * k1.sp));
                V so;
                const CV at_q2 = wedge(s, e1, k1, T, Q2, cv_axis(transverse(si, tpol), c), f, kw, so);
                const CV fld = wedge(s, e2, k2, Q1, X, at_q2, f, kw, so);
                emit(7, fld, so, k1.sp + k2.sp + k2.so, Q1, Q2, -1, a * E + b);
            }
        }
    }
    const int count = (int)paths.size();
    for (int i = 0; i < count && i < cap; ++i) out[i] = paths[i];
    return count;
}

// Coherent / incoherent accumulation of one pair's paths (deterministic.cu:829-970):
// out = [power_total, field_re, field_im, comp_power[4], comp_re[4], comp_im[4], diff_vec[6], count]
void oracle_pair_accumulate(const OraclePath* p, int count, int coherent, double sqrt_power, double* out) {
    double sre = 0, sim = 0, pinc = 0;
    double cre[4] = {0, 0, 0, 0}, cim[4] = {0, 0, 0, 0}, cpw[4] = {0, 0, 0, 0};
    double dv[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < count; ++i) {
        const int slot = p[i].component == 0 ? 0 : p[i].component == 1 ? 1 : p[i].component == 2 ? 2 : 3;
        cre[slot] += p[i].path_field[0];
        cim[slot] += p[i].path_field[1];
        cpw[slot] += p[i].gain;
        sre += p[i].path_field[0];
        sim += p[i].path_field[1];
        pinc += p[i].gain;
        if (p[i].component == 2) for (int j = 0; j < 6; ++j) dv[j] += p[i].field[j] * sqrt_power;
    }
    if (coherent) {
        out[0] = sre * sre + sim * sim;
        out[1] = sre;
        out[2] = sim;
    } else {
        out[0] = pinc;
        out[1] = std::sqrt(pinc);
        out[2] = 0.0;
    }
    for (int sl = 0; sl < 4; ++sl) {
        out[3 + sl] = coherent ? cre[sl] * cre[sl] + cim[sl] * cim[sl] : cpw[sl];
        out[7 + sl] = cre[sl];
        out[11 + sl] = cim[sl];
    }
    for (int j = 0; j < 6; ++j) out[15 + j] = dv[j];
    out[21] = count;
}

// Transmitter-launched discovery ray i of ray_count (Fibonacci sphere),
// closest-hit specular continuation; writes the group per bounce (-1 after escape).
void oracle_discover_ray(const Scene* sp, const double* tx, int i, int ray_count, int max_depth, int* out) {
    const Scene& s = *sp;
    const double golden = PI * (3.0 - std::sqrt(5.0));
    const double z = 1.0 - (2.0 * i + 1.0) / (double)ray_count;
    const double rad = std::sqrt(std::max(0.0, 1.0 - z * z));
    const double th = golden * i;
    V d{rad * std::cos(th), rad * std::sin(th), z};
    V o{tx[0], tx[1], tx[2]};
    int ignore = -1;
    int b = 0;
    for (; b < max_depth; ++b) {
        double best = 1e300;
        int bf = -1;
        for (int f = 0; f < s.face_count; ++f) {
            if (s.face_group[f] == ignore) continue;
            double t;
            if (seg_hits_tri(s, f, o, d, 1.0e-4, best, &t)) { best = t; bf = f; }
        }
        if (bf < 0) break;
        const int g = s.face_group[bf];
        out[b] = g;
        double off;
        const V n = plane_n(s, g, off);
        o = o + d * best;
        d = d - n * (2.0 * dot(d, n));
        ignore = g;
    }
    for (; b < max_depth; ++b) out[b] = -1;
}

// Every discovery ray of one transmitter (OpenMP over rays): out [ray_count, max_depth].
void oracle_discover_all(const Scene* sp, const double* tx, int ray_count, int max_depth, int* out) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < ray_count; ++i) oracle_discover_ray(sp, tx, i, ray_count, max_depth, out + (size_t)i * max_depth);
}

// Pair topology signature used for the ambiguity probe.
static bool same_topology(const std::vector<OraclePath>& a, const std::vector<OraclePath>& b) {
    if (a.size() != b.size()) return false;
    for (size_t i = 0; i < a.size(); ++i) {
        if (a[i].component != b[i].component || a[i].depth != b[i].depth || a[i].edge != b[i].edge) return false;
        for (int j = 0; j < MAXD; ++j) if (a[i].prim[j] != b[i].prim[j]) return false;
    }
    return true;
}

// Radiomap of n receivers for one transmitter (OpenMP over receivers).
// out: [n, 22] accumulate records; paths (optional): [n, cap] path records with
// counts[n]; ambiguous (optional): 1 where the topology changes when the
// receiver moves by +-perturb along (1,1,1)/sqrt(3), i.e. some visibility or
// face decision lies within ~perturb of its boundary.
void oracle_radiomap(const Scene* sp, const Solve* base, const double* rx_pos, int n, int coherent, double* out,
                     OraclePath* paths, int cap, int* counts, double perturb, int* ambiguous) {
    const double sqp = std::sqrt(std::max(base->tx_power, 0.0));
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < n; ++i) {
        Solve q = *base;
        q.rx[0] = rx_pos[3 * i]; q.rx[1] = rx_pos[3 * i + 1]; q.rx[2] = rx_pos[3 * i + 2];
        std::vector<OraclePath> buf(1024);
        int c = oracle_pair_paths(sp, &q, buf.data(), (int)buf.size());
        if (c > (int)buf.size()) { buf.resize(c); c = oracle_pair_paths(sp, &q, buf.data(), c); }
        buf.resize(c);
        oracle_pair_accumulate(buf.data(), c, coherent, sqp, out + 22 * (size_t)i);
        if (counts) counts[i] = c;
        if (paths) for (int j = 0; j < c && j < cap; ++j) paths[(size_t)i * cap + j] = buf[j];
        if (ambiguous) {
            // perturb the receiver and the transmitter separately, both signs: a
            // decision on a leg that ends at a fixed point (a clamped edge vertex)
            // only moves with the transmitter
            int amb = 0;
            for (int probe = 0; probe < 4 && !amb; ++probe) {
                Solve qp = q;
                const double dlt = (probe % 2 ? -1.0 : 1.0) * perturb / std::sqrt(3.0);
                double* p = probe < 2 ? qp.rx : qp.tx;
                p[0] += dlt; p[1] += dlt * 0.7; p[2] += dlt * 1.3;
                std::vector<OraclePath> b2(buf.size() + 64);
                int c2 = oracle_pair_paths(sp, &qp, b2.data(), (int)b2.size());
                if (c2 > (int)b2.size()) { b2.resize(c2); c2 = oracle_pair_paths(sp, &qp, b2.data(), c2); }
                b2.resize(c2);
                if (!same_topology(buf, b2)) amb = 1;
            }
            ambiguous[i] = amb;
        }
    }
}

int oracle_path_record_bytes(void) { return (int)sizeof(OraclePath); }

int oracle_abi_version(void) { return 1; }

}  // extern "C"
