// Locally-constant-field (LCFA) QED rates and samplers, device side.
//
// Quantum synchrotron emission and nonlinear Breit-Wheeler pair creation in the
// instantaneous field seen by a particle: the total rates are table lookups in
// ln(chi); spectra are sampled by inverse CDF (emission) or rejection (pairs).
// Behaviour follows src/lcfa/photon_emission/mod.rs and
// src/lcfa/pair_creation/mod.rs (cited per function); the tables are read from
// a pointer `tb` into the LCFA pack (dev_tables.h), normally staged in LDS.
#pragma once
#include "dev_common.h"
#include "dev_tables.h"

namespace ptg {

using TP = const double*;

// ---------------------------------------------------------------- Airy / Bessel K
// Ai(x), Ai'(x) for x >= 0: Maclaurin series below 1, generalised Gauss-Laguerre
// quadrature above (Gil, Segura & Temme 2002), node count chosen by x
// (special_functions/airy.rs:41-116).
PTG_HD __forceinline__ double airy_series(TP a, double x) {
    // sum of a[k] x^n[k] over the 20 tabulated terms with n < 20; the powers
    // ascend, and x^n[k] is carried over from x^n[k-1] by the same products
    // 1 * x * x ... that a fresh power would take
    double acc = 0.0, xp = 1.0;
    int have = 0;
    for (int k = 0; k < 20; ++k) {
        int p = int(a[20 + k]);
        if (p >= 20) break;
        for (; have < p; ++have) xp *= x;
        acc += a[k] * xp;
    }
    return acc;
}
// Above x = 1 the reference writes Ai = c s^(-1/6) e^(-s) F(1/s) and
// Ai' = -c' s^(1/6) e^(-s) G(1/s), s = 2 x^(3/2) / 3, with F and G sums of
// w_k (2 + x_k / s)^(-/+1/6) over 40, 16 or 4 nodes (Ai; x < 2, < 10, >= 10) and
// 32, 16, 8 or 4 (Ai'; x < 2, < 4, < 11, >= 11). Each sum is a smooth function
// of t = 1/s on its interval; AIRY_CHEB holds its Chebyshev expansion of degree
// 15 on one or two segments (tables/gen_airy_cheb.py: truncation error below
// 1e-17 relative), so a sum costs one Clenshaw recurrence of 16 terms instead
// of a logarithm and an exponential per node. Segment record: t_lo, c, 1/h,
// a_0 .. a_15, in u = (t - c) / h.
namespace airy_cheb {
constexpr int REC = 19;
constexpr int AI_2 = 0, AI_10 = 2, AI_INF = 4;                 // Ai: first segment of x < 2, < 10, >= 10
constexpr int AP_2 = 5, AP_4 = 7, AP_11 = 8, AP_INF = 9;       // Ai': x < 2, < 4, < 11, >= 11
}  // namespace airy_cheb
PTG_HD __forceinline__ double airy_cheb_sum(TP seg, double t) {
    const double u = (t - seg[1]) * seg[2];
    const double u2 = 2.0 * u;
    double b1 = 0.0, b2 = 0.0;
#pragma unroll
    for (int k = 15; k >= 1; --k) {
        const double b0 = u2 * b1 - b2 + seg[3 + k];
        b2 = b1;
        b1 = b0;
    }
    return u * b1 - b2 + seg[3];
}
// 0.262183997088323 s^(-1/6) and -0.2898380090915846 s^(1/6) (airy.rs:53,110) with
// s^(1/6) = (2/3)^(1/6) x^(1/4), x^(1/4) by two square roots
constexpr double AIRY_C_AI = 0.28051411772306184;     // 0.262183997088323 (2/3)^(-1/6)
constexpr double AIRY_C_AIP = -0.2708986212479171;   // -0.2898380090915846 (2/3)^(1/6)
PTG_HD __forceinline__ void airy_large(TP tb, double x, bool want_ai, bool want_aip, double& ai, double& aip) {
    using namespace lcfa_pack;
    const double rx = sqrt(x);
    const double s = 2.0 * x * rx / 3.0;
    const double t = 1.0 / s;
    const double q = sqrt(rx);
    const double e = exp(-s);
    TP C = tb + AIRY_CHEB;
    if (want_ai) {
        int seg = x < 2.0 ? airy_cheb::AI_2 : (x < 10.0 ? airy_cheb::AI_10 : airy_cheb::AI_INF);
        if (seg != airy_cheb::AI_INF && t >= C[(seg + 1) * airy_cheb::REC]) ++seg;
        ai = AIRY_C_AI / q * e * airy_cheb_sum(C + seg * airy_cheb::REC, t);
    }
    if (want_aip) {
        int seg = x < 2.0 ? airy_cheb::AP_2 : (x < 4.0 ? airy_cheb::AP_4 : (x < 11.0 ? airy_cheb::AP_11 : airy_cheb::AP_INF));
        if (seg == airy_cheb::AP_2 && t >= C[(seg + 1) * airy_cheb::REC]) ++seg;
        aip = AIRY_C_AIP * q * e * airy_cheb_sum(C + seg * airy_cheb::REC, t);
    }
}
PTG_HD inline double airy_ai(TP tb, double x) {
    using namespace lcfa_pack;
    if (x < 1.0) return airy_series(tb + AIRY_SERIES, x);
    double ai, aip;
    airy_large(tb, x, true, false, ai, aip);
    return ai;
}
PTG_HD inline double airy_ai_prime(TP tb, double x) {
    using namespace lcfa_pack;
    if (x < 1.0) return airy_series(tb + AIRYP_SERIES, x);
    double ai, aip;
    airy_large(tb, x, false, true, ai, aip);
    return aip;
}
// both at one argument (the uniform expansion of J_n needs the pair)
PTG_HD inline void airy_pair(TP tb, double x, double& ai, double& aip) {
    using namespace lcfa_pack;
    if (x < 1.0) {
        ai = airy_series(tb + AIRY_SERIES, x);
        aip = airy_series(tb + AIRYP_SERIES, x);
        return;
    }
    airy_large(tb, x, true, true, ai, aip);
}
// K_{1/3}(x) = pi (2 sqrt3 / x)^{1/3} Ai((1.5x)^{2/3});  K_{2/3}(x) = -pi 3^{-1/6} (2/x)^{2/3} Ai'(...)
// (airy.rs:118-136). Non-positive arguments give 0, as the reference's None -> unwrap_or(0.0).
__device__ inline double bessel_k13(TP tb, double x) {
    if (!(x > 0.0)) return 0.0;
    double z = pow(1.5 * x, 2.0 / 3.0);
    return PI_D * cbrt(2.0 * SQRT3_D / x) * airy_ai(tb, z);
}
__device__ inline double bessel_k23(TP tb, double x) {
    if (!(x > 0.0)) return 0.0;
    double z = pow(1.5 * x, 2.0 / 3.0);
    return -PI_D * 0.8326831776556043 * pow(2.0 / x, 2.0 / 3.0) * airy_ai_prime(tb, z);  // 3^{-1/6}
}

// ---------------------------------------------------------------- emission rate
// Quantum synchrotron rate per unit time (lcfa/photon_emission/mod.rs:13-33):
// h(chi) from the small-chi expansion, the ln-ln table, or the large-chi fit.
__device__ __forceinline__ double lcfa_h(TP tb, const LcfaScalars& sc, double chi) {
    if (chi < 0.01) return 5.0 * (PI_D / 3.0) * (1.0 - 8.0 * chi / (5.0 * SQRT3_D));
    if (chi >= 100.0) {
        double c = cbrt(chi), c2 = c * c;
        double h = -1019.4661473121777 + 1786.716527650374 * c2;
        h = 1750.6263395722715 + c2 * h;
        h = -2260.1819695887225 + c * h;
        return 0.00296527643253334 * h / (chi * chi);
    }
    TP T = tb + lcfa_pack::LN_H_CHI;
    double index = (log(chi) - T[0]) / sc.delta_ln_chi_h;
    double fl = floor(index);
    double w = index - fl;
    int i = int(fl);
    return exp((1.0 - w) * T[2 * i + 1] + w * T[2 * i + 3]);
}
__device__ __forceinline__ double lcfa_rate(TP tb, const LcfaScalars& sc, double chi, double gamma) {
    return SQRT3_D * ALPHA * chi * lcfa_h(tb, sc, chi) / (2.0 * PI_D * gamma * T_COMPTON);
}
// classical synchrotron rate (lcfa/photon_emission/classical.rs:10-13)
__device__ __forceinline__ double lcfa_classical_rate(double chi, double gamma) {
    return SQRT3_D * ALPHA * chi * (5.0 * PI_D / 3.0) / (2.0 * PI_D * gamma * T_COMPTON);
}

// ---------------------------------------------------------------- emission sampling
// Inverse CDF on a table that runs from global_zero to infinity, restricted to
// y > local_zero (lcfa/photon_emission/mod.rs:35-69). cdf = [coeff, power, 31x2].
__device__ inline double from_linear_cdf(TP cdf, double global_zero, double local_zero, double rand) {
    double coeff = cdf[0], power = cdf[1];
    TP t = cdf + 2;
    double r_zero;
    if (local_zero < t[0]) r_zero = coeff * pow(local_zero - global_zero, power);
    else if (!pwmci_eval(t, 31, local_zero, false, r_zero)) return local_zero;
    double r = r_zero + (1.0 - r_zero) * rand;
    double y;
    if (r <= t[1]) y = exp((log(r) - log(coeff)) / power) + global_zero;
    else if (!pwmci_invert(t, 31, r, false, y)) y = local_zero;
    return y;
}

// classical spectrum sample (lcfa/photon_emission/classical.rs:24-60)
__device__ inline void lcfa_classical_sample(TP tb, double chi, double gamma, double r1, double r2, double r3,
                                             double& omega, bool& ok, double& theta, double& cphi) {
    TP CS = tb + lcfa_pack::CLASSICAL;
    double delta = acos((-9.0 + 50.0 * r2 - 25.0 * r2 * r2) / 16.0);
    double zb = (2.0 + 4.0 * cos(delta / 3.0)) / (5.0 * (1.0 - r2));
    double z = zb * zb * zb;
    double ln_rand = log(r1);
    double x;
    if (ln_rand < CS[1]) x = 1.020377255 * pow(r1, 0.6);
    else {
        double lx;
        if (!pwmci_invert(CS, 41, ln_rand, false, lx)) lx = CS[2 * 40];
        x = exp(lx);
    }
    double u = 3.0 * chi * x / (2.0 * z);
    omega = u * gamma;
    double cos_theta = (gamma - pow(z, 2.0 / 3.0) / (2.0 * gamma)) / sqrt(gamma * gamma - 1.0);
    ok = cos_theta >= -1.0;
    theta = cos_theta >= 1.0 ? 0.0 : (ok ? acos(cos_theta) : 0.0);
    cphi = 2.0 * PI_D * r3;
}

// Photon energy (units of mc^2) and polar/azimuthal emission angles
// (lcfa/photon_emission/mod.rs:80-231). ok = false when the polar angle would
// exceed pi (the reference's None).
__device__ inline void lcfa_sample(TP tb, const LcfaScalars& sc, double chi, double gamma, double r1, double r2, double r3,
                                   double& omega, bool& ok, double& theta, double& cphi) {
    using namespace lcfa_pack;
    double lnchi = log(chi);
    if (lnchi <= sc.ln_chi_min) {
        double om;
        lcfa_classical_sample(tb, chi, gamma, r1, r2, r3, om, ok, theta, cphi);
        omega = om * gamma / (gamma + om);
        return;
    }
    double index = (lnchi - sc.ln_chi_min) / sc.ln_chi_step;
    double fl = floor(index);
    double weight = index - fl;
    int idx = int(fl);
    double u;
    if (idx >= 40) {
        // chi beyond the table: universal low-u part + chi-dependent peak (mod.rs:94-160)
        TP HC = tb + HIGH_CHI, HP = tb + HIGH_CHI_PEAK;
        double u_bdy = (chi - 8.0) / 8.0;
        double lub = log(u_bdy);
        double p_lower;
        if (lub < HC[0]) p_lower = 1.06327715546 * cbrt(u_bdy);
        else if (lub < HC[62]) { double y; pwmci_eval(HC, 32, lub, false, y); p_lower = 1.0 - exp(-exp(y)); }
        else p_lower = 1.0 - 0.265819288864 * pow(u_bdy, -2.0 / 3.0);
        double y_bdy = u_bdy / chi - 1.0 / 9.0;
        double scale = 0.84626000931 * pow(chi, -2.0 / 3.0);
        double lyb = log(y_bdy);
        double p_higher;
        if (lyb < HP[0]) p_higher = scale * (1.0 - 7.7567692587 * y_bdy);
        else if (lyb < HP[62]) { double t; pwmci_eval(HP, 32, lyb, true, t); p_higher = scale * exp(-exp(t)); }
        else p_higher = scale * 0.46534912384 * exp(-0.45 * u_bdy / chi) * pow(y_bdy, -5.0 / 3.0);
        double r = r1 * (p_lower + p_higher);
        if (r <= p_lower) {
            double target = log(-log(1.0 - r));
            if (target < HC[1]) { double uu = r / 1.06327715546; uu = uu * uu * uu; u = uu + 1.5 * uu * uu; }
            else if (target < HC[63]) { double x; pwmci_invert(HC, 32, target, false, x); u = exp(x); }
            else u = pow(0.265819288864 / (1.0 - r), 1.5);
        } else {
            double target = (p_lower + p_higher - r) / scale;
            double y;
            if (target < HP[1]) y = (1.0 - target) / 7.7567692587;
            else { double x; pwmci_invert(HP, 32, log(-log(target)), true, x); y = exp(x); }
            u = (y + 1.0 / 9.0) * chi;
        }
    } else {
        double lr = log(r1);
        double lu[2];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            TP cdf = tb + QUANTUM_CDF + 64 * (idx + k);
            TP t = cdf + 2;
            double v;
            if (lr <= t[1]) v = (lr - log(cdf[0])) / cdf[1];
            else if (!pwmci_invert(t, 31, lr, false, v)) v = t[60];
            lu[k] = v;
        }
        u = exp((1.0 - weight) * lu[0] + weight * lu[1]);
    }
    // polar angle from the conditional CDF in y = (gamma theta)^... (mod.rs:162-226)
    double beta = 2.0 * u / (3.0 * chi);
    double delta = (1.0 + (1.0 + u) * (1.0 + u)) * pow(beta, -2.0 / 3.0) / (1.0 + u);
    double dindex = (log(delta) - sc.ln_delta_min) / sc.ln_delta_step;
    double dfl = floor(dindex);
    double dw = dindex - dfl;
    int di = int(dfl);
    double y;
    if (di >= 30) y = from_linear_cdf(tb + Y_INF, 0.0, beta, r2);
    else {
        double g0 = pow(delta, -1.5);
        double ylo = from_linear_cdf(tb + Y_CDF + 64 * di, g0, beta, r2);
        double yhi = from_linear_cdf(tb + Y_CDF + 64 * (di + 1), g0, beta, r2);
        y = (1.0 - dw) * ylo + dw * yhi;
    }
    double z = fmax(y / beta, 1.0);
    double cos_theta = (gamma - pow(z, 2.0 / 3.0) / (2.0 * gamma)) / sqrt(gamma * gamma - 1.0);
    ok = cos_theta >= -1.0;
    theta = cos_theta >= 1.0 ? 0.0 : (ok ? acos(cos_theta) : 0.0);
    omega = gamma * u / (1.0 + u);
    cphi = 2.0 * PI_D * r3;
}

// Stokes parameters of the emitted photon in the global basis (e_1 in the x-z
// plane), from K_{1/3} and K_{2/3} (lcfa/photon_emission/mod.rs:239-298 quantum,
// classical.rs:67-110 classical).
__device__ inline sv4 lcfa_stokes(TP tb, d4 k, double chi, double gamma, d3 v, d3 w, bool classical) {
    v = unit(v);
    w = unit(w);
    d3 n = unit(sp(k));
    double u = classical ? k.t / gamma : k.t / (gamma - k.t);
    double beta = asin(dot(n, unit(cross(v, w))));
    double mu = sqrt(beta * beta + 1.0 / (gamma * gamma));
    double gm = gamma * mu;
    double eta = u * gm * gm * gm / (3.0 * chi);
    double k13 = bessel_k13(tb, eta), k23 = bessel_k23(tb, eta);
    double mk = mu * k23, bk = beta * k13;
    double g, x1, x3;
    if (classical) {
        g = bk * bk + mk * mk;
        x3 = 2.0 * beta * mu * k13 * k23 / g;
    } else {
        g = 0.5 * (u * mu) * (u * mu) * (k13 * k13 + k23 * k23) / (1.0 + u) + mk * mk + bk * bk;
        x3 = 2.0 * (1.0 + 0.5 * u * u / (1.0 + u)) * beta * mu * k13 * k23 / g;
    }
    x1 = (mk * mk - bk * bk) / g;
    double phi;
    if (classical) {
        phi = asin(dot(w - dot(n, w) * n, d3{0.0, 1.0, 0.0}));
    } else {
        d3 e1{k.z, 0.0, -k.x};
        e1 = nsq(e1) == 0.0 ? d3{1.0, 0.0, 0.0} : unit(e1);
        phi = acos(fmin(fmax(dot(e1, w), -1.0), 1.0));
        phi = dot(cross(w, e1), n) < 0.0 || (dot(cross(w, e1), n) == 0.0 && signbit(dot(cross(w, e1), n))) ? -phi : phi;
    }
    double s2, c2;
    sincos(2.0 * phi, &s2, &c2);
    return {1.0, c2 * x1, -s2 * x1, x3};
}

// Gaunt factor g(chi) = P_quantum / P_classical (lcfa/photon_emission/mod.rs:330-346)
__device__ inline double gaunt_factor(TP tb, double chi) {
    if (chi <= 1.0e-3) return 1.0 - 55.0 * SQRT3_D * chi / 16.0 + 48.0 * chi * chi;
    if (chi > 500.0) {
        double c = cbrt(chi), c2 = c * c;
        double g = -1.0000341464898563 + 0.5563810015417746 * c2;
        g = 2.6235881916750565 + c2 * g;
        g = -3.8106418922532588 + c * g;
        return g / (chi * chi * chi);
    }
    double y = 0.0;
    pwmci_eval(tb + lcfa_pack::GAUNT, 25, log(chi), false, y);
    return exp(y);
}

// ---------------------------------------------------------------- pair creation
// T(chi) for photons polarised parallel / perpendicular to the acceleration
// (lcfa/pair_creation/mod.rs:13-40).
__device__ __forceinline__ void auxiliary_t(TP tb, const LcfaScalars& sc, double chi, double& par, double& perp) {
    TP T = tb + lcfa_pack::LN_T_CHI;
    if (chi <= 0.01) { par = 0.0; perp = 0.0; return; }
    if (chi < 100.0) {
        double lc = log(chi);
        int i = int((lc - T[0]) / sc.delta_ln_chi_t);
        if (chi < 1.0) {
            double x0 = exp(T[3 * i]), x1 = exp(T[3 * i + 3]);
            double dx = (chi - x0) / (x1 - x0);
            double pa = (1.0 - dx) / T[3 * i + 1] + dx / T[3 * i + 4];
            double pe = (1.0 - dx) / T[3 * i + 2] + dx / T[3 * i + 5];
            par = exp(1.0 / pa);
            perp = exp(1.0 / pe);
        } else {
            double dx = (lc - T[3 * i]) / sc.delta_ln_chi_t;
            par = exp((1.0 - dx) * T[3 * i + 1] + dx * T[3 * i + 4]);
            perp = exp((1.0 - dx) * T[3 * i + 2] + dx * T[3 * i + 5]);
        }
        return;
    }
    double c = cbrt(chi);
    par = 0.3036898468348568 / c - 2.0 / (3.0 * chi);
    perp = 0.4555347702522852 / c - 2.0 / (3.0 * chi);
}

// Stokes parameters re-expressed in the basis whose first axis is `a`
// (perpendicular to k); returns cos 2theta, sin 2theta of the rotation
// (geometry/polarization.rs:58-101).
__device__ __forceinline__ sv4 stokes_in_basis(sv4 s, d3 a, d3 k, double& c2, double& s2) {
    double k_mag = hypot(k.x, k.z);
    double a_mag = sqrt(nsq(a));
    double cos_t, det;
    if (k_mag == 0.0) { cos_t = a.x / a_mag; det = -k.y * a.z + k.z * a.y; }
    else {
        cos_t = (k.z * a.x - k.x * a.z) / (k_mag * a_mag);
        det = a.y * (k.x * k.x + k.z * k.z) - k.y * (a.x * k.x + a.z * k.z);
    }
    double sign = signbit(det) ? 1.0 : -1.0;
    cos_t = fmin(fmax(cos_t, -1.0), 1.0);
    double sin_t = sign * sqrt(1.0 - cos_t * cos_t);
    c2 = cos_t * cos_t - sin_t * sin_t;
    s2 = 2.0 * cos_t * sin_t;
    return {s.i, c2 * s.q - s2 * s.u, s2 * s.q + c2 * s.u, s.v};
}

// Breit-Wheeler probability in dt and the no-decay Stokes update
// (lcfa/pair_creation/mod.rs:48-73).
__device__ __forceinline__ double pair_probability(TP tb, const LcfaScalars& sc, double ell0, d3 ell, sv4 sv0, double chi,
                                                   d3 a_perp, double dt, sv4& sv_new) {
    double c2, s2;
    sv4 sv = stokes_in_basis(sv0, a_perp, ell, c2, s2);
    double t_par, t_perp;
    auxiliary_t(tb, sc, chi, t_par, t_perp);
    double pre = ALPHA * chi * dt / (T_COMPTON * ell0);
    double prob = pre * (0.5 * (1.0 + sv.q) * t_par + 0.5 * (1.0 - sv.q) * t_perp);
    double prob_avg = pre * 0.5 * (t_par + t_perp);
    double delta = pre * 0.5 * (t_par - t_perp);
    double inv = 1.0 / (1.0 - prob);
    double q1 = (sv.q * (1.0 - prob_avg) - delta) * inv;
    double u1 = sv.u * (1.0 - prob_avg) * inv;
    sv_new = {sv.i, c2 * q1 + s2 * u1, -s2 * q1 + c2 * u1, sv.v * (1.0 - prob_avg) * inv};
    return prob;
}

// dW/ds up to a constant (lcfa/pair_creation/mod.rs:80-89). The node-only
// factors cosh(2t/3), cosh(t/3)/cosh(t), cosh(t) of the 32-point Gauss-Laguerre
// rule are precomputed in `qn` ([32][4]: t, w, cosh t, cosh 2t/3, cosh t/3 / cosh t -> packed 5).
__device__ inline double pair_spectrum(TP qn, double s, double chi, double sv1) {
    double ss = s * (1.0 - s);
    double xi = 2.0 / (3.0 * chi * ss);
    double A = 1.0 / ss - sv1 - 2.0;
    double acc = 0.0;
    for (int k = 0; k < 32; ++k) {
        TP q = qn + 5 * k;
        acc += q[1] * exp(-xi * q[2] + q[0]) * (A * q[3] + q[4]);
    }
    return acc;
}
__device__ inline double pair_spectrum_ceiling(TP qn, double chi, double sv1) {
    double chi_switch = exp((1.5 - sv1) / SQRT3_D);
    double mx;
    if (chi < chi_switch) mx = pair_spectrum(qn, 0.5, chi, sv1);
    else if (chi > 100.0) mx = pair_spectrum(qn, 4.0 / (3.0 * chi), chi, sv1);
    else {
        double m = -0.94866 + 0.170159 * sv1;
        mx = pair_spectrum(qn, 0.5 * pow(chi_switch, -m) * pow(chi, m), chi, sv1);
    }
    return 1.05 * mx;
}
__device__ inline double pair_angular_spectrum(TP tb, double y, double s, double chi, double sv1) {
    double xi = 2.0 / (3.0 * chi * s * (1.0 - s));
    double pre = s / (1.0 - s) + (1.0 - s) / s - sv1;
    double z = 1.0 + 4.0 * chi * y * y;
    return y * (1.0 + pre * pow(z, 2.0 / 3.0)) * bessel_k13(tb, xi * z);
}
__device__ inline double pair_angular_ceiling(TP tb, double s, double chi, double sv1) {
    double y_peak = 0.216 + (0.259 - 0.216) * (1.0 - tanh(pow(8.3 / chi, 2.0 / 3.0)));
    double y = y_peak * sqrt(1.0 - (1.0 - 2.0 * s) * (1.0 - 2.0 * s));
    return 1.05 * pair_angular_spectrum(tb, y, s, chi, sv1);
}

// Positron Lorentz factor, cos(theta) and azimuth (lcfa/pair_creation/mod.rs:145-217).
// Rejection loops are capped at `cap` trials; the cap is counted in `capped`.
__device__ inline void pair_sample(TP tb, TP qn, double ell0, d3 ell, sv4 sv0, double chi, d3 a_perp, Rng& rng,
                                   double& gamma_p, double& cos_theta, double& phi, uint32_t& trials, bool& capped) {
    const int cap = 1 << 16;
    double c2, s2;
    sv4 sv = stokes_in_basis(sv0, a_perp, ell, c2, s2);
    double mx = pair_spectrum_ceiling(qn, chi, sv.q);
    double s = 0.5;
    for (int t = 0;; ++t) {
        double a, b;
        rng_uniform2(rng, a, b);
        ++trials;
        if (b <= pair_spectrum(qn, a, chi, sv.q) / mx) { s = a; break; }
        if (t == cap) { capped = true; break; }
    }
    double mx2 = pair_angular_ceiling(tb, s, chi, sv.q);
    double z = 1.0;
    for (int t = 0;; ++t) {
        double y, b;
        rng_uniform2(rng, y, b);
        ++trials;
        if (b <= pair_angular_spectrum(tb, y, s, chi, sv.q) / mx2) { z = 1.0 + 4.0 * chi * y * y; break; }
        if (t == cap) { capped = true; break; }
    }
    // azimuth (mod.rs:145-173)
    double arg = 2.0 * z / (3.0 * chi * s * (1.0 - s));
    double k_ratio = arg < 1.0e-4 ? 0.6368498843179743 / cbrt(arg)
                                  : 1.0 + 1.0 / (1.4624087952220928 * cbrt(arg) + 1.023821552056939 * sqrt(arg) + 6.0 * arg);
    double z23 = pow(z, 2.0 / 3.0), z13 = cbrt(z);
    double ssq = s * s + (1.0 - s) * (1.0 - s), s1s = s * (1.0 - s);
    double A = 1.0 + z23 * ssq / s1s, Cc = z23, D = z23 - 1.0, E = z13 * (z23 - 1.0) * ssq * k_ratio / s1s;
    auto spec = [&](double p) {
        double sn, cs, s2p, c2p;
        sincos(p, &sn, &cs);
        sincos(2.0 * p, &s2p, &c2p);
        return A + (c2p * (1.0 - Cc) - Cc) * sv.q - D * s2p * sv.u + E * sn * sv.v;
    };
    double mxa = spec(0.0);
    for (int i = 1; i < 32; ++i) mxa = fmax(mxa, spec(2.0 * PI_D * double(i) / 32.0));
    mxa *= 1.1;
    phi = 0.0;
    for (int t = 0;; ++t) {
        double a, b;
        rng_uniform2(rng, a, b);
        ++trials;
        double p = 2.0 * PI_D * a;
        if (b <= spec(p) / mxa) { phi = p; break; }
        if (t == cap) { capped = true; break; }
    }
    gamma_p = s * ell0;
    double ct = (2.0 * gamma_p * gamma_p - z23) / (2.0 * gamma_p * gamma_p - 1.0);
    cos_theta = fmax(ct, -1.0);
}

}  // namespace ptg
