// Locally-monochromatic (LMA) plane wave and linearly polarised nonlinear
// Compton scattering, device side.
//
// Step path (every step of every electron): cycle-averaged ponderomotive push
// in proper time and the total emission rate from the LP rate table
// (src/field/plane_wave.rs:68-249, src/nonlinear_compton/lp/mod.rs:430-463).
// Event path (one thread per emission, in the event kernel): harmonic from the
// CDF table, (s, theta) by rejection on the double-differential rate built
// from the generalised Bessel function J(n, x, y), photon momentum by a boost
// from the zero-momentum frame, Stokes parameters
// (src/nonlinear_compton/mod.rs:58-138, lp/mod.rs:21-676,
// special_functions/double_bessel.rs).
#pragma once
#include "dev_common.h"
#include "dev_fields.h"

namespace ptg {

// ---------------------------------------------------------------- table layout
// The LMA pack (doubles): LP rate table (68 x 61) then the LP CDF table
// (4148 x 16 x 2), with their log-grid origins and steps as scalars.
namespace lma_pack {
constexpr int RATE = 0;                       // nonlinear_compton/lp/rate_table.rs:5
constexpr int RATE_COLS = 61, RATE_ROWS = 68;
constexpr int CDF = RATE + RATE_COLS * RATE_ROWS;  // nonlinear_compton/lp/cdf_table.rs:4
constexpr int CDF_COLS = 61, CDF_N = 4148;
constexpr int CL_N = 61;                          // classical tables in ln a (lp/classical/rate_table.rs:1)
constexpr int CL_RATE = CDF + CDF_N * 32;         // [61]
constexpr int CL_CDF = CL_RATE + CL_N;            // [61][16][2] (lp/classical/cdf_table.rs:4)
// the first 15 points of every CDF row as (ln n, ln(-ln(1 - F))), the coordinates
// the inversion interpolates in from a = 5 up; filled on the device from the rows
// (launch_lp_cdf_logs) with the same functions the inversion used to apply per call
constexpr int CDF_LN = CL_CDF + CL_N * 32;         // [CDF_N][15][2]
constexpr int SIZE = CDF_LN + CDF_N * 30;
}  // namespace lma_pack

struct LmaScalars {
    double rate_min[2], rate_step[2];  // ln a, ln eta (rate_table.rs:3-4)
    double cdf_min[2], cdf_step[2];    // (cdf_table.rs:2-3)
    double cl_rate_min, cl_rate_step;  // (classical/rate_table.rs:2-3)
    double cl_cdf_min, cl_cdf_step;    // (classical/cdf_table.rs:2-3)
};

// ---------------------------------------------------------------- complex functions
PTG_HD __forceinline__ double c_arg(cplx a) { return atan2(a.im, a.re); }
PTG_HD __forceinline__ double c_abs(cplx a) { return hypot(a.re, a.im); }
PTG_HD __forceinline__ cplx c_polar(double r, double th) { double s, c; sincos(th, &s, &c); return {r * c, r * s}; }
PTG_HD __forceinline__ cplx c_log(cplx a) { return {log(c_abs(a)), c_arg(a)}; }
PTG_HD __forceinline__ cplx c_sqrt(cplx a) {
    if (a.im == 0.0) {
        if (!__builtin_signbit(a.re)) return {sqrt(a.re), a.im};
        double im = sqrt(-a.re);
        return __builtin_signbit(a.im) ? cplx{0.0, -im} : cplx{0.0, im};
    }
    if (a.re == 0.0) {
        double x = sqrt(fabs(a.im) / 2.0);
        return __builtin_signbit(a.im) ? cplx{x, -x} : cplx{x, x};
    }
    return c_polar(sqrt(c_abs(a)), c_arg(a) / 2.0);
}
// Products with i and -i written out as full complex products, (0, 1) * a and
// (-0, -1) * a, so that the signs of zero parts match num-complex: they pick the
// branch of c_arg (and so of c_acos) on the real axis.
PTG_HD __forceinline__ cplx rmul_i(cplx a) { return {0.0 * a.re - a.im, 0.0 * a.im + a.re}; }
PTG_HD __forceinline__ cplx rmul_mi(cplx a) { return {-0.0 * a.re + a.im, -0.0 * a.im - a.re}; }
PTG_HD __forceinline__ cplx c_acos(cplx z) {  // -i ln(i sqrt(1 - z^2) + z)
    cplx zz{z.re * z.re - z.im * z.im, z.re * z.im + z.im * z.re};
    cplx w = rmul_i(c_sqrt(cplx{1.0 - zz.re, 0.0 - zz.im})) + z;
    return rmul_mi(c_log(w));
}
PTG_HD __forceinline__ cplx c_sin(cplx a) {
    double s, c;
    sincos(a.re, &s, &c);
    return {s * cosh(a.im), c * sinh(a.im)};
}
PTG_HD __forceinline__ cplx c_exp(cplx a) { return c_polar(exp(a.re), a.im); }

// ---------------------------------------------------------------- plane wave (plane_wave.rs)
// <a^2> and d<a^2>/dphase for the LMA envelope (plane_wave.rs:68-150).
__device__ __forceinline__ double pw_a_sqd(const LaserDev& L, double phase) {
    double norm = L.circular ? 1.0 : 0.5;
    double nc = L.n_cycles;
    double a02 = L.a0 * L.a0;
    if (L.envelope == ENV_COS2) {
        if (fabs(phase) < PI_D * nc) {
            double s, c;
            sincos_phase(phase / (2.0 * nc), &s, &c);
            double c2 = c * c;
            return norm * a02 * c2 * c2;
        }
        return 0.0;
    }
    if (L.envelope == ENV_FLATTOP) {
        double ap = fabs(phase);
        if (ap > PI_D * (nc + 1.0)) return 0.0;
        if (ap > PI_D * (nc - 1.0)) {
            double s, c;
            sincos_phase(0.25 * (ap - (nc - 1.0) * PI_D), &s, &c);
            double c2 = c * c;
            return norm * a02 * c2 * c2;
        }
        return norm * a02;
    }
    double q = phase / (PI_D * nc);
    return norm * a02 * exp2(-q * q);
}
// d<a^2>/d(phase) * k0 (the magnitude of the time and z components of the four-gradient)
__device__ __forceinline__ double pw_grad(const LaserDev& L, double phase) {
    double norm = L.circular ? 1.0 : 0.5;
    double nc = L.n_cycles;
    double a02 = L.a0 * L.a0;
    if (L.envelope == ENV_COS2) {
        if (fabs(phase) < PI_D * nc) {
            double s, c, s2, c2;
            sincos_phase(phase / nc, &s2, &c2);
            sincos_phase(phase / (2.0 * nc), &s, &c);
            return norm * L.k0 * a02 * s2 * c * c / nc;
        }
        return 0.0;
    }
    if (L.envelope == ENV_FLATTOP) {
        double ap = fabs(phase);
        if (ap > PI_D * (nc + 1.0) || ap < PI_D * (nc - 1.0)) return 0.0;
        double s, c;
        sincos_phase(0.25 * (ap - (nc - 1.0) * PI_D), &s, &c);
        return norm * L.k0 * a02 * copysign(1.0, phase) * s * c * c * c;
    }
    double q = phase / (PI_D * nc);
    return norm * L.k0 * a02 * 2.0 * LN2_D * phase * exp2(-q * q) / ((PI_D * nc) * (PI_D * nc));
}

// ---------------------------------------------------------------- FocusedLaser (focused_laser.rs)
// Gaussian beam of waist w0 whose <a^2> is the plane wave's envelope times
// exp(-2 rho^2 / (w0^2 W)) / W, W = 1 + (z / z_R)^2, and whose four-gradient has
// x, y components besides the t = z ones (focused_laser.rs:72-146).
__device__ __forceinline__ void fl_envelope(const LaserDev& L, double phase, double& env, double& grad) {
    const double nc = L.n_cycles;
    if (L.envelope == ENV_COS2) {
        if (fabs(phase) < PI_D * nc) {
            double c = cos(phase / (2.0 * nc));
            double c2 = c * c;
            env = c2 * c2;
            grad = 2.0 * L.k0 * tan(phase / (2.0 * nc)) * env / nc;
        } else {
            env = 0.0;
            grad = 0.0;
        }
    } else if (L.envelope == ENV_FLATTOP) {
        double ap = fabs(phase);
        if (ap > PI_D * (nc + 1.0)) { env = 0.0; grad = 0.0; }
        else if (ap > PI_D * (nc - 1.0)) {
            double arg = 0.25 * (ap - (nc - 1.0) * PI_D);
            double sn, cs;
            sincos(arg, &sn, &cs);
            double c2 = cs * cs;
            env = c2 * c2;
            grad = L.k0 * copysign(1.0, phase) * sn * (c2 * cs);
        } else { env = 1.0; grad = 0.0; }
    } else {
        env = exp(-4.0 * LN2_D * (phase * phase) / L.tau_sqd);
        grad = 8.0 * LN2_D * L.k0 * phase * env / L.tau_sqd;
    }
}
__device__ __forceinline__ double fl_beam(const LaserDev& L, d4 r, double& width_sqd, double& rho_sqd) {
    double zz = r.z / L.zr;
    width_sqd = 1.0 + zz * zz;
    rho_sqd = (r.x * r.x + r.y * r.y) / L.waist_sqd;
    double norm = L.circular ? 1.0 : 0.5;
    return norm * (L.a0 * L.a0) * exp(-2.0 * rho_sqd / width_sqd) / width_sqd;
}
__device__ __forceinline__ double fl_a_sqd(const LaserDev& L, d4 r) {
    double w2, rho2;
    double beam = fl_beam(L, r, w2, rho2);
    double env, grad;
    fl_envelope(L, L.k0 * (r.t - r.z), env, grad);
    return beam * env;
}
// minus the four-gradient of <a^2>: (g_tz, g_x, g_y, g_tz)
__device__ __forceinline__ void fl_neg_grad(const LaserDev& L, d4 r, double& g_tz, double& g_x, double& g_y) {
    double w2, rho2;
    double beam = fl_beam(L, r, w2, rho2);
    double gb0 = -4.0 * beam * r.x / (L.waist_sqd * w2);
    double gb1 = -4.0 * beam * r.y / (L.waist_sqd * w2);
    double gb2 = (2.0 * beam * r.z / ((L.zr * L.zr) * w2)) * (2.0 * rho2 / w2 - 1.0);
    double env, ge;
    fl_envelope(L, L.k0 * (r.t - r.z), env, ge);
    g_tz = -(beam * ge + gb2 * env);
    g_x = -(gb0 * env);
    g_y = -(gb1 * env);
}
// <a^2> at r for either LMA field
__device__ __forceinline__ double lma_a_sqd(const LaserDev& L, d4 r) {
    return L.kind == K_FOCUSED_LASER ? fl_a_sqd(L, r) : pw_a_sqd(L, L.k0 * (r.t - r.z));
}
// The plane-wave leapfrog with the focused beam's four-gradient (focused_laser.rs:185-225).
__device__ inline double fl_push(const LaserDev& L, d4& r, d4& u, double rqm, double dt, int ll) {
    double dtau = dt / u.t;
    double sc = rqm / RQM_ELECTRON;
    double scale = sc * sc;
    const double kc = C_LIGHT * T_COMPTON;
    const double half_c = 0.5 * C_LIGHT;
    d4 u_mid{0.0, 0.0, 0.0, 0.0};
    if (ll) {
        double gtz, gx, gy;
        fl_neg_grad(L, r, gtz, gx, gy);
        double eta = kc * L.k0 * (u.t - u.z);
        double gfac = -2.0 * ALPHA * fl_a_sqd(L, r) * eta * eta / (3.0 * T_COMPTON) * scale * scale;
        double ft = half_c * scale * gtz, fx = half_c * scale * gx, fy = half_c * scale * gy;
        u_mid = {u.t + 0.5 * (ft + gfac * u.t) * dtau, u.x + 0.5 * (fx + gfac * u.x) * dtau, u.y + 0.5 * (fy + gfac * u.y) * dtau,
                 u.z + 0.5 * (ft + gfac * u.z) * dtau};
    }
    double h = 0.5 * C_LIGHT * dtau;
    r = {r.t + h * u.t, r.x + h * u.x, r.y + h * u.y, r.z + h * u.z};
    double gtz, gx, gy;
    fl_neg_grad(L, r, gtz, gx, gy);
    double ft = half_c * scale * gtz, fx = half_c * scale * gx, fy = half_c * scale * gy;
    double dwork;
    if (ll) {
        double eta = kc * L.k0 * (u_mid.t - u_mid.z);
        double asq = fl_a_sqd(L, r);
        double gfac = -2.0 * ALPHA * asq * eta * eta / (3.0 * T_COMPTON) * scale * scale;
        u = {u.t + (ft + gfac * u_mid.t) * dtau, u.x + (fx + gfac * u_mid.x) * dtau, u.y + (fy + gfac * u_mid.y) * dtau,
             u.z + (ft + gfac * u_mid.z) * dtau};
        double omega = C_LIGHT * L.k0;
        double delta = L.circular ? 1.0 : 0.75;
        dwork = (ft + 2.0 * ALPHA * omega * eta * delta * asq * asq / 3.0) * dtau;
    } else {
        u = {u.t + ft * dtau, u.x + fx * dtau, u.y + fy * dtau, u.z + ft * dtau};
        dwork = ft * dtau;
    }
    r = {r.t + h * u.t, r.x + h * u.x, r.y + h * u.y, r.z + h * u.z};
    double asq = fl_a_sqd(L, r);
    u.t = sqrt(1.0 + asq + u.x * u.x + u.y * u.y + u.z * u.z);
    return dwork;
}

// Leapfrog in proper time with the ponderomotive force, optional cycle-averaged
// Landau-Lifshitz reaction, then the mass-shell fix u^2 = 1 + <a^2>
// (plane_wave.rs:196-249). r = (ct, x, y, z), u = (u0, u). Returns the work.
__device__ __forceinline__ double pw_push(const LaserDev& L, d4& r, d4& u, double rqm, double dt, int ll) {
    double dtau = dt / u.t;
    double sc = rqm / RQM_ELECTRON;
    double scale = sc * sc;
    const double kc = C_LIGHT * T_COMPTON;
    d4 u_mid{0.0, 0.0, 0.0, 0.0};
    if (ll) {
        double ph = L.k0 * (r.t - r.z);
        double f = -0.5 * C_LIGHT * scale * pw_grad(L, ph);  // both t and z components
        double eta = kc * L.k0 * (u.t - u.z);
        double gfac = -2.0 * ALPHA * pw_a_sqd(L, ph) * eta * eta / (3.0 * T_COMPTON) * scale * scale;
        u_mid = {u.t + 0.5 * (f + gfac * u.t) * dtau, u.x + 0.5 * (gfac * u.x) * dtau, u.y + 0.5 * (gfac * u.y) * dtau,
                 u.z + 0.5 * (f + gfac * u.z) * dtau};
    }
    double h = 0.5 * C_LIGHT * dtau;
    r = {r.t + h * u.t, r.x + h * u.x, r.y + h * u.y, r.z + h * u.z};
    double ph = L.k0 * (r.t - r.z);
    double f = -0.5 * C_LIGHT * scale * pw_grad(L, ph);
    double dwork;
    if (ll) {
        double eta = kc * L.k0 * (u_mid.t - u_mid.z);
        double asq = pw_a_sqd(L, ph);
        double gfac = -2.0 * ALPHA * asq * eta * eta / (3.0 * T_COMPTON) * scale * scale;
        u = {u.t + (f + gfac * u_mid.t) * dtau, u.x + gfac * u_mid.x * dtau, u.y + gfac * u_mid.y * dtau,
             u.z + (f + gfac * u_mid.z) * dtau};
        double omega = C_LIGHT * L.k0;
        double delta = L.circular ? 1.0 : 0.75;
        dwork = (f + 2.0 * ALPHA * omega * eta * delta * asq * asq / 3.0) * dtau;
    } else {
        u = {u.t + f * dtau, u.x, u.y, u.z + f * dtau};
        dwork = f * dtau;
    }
    r = {r.t + h * u.t, r.x + h * u.x, r.y + h * u.y, r.z + h * u.z};
    double asq = pw_a_sqd(L, L.k0 * (r.t - r.z));
    u.t = sqrt(1.0 + asq + u.x * u.x + u.y * u.y + u.z * u.z);
    return dwork;
}
// ---------------------------------------------------------------- mixed precision
// The plane-wave push with the pusher algebra (envelope, its gradient, the
// radiation-reaction factor) in FP32, while the phase k.r, the position and
// momentum updates and the mass shell stay in FP64 (the phase is formed in
// FP64 and only its reduced envelope argument, |arg| <= pi/2 for cos^2, goes to
// FP32). Same steps as pw_push (plane_wave.rs:196-249).
__device__ __forceinline__ float pw_a_sqd_f(const LaserDev& L, double phase) {
    const float norm = L.circular ? 1.0f : 0.5f;
    const double nc = L.n_cycles;
    const float a02 = float(L.a0 * L.a0);
    if (L.envelope == ENV_COS2) {
        if (fabs(phase) < PI_D * nc) {
            float sn, cs;
            sincosf(float(phase / (2.0 * nc)), &sn, &cs);
            float c2 = cs * cs;
            return norm * a02 * c2 * c2;
        }
        return 0.0f;
    }
    if (L.envelope == ENV_FLATTOP) {
        double ap = fabs(phase);
        if (ap > PI_D * (nc + 1.0)) return 0.0f;
        if (ap > PI_D * (nc - 1.0)) {
            float sn, cs;
            sincosf(float(0.25 * (ap - (nc - 1.0) * PI_D)), &sn, &cs);
            float c2 = cs * cs;
            return norm * a02 * c2 * c2;
        }
        return norm * a02;
    }
    float q = float(phase / (PI_D * nc));
    return norm * a02 * exp2f(-q * q);
}
__device__ __forceinline__ float pw_grad_f(const LaserDev& L, double phase) {
    const float norm = L.circular ? 1.0f : 0.5f;
    const double nc = L.n_cycles;
    const float a02 = float(L.a0 * L.a0), k0 = float(L.k0), ncf = float(nc);
    if (L.envelope == ENV_COS2) {
        if (fabs(phase) < PI_D * nc) {
            float sn, cs, s2, c2;
            sincosf(float(phase / nc), &s2, &c2);
            sincosf(float(phase / (2.0 * nc)), &sn, &cs);
            return norm * k0 * a02 * s2 * cs * cs / ncf;
        }
        return 0.0f;
    }
    if (L.envelope == ENV_FLATTOP) {
        double ap = fabs(phase);
        if (ap > PI_D * (nc + 1.0) || ap < PI_D * (nc - 1.0)) return 0.0f;
        float sn, cs;
        sincosf(float(0.25 * (ap - (nc - 1.0) * PI_D)), &sn, &cs);
        return norm * k0 * a02 * copysignf(1.0f, float(phase)) * sn * cs * cs * cs;
    }
    float q = float(phase / (PI_D * nc));
    const float w = float(PI_D * nc);
    return norm * k0 * a02 * 2.0f * float(LN2_D) * float(phase) * exp2f(-q * q) / (w * w);
}
__device__ __forceinline__ double pw_push_mixed(const LaserDev& L, d4& r, d4& u, double rqm, double dt, int ll) {
    double dtau = dt / u.t;
    const float sc = float(rqm / RQM_ELECTRON);
    const float scale = sc * sc;
    const float kc = float(C_LIGHT * T_COMPTON * L.k0);
    const float rr = float(2.0 * ALPHA / (3.0 * T_COMPTON));
    d4 u_mid{0.0, 0.0, 0.0, 0.0};
    if (ll) {
        double ph = L.k0 * (r.t - r.z);
        float f = -0.5f * float(C_LIGHT) * scale * pw_grad_f(L, ph);
        float eta = kc * float(u.t - u.z);
        float gfac = -rr * pw_a_sqd_f(L, ph) * eta * eta * scale * scale;
        u_mid = {u.t + double(0.5f * (f + gfac * float(u.t))) * dtau, u.x + double(0.5f * gfac * float(u.x)) * dtau,
                 u.y + double(0.5f * gfac * float(u.y)) * dtau, u.z + double(0.5f * (f + gfac * float(u.z))) * dtau};
    }
    double h = 0.5 * C_LIGHT * dtau;
    r = {r.t + h * u.t, r.x + h * u.x, r.y + h * u.y, r.z + h * u.z};
    double ph = L.k0 * (r.t - r.z);
    float f = -0.5f * float(C_LIGHT) * scale * pw_grad_f(L, ph);
    double dwork;
    if (ll) {
        float eta = kc * float(u_mid.t - u_mid.z);
        float asq = pw_a_sqd_f(L, ph);
        float gfac = -rr * asq * eta * eta * scale * scale;
        u = {u.t + double(f + gfac * float(u_mid.t)) * dtau, u.x + double(gfac * float(u_mid.x)) * dtau,
             u.y + double(gfac * float(u_mid.y)) * dtau, u.z + double(f + gfac * float(u_mid.z)) * dtau};
        const float omega = float(C_LIGHT * L.k0);
        const float delta = L.circular ? 1.0f : 0.75f;
        dwork = double(f + 2.0f * float(ALPHA) * omega * eta * delta * asq * asq / 3.0f) * dtau;
    } else {
        u = {u.t + double(f) * dtau, u.x, u.y, u.z + double(f) * dtau};
        dwork = double(f) * dtau;
    }
    r = {r.t + h * u.t, r.x + h * u.x, r.y + h * u.y, r.z + h * u.z};
    double asq = double(pw_a_sqd_f(L, L.k0 * (r.t - r.z)));
    u.t = sqrt(1.0 + asq + u.x * u.x + u.y * u.y + u.z * u.z);
    return dwork;
}
// push in either LMA field (the kind is uniform over a launch)
__device__ __forceinline__ double lma_push(const LaserDev& L, d4& r, d4& u, double rqm, double dt, int ll) {
    return L.kind == K_FOCUSED_LASER ? fl_push(L, r, u, rqm, dt, ll) : pw_push(L, r, u, rqm, dt, ll);
}

// ---------------------------------------------------------------- LP rate (lp/mod.rs:430-463, lp/linear.rs:6-45)
__device__ __forceinline__ double lp_linear_rate(double a, double eta) {
    double eta2 = eta * eta;
    if (eta > 1.0e-3)
        return (a * a / 8.0) * (0.5 + 4.0 / eta - 0.5 / ((1.0 + 2.0 * eta) * (1.0 + 2.0 * eta)) +
                                (1.0 - 2.0 / eta - 2.0 / eta2) * log(1.0 + 2.0 * eta));
    return (a * a / 8.0) * (8.0 * eta / 3.0 - 16.0 * eta2 / 3.0 + 208.0 * eta2 * eta / 15.0);
}
// f(a, eta) such that dP/dphase = alpha f / eta; returns false when (a, eta)
// is outside the table (the reference then reports and uses probability 0).
__device__ __forceinline__ bool lp_rate(const double* rt, const LmaScalars& s, double a, double eta, double& out) {
    double x = log(a), y = log(eta);
    if (x < s.rate_min[0]) { out = lp_linear_rate(a, eta); return true; }
    double fx = (x - s.rate_min[0]) / s.rate_step[0];
    if (!(fx >= 0.0 && fx < double(lma_pack::RATE_COLS - 1))) return false;  // also NaN
    int ix = int(fx);
    double dx = fx - double(ix);
    if (y < s.rate_min[1]) {
        double f = (1.0 - dx) * rt[ix] + dx * rt[ix + 1];
        out = exp(f - s.rate_min[1]) * eta;
        return true;
    }
    double fy = (y - s.rate_min[1]) / s.rate_step[1];
    if (!(fy >= 0.0 && fy < double(lma_pack::RATE_ROWS - 1))) return false;
    int iy = int(fy);
    double dy = fy - double(iy);
    const double* t0 = rt + iy * lma_pack::RATE_COLS + ix;
    const double* t1 = t0 + lma_pack::RATE_COLS;
    double f = (1.0 - dx) * (1.0 - dy) * t0[0] + dx * (1.0 - dy) * t0[1] + (1.0 - dx) * dy * t1[0] + dx * dy * t1[1];
    out = exp(f);
    return true;
}

// The LP rate lookup of lp_rate on an FP32 copy of the table, in FP32; the
// linear-Compton branch below the table stays FP64 (it cancels strongly).
// FP32 table and interpolation; the cell index and the table-range decisions are
// taken in FP64, as lp_rate takes them, so that both paths agree on which
// (a, eta) are inside the table (at the table's edge an FP32 index can round past
// the last column).
__device__ __forceinline__ bool lp_rate_f(const float* rt, const LmaScalars& s, double a, double eta, float& out) {
    const double x = log(a), y = log(eta);
    if (x < s.rate_min[0]) { out = float(lp_linear_rate(a, eta)); return true; }
    const double fx = (x - s.rate_min[0]) / s.rate_step[0];
    if (!(fx >= 0.0 && fx < double(lma_pack::RATE_COLS - 1))) return false;
    const int ix = int(fx);
    const float dx = float(fx - double(ix));
    if (y < s.rate_min[1]) {
        float f = (1.0f - dx) * rt[ix] + dx * rt[ix + 1];
        out = expf(f - float(s.rate_min[1])) * float(eta);
        return true;
    }
    const double fy = (y - s.rate_min[1]) / s.rate_step[1];
    if (!(fy >= 0.0 && fy < double(lma_pack::RATE_ROWS - 1))) return false;
    const int iy = int(fy);
    const float dy = float(fy - double(iy));
    const float* t0 = rt + iy * lma_pack::RATE_COLS + ix;
    const float* t1 = t0 + lma_pack::RATE_COLS;
    out = expf((1.0f - dx) * (1.0f - dy) * t0[0] + dx * (1.0f - dy) * t0[1] + (1.0f - dx) * dy * t1[0] + dx * dy * t1[1]);
    return true;
}

// ---------------------------------------------------------------- harmonic index (lp/mod.rs:465-553)
// The k-th point of a CDF row [16][2] in the coordinates of the inversion from
// a = 5 up (ln n, ln(-ln(1 - F))).
__device__ __forceinline__ void lp_cdf_log_point(const double* t16, int k, double* out2) {
    out2[0] = log(t16[2 * k]);
    out2[1] = log(-log(1.0 - t16[2 * k + 1]));
}
// `l30` is the row's first 15 points in those coordinates (lma_pack::CDF_LN):
// formed per call they were 45 logarithms a row, four rows per emission.
template <class P>
__device__ inline double harmonic_from_table(P t16, P l30, double a, double frac) {
    if (frac <= t16[1]) return 0.9;
    double n = 0.0;
    if (a < 5.0) {
        pwmci_invert(t16, 16, frac, false, n);
        return n;
    }
    // rescaled table: ln n against ln(-ln(1 - F)), first 15 points
    pwmci_invert(l30, 15, log(-log(1.0 - frac)), true, n);
    return exp(n);
}
// Returns 0 when (a, eta) is outside the table, where the reference's lookup
// would panic; a valid index is at least 1.
// The interpolated harmonic before rounding up (lp/mod.rs:478-553): 1 at or below
// the table in a, 0 outside it.
template <class P>
__device__ inline double lp_harmonic_cont(P cdf, const LmaScalars& sc, double a, double eta, double frac) {
    double la = log(a), le = log(eta);
    if (la <= sc.cdf_min[0]) return 1.0;
    double fx = (la - sc.cdf_min[0]) / sc.cdf_step[0];
    if (!(fx >= 0.0 && fx < double(lma_pack::CDF_COLS - 1))) return 0;
    if (!(le <= sc.cdf_min[1]) && !((le - sc.cdf_min[1]) / sc.cdf_step[1] < double(lma_pack::CDF_N / lma_pack::CDF_COLS - 1)))
        return 0;
    int ix = int(fx);
    double dx = fx - double(ix);
    const P ln = cdf + (lma_pack::CDF_LN - lma_pack::CDF);  // cdf is the pack's CDF region
    auto inv = [&](int row) { return harmonic_from_table(cdf + 32 * row, ln + 30 * row, a, frac); };
    double n;
    if (le <= sc.cdf_min[1]) {
        n = inv(ix) * (1.0 - dx) + inv(ix + 1) * dx;
    } else {
        double fy = (le - sc.cdf_min[1]) / sc.cdf_step[1];
        int iy = int(fy);
        double dy = fy - double(iy);
        const int C = lma_pack::CDF_COLS;
        n = inv(C * iy + ix) * ((1.0 - dx) * (1.0 - dy)) + inv(C * iy + ix + 1) * (dx * (1.0 - dy)) +
            inv(C * (iy + 1) + ix) * ((1.0 - dx) * dy) + inv(C * (iy + 1) + ix + 1) * (dx * dy);
    }
    return n;
}
template <class P>
__device__ inline int lp_harmonic(P cdf, const LmaScalars& sc, double a, double eta, double frac) {
    double n = ceil(lp_harmonic_cont(cdf, sc, a, eta, frac));
    return (n >= 1.0 && n <= 1.0e7) ? int(n) : 0;
}
// lp_harmonic at the hand of a whole wave (all lanes call it with the same
// arguments and return the same value): lane q of every quad inverts corner q
// of the interpolation (lanes 0-3 supply the values), so the wave waits for one
// table search and root iteration instead of four in a row; the corners are
// combined as lp_harmonic_cont combines them.
__device__ inline int lp_harmonic_wave(const double* cdf, const LmaScalars& sc, double a, double eta, double frac) {
    const int q = int(__lane_id()) & 3;
    double la = log(a), le = log(eta);
    if (la <= sc.cdf_min[0]) return 1;
    double fx = (la - sc.cdf_min[0]) / sc.cdf_step[0];
    if (!(fx >= 0.0 && fx < double(lma_pack::CDF_COLS - 1))) return 0;
    if (!(le <= sc.cdf_min[1]) && !((le - sc.cdf_min[1]) / sc.cdf_step[1] < double(lma_pack::CDF_N / lma_pack::CDF_COLS - 1)))
        return 0;
    int ix = int(fx);
    double dx = fx - double(ix);
    const double* ln = cdf + (lma_pack::CDF_LN - lma_pack::CDF);
    auto inv = [&](int row) { return harmonic_from_table(cdf + 32 * row, ln + 30 * row, a, frac); };
    double n;
    if (le <= sc.cdf_min[1]) {
        const double v = inv(ix + (q & 1));
        const double v0 = __shfl(v, 0, 64), v1 = __shfl(v, 1, 64);
        n = v0 * (1.0 - dx) + v1 * dx;
    } else {
        double fy = (le - sc.cdf_min[1]) / sc.cdf_step[1];
        int iy = int(fy);
        double dy = fy - double(iy);
        const int C = lma_pack::CDF_COLS;
        const double v = inv(C * (iy + (q >> 1)) + ix + (q & 1));
        const double v0 = __shfl(v, 0, 64), v1 = __shfl(v, 1, 64), v2 = __shfl(v, 2, 64), v3 = __shfl(v, 3, 64);
        n = v0 * ((1.0 - dx) * (1.0 - dy)) + v1 * (dx * (1.0 - dy)) + v2 * ((1.0 - dx) * dy) + v3 * (dx * dy);
    }
    n = ceil(n);
    return (n >= 1.0 && n <= 1.0e7) ? int(n) : 0;
}

// ---------------------------------------------------------------- generalised Bessel function J(n, x, y)
// J(n, x, y) = sum_k J_{n+2k}(x) J_k(y) for the five orders n-2..n+2.
//
// Recurrence method (double_bessel.rs:142-246): a forward pass builds the
// continued-fraction coefficients xi (3) and lambda (2) for every order k in
// [k_min, k_min + len); a backward pass runs the two three-/two-term
// recurrences f and g, which are matched at the turning point and normalised
// by sum_k J = 1 (sign) and sum_k J^2 = 1 (magnitude). Only the forward
// coefficients are stored, in per-lane scratch `ws` with stride `st` between
// consecutive doubles of one lane (so a wave's accesses are contiguous); the
// backward pass keeps running sums instead of the reference's stored f, g, h.
// The reference's backward recurrence for J(n - 2 .. n + 2, x, y)
// (double_bessel.rs:142-246) over the orders of its workspace. The kernels
// evaluate J by quadrature (dbessel_quadrature below); the recurrence stays as
// a probe that the parity tests hold the quadrature against.
struct DBesselRec {
    int n, k_min, len;
    double* ws;   // 5 * len doubles, element (i, c) at ws[(5 * i + c) * st]
    size_t st;
};
PTG_HD __forceinline__ DBesselRec dbessel_rec_at(int n, double* ws, size_t st) {
    double xm = double(n) * 1.4142135623730951, ym = double(n) * 0.5;
    int n_minus_bound = int(floor(-2.0 * ym - xm));
    int n_plus_bound = xm > 4.0 * (2.0 - 1.4142135623730951) * ym ? int(ceil(xm)) : int(ceil(2.0 * ym + xm * xm / (16.0 * ym)));
    return DBesselRec{n, n_minus_bound - 100, n_plus_bound - n_minus_bound + 200, ws, st};
}
PTG_HD __forceinline__ int dbessel_len(int n) {
    double xm = double(n) * 1.4142135623730951, ym = double(n) * 0.5;
    int lo = int(floor(-2.0 * ym - xm));
    int hi = xm > 4.0 * (2.0 - 1.4142135623730951) * ym ? int(ceil(xm)) : int(ceil(2.0 * ym + xm * xm / (16.0 * ym)));
    return hi - lo + 200;
}

// One order of the forward pass: the coefficients of order k from those of k - 1.
PTG_HD __forceinline__ void dbessel_forward(double x, double y, int k, double& x0, double& x1, double& x2, double& l0,
                                            double& l1) {
    double n0 = -x - 4.0 * y * y / x2;
    double n1 = 2.0 * double(k - 1) - 2.0 * y * x0 / x2;
    double n2 = -x - 2.0 * y * x1 / x2;
    double m0 = n0 - 2.0 * y * n2 / l1;
    double m1 = n1 - l0 * n2 / l1;
    x0 = n0; x1 = n1; x2 = n2; l0 = m0; l1 = m1;
}
// The backward pass over the forward coefficients W(i, c), read in descending order of i.
template <class DB, class WF>
PTG_HD __forceinline__ void dbessel_backward(const DB& D, double x, double y, WF& W, double j[5]) {
    const int len = D.len;
    int n_minus = int(floor(-2.0 * y - x));
    int n_plus = 8.0 * y > x ? int(ceil(2.0 * y + x * x / (16.0 * y))) : int(ceil(-2.0 * y + x));
    int i_match = (n_plus + n_minus) / 2 - D.k_min;
    int i_target = D.n - D.k_min;
    double initial = D.n < 2000 ? 1.0e-100 : 1.0e-300;
    // backward pass; f[i] from f[i+1..i+3], g[i] from g[i+1..i+2]
    double f3 = 0.0, f2 = 1.0 * initial, f1 = 2.0 * initial;  // f[len-1], f[len-2], f[len-3]
    double g2 = 0.0, g1 = 1.0 / initial;                       // g[len-1], g[len-2]
    double g0 = -(2.0 * y * g2 + W(len - 2, 3) * g1) / W(len - 2, 4);  // g[len-3]
    // tail values (i >= i_match) keep f; head values (i < i_match) keep g.
    // Sums over the tail are accumulated in f units and rescaled with f.
    double ft_sum = 0.0, ft_rss = 0.0;                  // sum f, sqrt(sum f^2) over i >= i_match
    double gh_sum = 0.0, gh_rss = 0.0;                  // sum g, sqrt(sum g^2) over i < i_match
    double tgt[5] = {0, 0, 0, 0, 0};
    bool tgt_is_f[5] = {false, false, false, false, false};
    double f_match = 0.0, g_match = 0.0;
    bool have_match = false;
    // the reference zeroes f[i_match] and g[i_match] before the pass; they are
    // then recomputed in order, so the sentinels only matter for the rescaling test
    auto visit_tail = [&](int i, double fv) {
        ft_sum += fv;
        ft_rss = hypot(ft_rss, fv);
        int q = i - (i_target - 2);
        if (q >= 0 && q < 5) { tgt[q] = fv; tgt_is_f[q] = true; }
    };
    // elements len-1, len-2, len-3 (f known), all >= i_match
    visit_tail(len - 1, f3);
    visit_tail(len - 2, f2);
    visit_tail(len - 3, f1);
    // g values at len-1.. are only used for i < i_match, never here
    double gp2 = g1, gp1 = g0;  // g[i+2], g[i+1] as i descends from len-4
    for (int i = len - 4; i >= 0; --i) {
        double fi = -(2.0 * y * f3 + W(i + 2, 0) * f2 + W(i + 2, 1) * f1) / W(i + 2, 2);
        double gi = -(2.0 * y * gp2 + W(i + 1, 3) * gp1) / W(i + 1, 4);
        f3 = f2; f2 = f1; f1 = fi;
        gp2 = gp1; gp1 = gi;
        if (i >= i_match) {
            if (i == i_match) { f_match = fi; g_match = gi; have_match = true; }
            visit_tail(i, fi);
        } else {
            gh_sum += gi;
            gh_rss = hypot(gh_rss, gi);
            int q = i - (i_target - 2);
            if (q >= 0 && q < 5) { tgt[q] = gi; tgt_is_f[q] = false; }
        }
        // rescaling (double_bessel.rs:199-204): f of elements >= i by 1e-100, g by 1e100
        double gm = have_match ? g_match : 0.0;
        if (fi > 1.0 && fabs(gm) < 1.0e200) {
            f3 *= 1.0e-100; f2 *= 1.0e-100; f1 *= 1.0e-100;
            gp2 *= 1.0e100; gp1 *= 1.0e100;
            ft_sum *= 1.0e-100;
            ft_rss *= 1.0e-100;
            if (have_match) { f_match *= 1.0e-100; g_match *= 1.0e100; }
            for (int q = 0; q < 5; ++q)
                if (tgt_is_f[q]) tgt[q] *= 1.0e-100;
            if (i < i_match) {  // head g values computed so far are also scaled in the reference
                gh_sum *= 1.0e100;
                gh_rss *= 1.0e100;
                for (int q = 0; q < 5; ++q)
                    if (!tgt_is_f[q] && (q + i_target - 2) >= i) tgt[q] *= 1.0e100;
            }
        }
    }
    // h = g for i < i_match, (g_match f) / f_match for i >= i_match, in that order:
    // g_match / f_match alone can overflow (e.g. 1e207 / 1e-191) where the
    // products stay finite
    double sum_linear = gh_sum + (g_match * ft_sum) / f_match;
    double rss = hypot(gh_rss, fabs(g_match * ft_rss) / fabs(f_match));
    for (int q = 0; q < 5; ++q) {
        double h = tgt_is_f[q] ? (g_match * tgt[q]) / f_match : tgt[q];
        double val = copysign(1.0, h / sum_linear) * fabs(h) / rss;
        j[q] = __builtin_isfinite(val) ? val : 0.0;
    }
}
PTG_HD inline void dbessel_recurrence(const DBesselRec& D, double x, double y, double j[5]) {
    const int len = D.len;
    const size_t st = D.st;
    double* ws = D.ws;
    auto S = [&](int i, int c) -> double& { return ws[(size_t(5) * i + c) * st]; };
    // forward pass (element 0 holds the initial value 1 for all coefficients)
    double x0 = 1.0, x1 = 1.0, x2 = 1.0, l0 = 1.0, l1 = 1.0;
    S(0, 0) = 1.0; S(0, 1) = 1.0; S(0, 2) = 1.0; S(0, 3) = 1.0; S(0, 4) = 1.0;
    for (int i = 1; i < len; ++i) {
        dbessel_forward(x, y, i + D.k_min, x0, x1, x2, l0, l1);
        S(i, 0) = x0; S(i, 1) = x1; S(i, 2) = x2; S(i, 3) = l0; S(i, 4) = l1;
    }
    auto W = [&](int i, int c) -> double { return ws[(size_t(5) * i + c) * st]; };
    dbessel_backward(D, x, y, W, j);
}

// Saddle-point approximation of J(n, x, y) for large n (double_bessel.rs:272-355).
PTG_HD inline double dbessel_spa(int ni, double x, double y) {
    double n = double(ni);
    if (n <= -2.0 * y + x) {
        double cos_theta = y > 0.0 ? x / (8.0 * y) - sqrt(x * x / (64.0 * y * y) + 0.5 - n / (4.0 * y)) : n / x;
        double theta = acos(cos_theta);
        double st, ct;
        sincos(theta, &st, &ct);
        double s2 = 2.0 * st * ct;
        double f = x * st - y * s2;
        double d2f = -x * st + 4.0 * y * s2;
        double phi = (PI_D / 4.0) * copysign(1.0, d2f);
        return cos(f - n * theta + phi) * sqrt(2.0 / (PI_D * fabs(d2f)));
    } else if (8.0 * y < x || n > 2.0 * y + x * x / (16.0 * y)) {
        double disc = x * x / (64.0 * y * y) + 0.5 - n / (4.0 * y);
        cplx cos_theta;
        if (__builtin_isnan(disc)) cos_theta = {n / x, 0.0};
        else if (disc < 0.0) cos_theta = {x / (8.0 * y), sqrt(-disc)};
        else cos_theta = {x / (8.0 * y) - sqrt(disc), 0.0};
        cplx theta = c_acos(cos_theta);
        cplx s1 = c_sin(theta), s2 = c_sin(2.0 * theta);
        cplx f = rmul_i(cplx{-n * theta.re + x * s1.re - y * s2.re, -n * theta.im + x * s1.im - y * s2.im});
        cplx d2f = rmul_i(cplx{-x * s1.re + 4.0 * y * s2.re, -x * s1.im + 4.0 * y * s2.im});
        cplx phase = f + cplx{0.0, (PI_D - c_arg(d2f)) / 2.0};
        double amp = sqrt(2.0 / (PI_D * c_abs(d2f)));
        double re;
        if (f.im == 0.0 && d2f.im == 0.0) re = 0.5 * exp(-f.re) * amp;
        else re = c_exp(phase).re * amp;
        return disc != 0.0 ? re : 0.0;
    } else {
        double r = sqrt(x * x / (64.0 * y * y) + 0.5 - n / (4.0 * y));
        double out = 0.0;
        for (int k = 0; k < 2; ++k) {
            double theta = acos(x / (8.0 * y) + (k ? r : -r));
            double st, ct;
            sincos(theta, &st, &ct);
            double s2 = 2.0 * st * ct;
            double f = x * st - y * s2;
            double d2f = -x * st + 4.0 * y * s2;
            double phi = (PI_D / 4.0) * copysign(1.0, d2f);
            out += cos(f - n * theta + phi) * sqrt(2.0 / (PI_D * fabs(d2f)));
        }
        return out;
    }
}
PTG_HD inline bool dbessel_use_spa(int n, double x, double y) {
    if (n <= 10) return false;
    double ne = double(n);
    ne = ne * (1.0 - sqrt(10.0 / ne));
    double x_crit = y > ne / 6.0 ? 4.0 * sqrt(y * (ne - 2.0 * y)) : ne + 2.0 * y;
    return x < x_crit;
}
// J(n - 2 .. n + 2, x, y) by the trapezoid rule. J(m, x, y) = (1/pi) times the
// integral over [0, pi] of cos(m phi - x sin phi + y sin 2 phi) (the integrand is
// even and 2 pi periodic), and the trapezoid rule on K intervals of [0, pi] is
// the 2K-point rule over the full period: its error is the integrand's Fourier
// content beyond order 2K, which ends near |m| + x + 2y and decays within a few
// multiples of |x|^(1/3) and |y|^(1/3) past it (x < 0 is the angle theta in the
// second or third quadrant: J(m, -x, y) = (-1)^m J(m, x, y), which the integral
// gives by itself). The node count below keeps the error under ~2e-13 absolute
// over |x| <= n sqrt 2, 0 <= y <= n / 2; the
// rounding floor is ~1e-15 absolute, so a J far below that (deep in its
// evanescent region, where it feeds no rate) is not resolved relatively, as it
// is by the recurrence. The nodes phi_k = k pi / K enter through e^(i phi) and
// e^(i n phi), advanced by rotation and recomputed every 32 nodes; each node
// costs one sin/cos pair of the bounded phase y sin 2 phi - x sin phi.
__device__ __forceinline__ int dbessel_quad_nodes(int n, double x, double y) {
    const double ax = fabs(x), ay = fabs(y);
    return int(ceil(0.5 * (double(n) + 12.0 + ax + 2.0 * ay + 10.0 * (cbrt(ax) + 2.0 * cbrt(ay)))));
}
__device__ inline void dbessel_quadrature(int n, double x, double y, double j[5]) {
    const int K = dbessel_quad_nodes(n, x, y);
    const long long two_k = 2LL * K;
    const long long nm = ((long long)n % two_k + two_k) % two_k;  // n phi_1 = pi nm / K (mod 2 pi)
    const double inv_k = 1.0 / double(K);
    double sh, ch, snh, cnh;
    sincos_phase(PI_D * inv_k, &sh, &ch);
    sincos_phase(PI_D * double(nm) * inv_k, &snh, &cnh);
    double cp = 1.0, sp = 0.0, cq = 1.0, sq = 0.0;  // e^(i phi_k), e^(i n phi_k)
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0, a4 = 0.0;
    for (int k = 0; k <= K; ++k) {
        if (k > 0 && (k & 31) == 0) {
            sincos_phase(PI_D * double(k) * inv_k, &sp, &cp);
            sincos_phase(PI_D * double((nm * k) % two_k) * inv_k, &sq, &cq);
        }
        const double s2 = 2.0 * sp * cp, c2 = (cp - sp) * (cp + sp);
        double sv, cv;
        sincos_phase(y * s2 - x * sp, &sv, &cv);
        const double w = (k == 0 || k == K) ? 0.5 : 1.0;
        const double cr = w * (cq * cv - sq * sv), sr = w * (sq * cv + cq * sv);  // w e^(i psi_k)
        a0 += cr * c2 + sr * s2;  // cos(psi - 2 phi)
        a1 += cr * cp + sr * sp;  // cos(psi - phi)
        a2 += cr;
        a3 += cr * cp - sr * sp;  // cos(psi + phi)
        a4 += cr * c2 - sr * s2;  // cos(psi + 2 phi)
        const double cp1 = cp * ch - sp * sh;
        sp = sp * ch + cp * sh;
        cp = cp1;
        const double cq1 = cq * cnh - sq * snh;
        sq = sq * cnh + cq * snh;
        cq = cq1;
    }
    j[0] = a0 * inv_k; j[1] = a1 * inv_k; j[2] = a2 * inv_k; j[3] = a3 * inv_k; j[4] = a4 * inv_k;
}

// The same quadrature shared by the 64 lanes of a wave, which all call it with
// the same (n, x, y): lane l sums the contiguous nodes [l C, (l + 1) C), C =
// ceil((K + 1) / 64), starting its rotations from a direct evaluation and
// re-anchoring every 32 nodes, and a butterfly over the wave adds the partial
// sums. Each butterfly step adds the same two values in every lane of a pair,
// so every lane returns the same bits.
__device__ inline void dbessel_quadrature_wave(int n, double x, double y, double j[5]) {
    const int lane = int(__lane_id());
    const int K = dbessel_quad_nodes(n, x, y);
    const int C = (K + 64) / 64;
    const int k_lo = lane * C, k_hi = min(k_lo + C, K + 1);
    const long long two_k = 2LL * K;
    const long long nm = ((long long)n % two_k + two_k) % two_k;
    const double inv_k = 1.0 / double(K);
    double sh, ch, snh, cnh;
    sincos_phase(PI_D * inv_k, &sh, &ch);
    sincos_phase(PI_D * double(nm) * inv_k, &snh, &cnh);
    double cp = 1.0, sp = 0.0, cq = 1.0, sq = 0.0;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0, a4 = 0.0;
    for (int k = k_lo; k < k_hi; ++k) {
        if (((k - k_lo) & 31) == 0) {
            sincos_phase(PI_D * double(k) * inv_k, &sp, &cp);
            sincos_phase(PI_D * double((nm * k) % two_k) * inv_k, &sq, &cq);
        }
        const double s2 = 2.0 * sp * cp, c2 = (cp - sp) * (cp + sp);
        double sv, cv;
        sincos_phase(y * s2 - x * sp, &sv, &cv);
        const double w = (k == 0 || k == K) ? 0.5 : 1.0;
        const double cr = w * (cq * cv - sq * sv), sr = w * (sq * cv + cq * sv);
        a0 += cr * c2 + sr * s2;
        a1 += cr * cp + sr * sp;
        a2 += cr;
        a3 += cr * cp - sr * sp;
        a4 += cr * c2 - sr * s2;
        const double cp1 = cp * ch - sp * sh;
        sp = sp * ch + cp * sh;
        cp = cp1;
        const double cq1 = cq * cnh - sq * snh;
        sq = sq * cnh + cq * snh;
        cq = cq1;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a0 += __shfl_xor(a0, off, 64);
        a1 += __shfl_xor(a1, off, 64);
        a2 += __shfl_xor(a2, off, 64);
        a3 += __shfl_xor(a3, off, 64);
        a4 += __shfl_xor(a4, off, 64);
    }
    j[0] = a0 * inv_k; j[1] = a1 * inv_k; j[2] = a2 * inv_k; j[3] = a3 * inv_k; j[4] = a4 * inv_k;
}

// The reference's evaluation (double_bessel.rs:22-37): the saddle point where
// dbessel_use_spa holds, else the recurrence (probe of the parity tests).
PTG_HD inline void dbessel_eval(const DBesselRec& D, double x, double y, double j[5]) {
    if (dbessel_use_spa(D.n, x, y))
        for (int q = 0; q < 5; ++q) j[q] = dbessel_spa(D.n - 2 + q, x, y);
    else
        dbessel_recurrence(D, x, y, j);
}
// The kernels' evaluation at harmonic n: the same saddle point, and the
// quadrature where the reference recurs.
struct DBessel { int n; };
__device__ __forceinline__ DBessel dbessel_of(int n) { return DBessel{n}; }
__device__ inline void dbessel_eval(const DBessel& D, double x, double y, double j[5]) {
    if (dbessel_use_spa(D.n, x, y))
        for (int q = 0; q < 5; ++q) j[q] = dbessel_spa(D.n - 2 + q, x, y);
    else
        dbessel_quadrature(D.n, x, y, j);
}
// The same at the hand of a whole wave (all lanes, same arguments): the
// quadrature is shared out over the lanes (dbessel_quadrature_wave).
struct DBesselWave { int n; };
__device__ inline void dbessel_eval(const DBesselWave& D, double x, double y, double j[5]) {
    if (dbessel_use_spa(D.n, x, y))
        for (int q = 0; q < 5; ++q) j[q] = dbessel_spa(D.n - 2 + q, x, y);
    else
        dbessel_quadrature_wave(D.n, x, y, j);
}

// ---------------------------------------------------------------- LP double-differential rate (lp/mod.rs:21-140)
template <class DB>
__device__ inline double lp_ddpr(const DB& D, double a, double eta, double s, double theta) {
    int n = D.n;
    double sn = 2.0 * double(n) * eta / (1.0 + 0.5 * a * a);
    double wn = fmin(s / (sn * (1.0 - s)), 1.0);
    double alpha = 2.0 * double(n) * a * cos(theta) * sqrt(wn * (1.0 - wn)) / sqrt(1.0 + 0.5 * a * a);
    double beta = a * a * s / (8.0 * eta * (1.0 - s));
    double j[5];
    dbessel_eval(D, alpha, beta, j);
    double g0, g1, g2;
    if (n % 2 == 0) { g0 = j[2]; g1 = -0.5 * (j[1] + j[3]); g2 = 0.25 * (j[0] + 2.0 * j[2] + j[4]); }
    else { g0 = -j[2]; g1 = 0.5 * (j[1] + j[3]); g2 = -0.25 * (j[0] + 2.0 * j[2] + j[4]); }
    return (-g0 * g0 - a * a * (1.0 + 0.5 * s * s / (1.0 - s)) * (g0 * g2 - g1 * g1)) / (2.0 * PI_D);
}
__device__ inline int lp_sum_limit(double a, double eta) {
    double m = 3.0 - 0.4 * pow(eta, 0.25);
    return int(ceil(6.0 * (1.0 + a * a) + 2.0 * pow(a, m)));
}
// Tabulated rejection ceiling of LP emission at high harmonics. The reference
// computes, for every emission, the largest of 32-64 samples of the
// double-differential rate (lp/mod.rs:53-140); from harmonic 200 up each sample
// is a J(n, x, y) quadrature of ~n nodes, and the samples are ~2/3 of an
// event's work. That largest sample varies smoothly with (a, eta, n) within the
// reference's regimes, so a table of it over ln a, ln eta and ln n (generated
// by the library's own sampling code, tables/gen_lp_ceiling.py) gives the
// ceiling by trilinear interpolation of its logarithm, times a margin kept per
// cell: the largest ratio of the sample to the interpolated value at eight
// interior points of the cell (x 1.02, at least 1). Cells that needed more than
// 1.25 (where the reference changes its sample points, and near the harmonic
// cutoff) hold 0 and leave their events to the direct ceiling. Nodes: a = 3 ...
// 20, eta = 1e-3 ... 2, n = 200 ... 40000 (rounded to integers), all
// log-spaced; nodes <= 0 are not tabulated (harmonics beyond twice the sum
// limit). Layout: the nodes [NA][NE][NN], then the cell margins
// [NA - 1][NE - 1][NN - 1].
namespace lp_ceil {
constexpr int NA = 33, NE = 44, NN = 110;
constexpr double LN_A0 = 1.0986122886681098;                        // ln 3
constexpr double LN_A_STEP = 1.8971199848858813 / double(NA - 1);   // ln (20 / 3)
constexpr double LN_E0 = -6.907755278982137;                        // ln 1e-3
constexpr double LN_E_STEP = 7.600902459542082 / double(NE - 1);    // ln 2000
constexpr double LN_N0 = 5.298317366548036;                         // ln 200
constexpr double LN_N_STEP = 5.298317366548036 / double(NN - 1);    // ln 200
constexpr int NODES = NA * NE * NN;
constexpr int SIZE = NODES + (NA - 1) * (NE - 1) * (NN - 1);
}  // namespace lp_ceil

// The tabulated ceiling sample (interpolated largest sample x the cell's
// margin) at (n, a, eta); false outside the table or where the cell is not
// tabulated.
__device__ inline bool lp_ceiling_lookup(const double* T, int n, double a, double eta, double& m) {
    using namespace lp_ceil;
    const double fa = (log(a) - LN_A0) / LN_A_STEP, fe = (log(eta) - LN_E0) / LN_E_STEP;
    const double fn = (log(double(n)) - LN_N0) / LN_N_STEP;
    if (!(fa >= 0.0 && fa < double(NA - 1) && fe >= 0.0 && fe < double(NE - 1) && fn >= 0.0 && fn < double(NN - 1)))
        return false;
    const int ia = int(fa), ie = int(fe), in = int(fn);
    const double mu = T[NODES + (ia * (NE - 1) + ie) * (NN - 1) + in];
    if (!(mu > 0.0)) return false;
    const double da = fa - double(ia), de = fe - double(ie), dn = fn - double(in);
    double acc = 0.0;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int xa = c & 1, xe = (c >> 1) & 1, xn = c >> 2;
        const double v = T[((ia + xa) * NE + ie + xe) * NN + in + xn];
        if (!(v > 0.0)) return false;
        acc += (xa ? da : 1.0 - da) * (xe ? de : 1.0 - de) * (xn ? dn : 1.0 - dn) * log(v);
    }
    m = mu * exp(acc);
    return true;
}

// The points at which lp_ceiling samples the double-differential rate
// (lp/mod.rs:53-140): how many, and the k-th (s, theta). Split out so that a
// wave can evaluate them in parallel.
__device__ inline int lp_ceiling_count(int n, double a, double eta) {
    if (n == 1) return 2;
    if (a <= 2.0) return 32;
    if (a <= 5.0) {
        int n_switch = int(ceil(1.3 * pow(a, 1.2)));
        if (n < n_switch) return 33;
        return n < lp_sum_limit(a, eta) / 2 ? 32 : 64;
    }
    int n_switch = int(ceil(a * a));
    if (n < n_switch) return 64;
    return n > 10 * n_switch ? 64 : (n > 3 * n_switch ? 42 : 32);
}
__device__ inline void lp_ceiling_arg(int n, double a, double eta, int k, double& s, double& theta) {
    double sn = 2.0 * double(n) * eta / (1.0 + 0.5 * a * a);
    double smax = sn / (1.0 + sn);
    theta = 0.0;
    if (n == 1) {
        s = (k == 0 ? 0.01 : 0.99) * smax;
        theta = PI_D / 2.0;
    } else if (a <= 2.0) {
        s = smax * (0.4 + 0.6 * double(k + 1) / 32.0);
    } else if (a <= 5.0) {
        int n_switch = int(ceil(1.3 * pow(a, 1.2)));
        if (n < n_switch) {
            theta = (PI_D / 2.0) * (1.0 - log(double(n)) / log(double(n_switch)));
            s = smax * double(k) / 32.0;
        } else {
            double x_min = fmax(4.0 * double(n) * eta / (3.0 * a * a + 4.0 * double(n) * eta), 0.4);
            int m = n < lp_sum_limit(a, eta) / 2 ? 32 : 64;
            s = smax * (x_min + (1.0 - x_min) * double(k + 1) / double(m));
        }
    } else {
        int n_switch = int(ceil(a * a));
        if (n < n_switch) {
            double ns = double(n_switch);
            theta = (PI_D / 2.0) * (pow(double(n), -0.7) - pow(ns, -0.7)) / (1.0 - pow(ns, -0.7));
            s = smax * (0.3 + 0.7 * double(k + 1) / 64.0);
        } else {
            double x_min = fmax(4.0 * double(n) * eta / (3.0 * a * a + 4.0 * double(n) * eta), 0.3);
            double x_peak = (1.0 + sn) / (2.0 + sn);
            int len = n > 10 * n_switch ? 64 : (n > 3 * n_switch ? 42 : 32);
            s = smax * (x_min + (x_peak - x_min) * double(k + 1) / double(len));
        }
    }
}

// Largest theta contributing at each s (lp/mod.rs:143-301).
PTG_HD inline double ln_dbessel_x_zero(int ni, double y) {
    double n = double(ni);
    double z = sqrt(n / (4.0 * y) - 0.5);
    cplx theta{PI_D / 2.0, log(sqrt(1.0 + z * z) - z)};
    cplx s2 = c_sin(2.0 * theta);
    cplx f = rmul_i(cplx{-n * theta.re - y * s2.re, -n * theta.im - y * s2.im});
    cplx f2 = rmul_i(4.0 * y * s2);
    double e = 0.5 * (PI_D - c_arg(f2));
    cplx phase{f.re, f.im + e};
    return log(sqrt(2.0 / (PI_D * c_abs(f2)))) + phase.re + log(fabs(cos(phase.im)));
}
// Point i of 16 of the bound (before the running minimum): s_i and theta_i.
__device__ inline void theta_bound_point(int n, double a, double eta, int i, double& s_i, double& f_i) {
    double sn = 2.0 * double(n) * eta / (1.0 + 0.5 * a * a);
    double ln_j_crit = log(0.443) - log(double(n)) / 3.0;
    double z = cos(PI_D * double(i) / 15.0);
    double s = 0.5 * (z + 1.0) * sn / (1.0 + sn);
    s_i = s;
    double wn = fmin(s / (sn * (1.0 - s)), 1.0);
    double x = 2.0 * double(n) * a * sqrt(wn * (1.0 - wn)) / sqrt(1.0 + 0.5 * a * a);
    double y = a * a * s / (8.0 * eta * (1.0 - s));
    double x_crit = y > double(n) / 6.0 ? 4.0 * sqrt(y) * sqrt(double(n) - 2.0 * y) : double(n) + 2.0 * y;
    double ln_j_bdy = fmin(ln_dbessel_x_zero(n % 2 == 0 ? n : n + 1, y), ln_j_crit);
    double cos_theta = fmax(1.0 - x_crit * log(0.01) / (x * (ln_j_bdy - ln_j_crit)), -1.0);
    double f = acos(cos_theta);
    f_i = isnan(f) ? PI_D : f;
}
// the bound is non-increasing in s: a running minimum from the largest s down
__device__ __forceinline__ void theta_bound_running_min(double* f) {
    double f_min = f[0];
    for (int i = 1; i < 16; ++i) {
        if (f[i] < f_min) f_min = f[i];
        else f[i] = f_min;
    }
}
// the bound at s (or v), interpolated between its 16 points (held in LDS, one copy per wave)
__device__ __forceinline__ double theta_bound_at(const double* ts, const double* tf, double arg) {
    double val = tf[15];
    for (int i = 1; i < 16; ++i) {
        if (arg > ts[i]) {
            double w = (arg - ts[i - 1]) / (ts[i] - ts[i - 1]);
            val = w * tf[i] + (1.0 - w) * tf[i - 1];
            break;
        }
    }
    return fmin(val, PI_D / 2.0);
}

// ---------------------------------------------------------------- LP, classical limit (lp/classical.rs)
// theta bound at fixed v = s / s_n (lp/mod.rs:211-284)
__device__ inline void theta_bound_low_eta_point(int n, double a, int i, double& v_i, double& f_i) {
    double ln_j_crit = log(0.443) - log(double(n)) / 3.0;
    double z = cos(PI_D * double(i) / 15.0);
    double v = 0.5 * (z + 1.0);
    v_i = v;
    double x = 2.0 * double(n) * a * sqrt(v * (1.0 - v) / (1.0 + 0.5 * a * a));
    double y = 0.25 * double(n) * a * a * v / (1.0 + 0.5 * a * a);
    double x_crit = y > double(n) / 6.0 ? 4.0 * sqrt(y) * sqrt(double(n) - 2.0 * y) : double(n) + 2.0 * y;
    double ln_j_bdy = fmin(ln_dbessel_x_zero(n % 2 == 0 ? n : n + 1, y), ln_j_crit);
    double cos_theta = fmax(1.0 - x_crit * log(0.01) / (x * (ln_j_bdy - ln_j_crit)), -1.0);
    double f = acos(cos_theta);
    f_i = isnan(f) ? PI_D : f;
}
// rate differential in v and theta, 0 < theta < pi/2 (lp/classical.rs:17-40)
template <class DB>
__device__ inline double lp_cl_ddpr(const DB& D, double a, double v, double theta) {
    int n = D.n;
    double x = 2.0 * double(n) * a * cos(theta) * sqrt(v * (1.0 - v) / (1.0 + 0.5 * a * a));
    double y = 0.25 * double(n) * a * a * v / (1.0 + 0.5 * a * a);
    double j[5];
    dbessel_eval(D, x, y, j);
    double g0, g1, g2;
    if (n % 2 == 0) { g0 = j[2]; g1 = -0.5 * (j[1] + j[3]); g2 = 0.25 * (j[0] + 2.0 * j[2] + j[4]); }
    else { g0 = -j[2]; g1 = 0.5 * (j[1] + j[3]); g2 = -0.25 * (j[0] + 2.0 * j[2] + j[4]); }
    return (-g0 * g0 - a * a * (g0 * g2 - g1 * g1)) / (2.0 * PI_D);
}
// The points at which the classical ceiling samples the rate (lp/classical.rs:43-110)
__device__ inline int lp_cl_ceiling_count(int n, double a) {
    if (n == 1) return 1;
    if (a <= 1.5) return 16;
    if (a < 5.0) return a > 4.0 ? 32 : 16;
    return a > 8.0 ? 64 : 32;
}
__device__ inline void lp_cl_ceiling_arg(int n, double a, int k, double& v, double& theta) {
    theta = 0.0;
    if (n == 1) {
        v = 0.01;
        theta = PI_D / 2.0;
    } else if (a <= 1.5) {
        double v_min = 0.1 + (9.0 / sqrt(double(n)) - 1.0) / 10.0;
        double v_max = fmin(v_min + 0.4, 1.0);
        v = v_min + (v_max - v_min) * double(k + 1) / 16.0;
    } else if (a < 5.0) {
        int n_switch = int(3.0 * pow(a / 2.0, 4.0 / 3.0));
        theta = n > n_switch ? 0.0 : (PI_D / 2.0) * (1.0 - log(double(n)) / log(double(n_switch)));
        double v_min = fmin(0.35 + (7.0 / sqrt(double(n)) - 1.0) / 9.0, 0.55);
        double v_max = 0.55 + (7.0 / sqrt(double(n)) - 1.0) / 9.0;
        int num = a > 4.0 ? 32 : 16;
        v = v_min + (v_max - v_min) * double(k) / double(num);
    } else {
        double n_switch = 10.0 * ((a / 5.0) * (a / 5.0));
        if (!(n > int(n_switch))) {
            double pw = 0.3 + (0.7 - 0.3) * (log(a) - log(5.0)) / (log(20.0) - log(5.0));
            theta = (PI_D / 2.0) * (pow(double(n), -pw) - pow(n_switch, -pw)) / (1.0 - pow(n_switch, -pw));
        }
        double v_min = 0.3 + 0.4 * log(double(n)) * pow(double(n), -0.7);
        double v_max = v_min + 0.2;
        int num = a > 8.0 ? 64 : 32;
        v = v_min + (v_max - v_min) * double(k) / double(num);
    }
}
// 1.1 x the largest of up to 64 samples (lp/classical.rs:43-110)
__device__ inline double lp_cl_ceiling(const DBessel& D, double a) {
    double mx = NAN;
    int m = lp_cl_ceiling_count(D.n, a);
    for (int k = 0; k < m; ++k) {
        double v, th;
        lp_cl_ceiling_arg(D.n, a, k, v, th);
        mx = fmax(mx, lp_cl_ddpr(D, a, v, th));
    }
    return 1.1 * mx;
}
// f(a, eta), dP/dphase = alpha f / eta (lp/classical.rs:189-213); `P` = the LMA pack
__device__ inline bool lp_cl_rate(const double* P, const LmaScalars& s, double a, double eta, double& out) {
    double x = log(a);
    if (x < s.cl_rate_min) { out = a * a * eta / 3.0; return true; }
    double fx = (x - s.cl_rate_min) / s.cl_rate_step;
    if (!(fx < double(lma_pack::CL_N - 1))) return false;
    int ix = int(fx);
    double dx = fx - double(ix);
    const double* T = P + lma_pack::CL_RATE;
    out = 2.0 * eta / (1.0 + 0.5 * a * a) * exp((1.0 - dx) * T[ix] + dx * T[ix + 1]);
    return true;
}
// harmonic from the [16][2] CDFs of the two neighbouring a columns (lp/classical.rs:215-252);
// 0 outside the table
__device__ inline int lp_cl_harmonic(const double* P, const LmaScalars& s, double a, double frac) {
    double la = log(a);
    if (la <= s.cl_cdf_min) return 1;
    double fx = (la - s.cl_cdf_min) / s.cl_cdf_step;
    if (!(fx < double(lma_pack::CL_N - 1))) return 0;
    int ix = int(fx);
    double dx = fx - double(ix);
    double n = 0.0;
    for (int k = 0; k < 2; ++k) {
        const double* t = P + lma_pack::CL_CDF + 32 * (ix + k);
        double nk = 0.9;
        if (frac > t[1] && !pwmci_invert(t, 16, frac, false, nk)) return 0;
        n += nk * (k == 0 ? 1.0 - dx : dx);
    }
    n = ceil(n);
    return (n >= 1.0 && n <= 1.0e7) ? int(n) : 0;
}
// Stokes parameters of a classical LP photon at (v, theta) (lp/classical.rs:296-330).
// The reference evaluates J at the signed x of the final quadrant and applies the
// parity of a negative argument for every quadrant (lp/classical.rs:300-314); so here.
template <class DB>
__device__ inline sv4 lp_cl_stokes(const DB& D, double a, double v, double theta) {
    int n = D.n;
    double x = 2.0 * double(n) * a * cos(theta) * sqrt(v * (1.0 - v) / (1.0 + 0.5 * a * a));
    double y = 0.25 * double(n) * a * a * v / (1.0 + 0.5 * a * a);
    double j[5];
    dbessel_eval(D, x, y, j);
    double A0, A1, A2;
    if (n % 2 == 0) { A0 = j[2]; A1 = -0.5 * (j[1] + j[3]); A2 = 0.25 * (j[0] + 2.0 * j[2] + j[4]); }
    else { A0 = -j[2]; A1 = 0.5 * (j[1] + j[3]); A2 = -0.25 * (j[0] + 2.0 * j[2] + j[4]); }
    double u = sqrt((1.0 + 0.5 * a * a) * (1.0 - v) / v);
    double st_, ct_;
    sincos(theta, &st_, &ct_);
    double A0a = A0 / a;
    double xi0 = 2.0 * (A1 * A1 - A0 * A2) - 2.0 * A0a * A0a;
    double xi1 = 2.0 * (A1 * A1 - A0 * A2) - 2.0 * (1.0 + 2.0 * u * u * st_ * st_) * A0a * A0a;
    double xi2 = 2.0 * (u * A0a) * (u * A0a) * (2.0 * st_ * ct_) + 4.0 * u * A0 * A1 * st_ / a;
    return {1.0, xi1 / xi0, xi2 / xi0, 0.0};
}
// Stokes parameters of an LP photon at (s, theta) (lp/mod.rs:626-676)
template <class DB>
__device__ inline sv4 lp_stokes(const DB& D, double a, double eta, double s, double theta) {
    int n = D.n;
    double sn = 2.0 * double(n) * eta / (1.0 + 0.5 * a * a);
    double wn = fmin(s / (sn * (1.0 - s)), 1.0);
    double x = -2.0 * double(n) * a * cos(theta) * sqrt(wn * (1.0 - wn)) / sqrt(1.0 + 0.5 * a * a);
    double y = a * a * s / (8.0 * eta * (1.0 - s));
    double j[5];
    dbessel_eval(D, fabs(x), y, j);
    double A0, A1, A2;
    if (x > 0.0) { A0 = j[2]; A1 = 0.5 * (j[1] + j[3]); A2 = 0.25 * (j[0] + 2.0 * j[2] + j[4]); }
    else if (n % 2 == 0) { A0 = j[2]; A1 = -0.5 * (j[1] + j[3]); A2 = 0.25 * (j[0] + 2.0 * j[2] + j[4]); }
    else { A0 = -j[2]; A1 = 0.5 * (j[1] + j[3]); A2 = -0.25 * (j[0] + 2.0 * j[2] + j[4]); }
    double u = sqrt(2.0 * double(n) * eta * (1.0 - s) / s - (1.0 + 0.5 * a * a));
    double st_, ct_;
    sincos(theta, &st_, &ct_);
    double A0a = A0 / a;
    double xi0 = (1.0 - s + 1.0 / (1.0 - s)) * (A1 * A1 - A0 * A2) - 2.0 * A0a * A0a;
    double xi1 = 2.0 * (A1 * A1 - A0 * A2) - 2.0 * (1.0 + 2.0 * u * u * st_ * st_) * A0a * A0a;
    double xi2 = -(2.0 * (u * A0a) * (u * A0a) * (2.0 * st_ * ct_) + 4.0 * u * A0 * A1 * st_ / a);
    return {1.0, xi1 / xi0, xi2 / xi0, 0.0};
}


// ---------------------------------------------------------------- photon from the ZMF (nonlinear_compton/mod.rs:58-138)
__device__ __forceinline__ d4 boost4(d4 s, d4 u) {  // four_vector.rs:68-80
    double gamma = u.t;
    double beta = sqrt(1.0 - 1.0 / (gamma * gamma));
    d3 n = unit(sp(u));
    d3 z = sp(s);
    double nz = dot(n, z);
    double c1 = (gamma - 1.0) * nz - gamma * beta * s.t;
    return {gamma * (s.t - beta * nz), z.x + c1 * n.x, z.y + c1 * n.y, z.z + c1 * n.z};
}

// Stokes parameters from the LMA basis to the global basis (polarization.rs:104-157).
__device__ __forceinline__ sv4 stokes_from_lma(sv4 s, d4 k) {
    double k_minus = k.t - k.z;
    double cos_t = (k.z * k_minus - k.x * k.x) / (k_minus * hypot(k.x, k.z));
    double sign = signbit(k.x * k.y) ? -1.0 : 1.0;
    cos_t = fmin(fmax(cos_t, -1.0), 1.0);
    double sin_t = -(sign * sqrt(1.0 - cos_t * cos_t));
    double c2 = cos_t * cos_t - sin_t * sin_t, s2 = 2.0 * cos_t * sin_t;
    return {s.i, c2 * s.q - s2 * s.u, s2 * s.q + c2 * s.u, s.v};
}
// polarization.rs:116-131 (in_lma_basis): the inverse rotation of stokes_from_lma
__device__ __forceinline__ sv4 stokes_to_lma(sv4 s, d4 k) {
    double k_minus = k.t - k.z;
    double cos_t = (k.z * k_minus - k.x * k.x) / (k_minus * hypot(k.x, k.z));
    double sign = signbit(k.x * k.y) ? -1.0 : 1.0;
    cos_t = fmin(fmax(cos_t, -1.0), 1.0);
    double sin_t = sign * sqrt(1.0 - cos_t * cos_t);
    double c2 = cos_t * cos_t - sin_t * sin_t, s2 = 2.0 * cos_t * sin_t;
    return {s.i, c2 * s.q - s2 * s.u, s2 * s.q + c2 * s.u, s.v};
}
__device__ __forceinline__ sv4 stokes_rotate(sv4 s, double theta) {  // polarization.rs:44-53
    double s2, c2;
    sincos(2.0 * theta, &s2, &c2);
    return {s.i, c2 * s.q + s2 * s.u, -s2 * s.q + c2 * s.u, s.v};
}

}  // namespace ptg
