// Laser fields and the Lorentz-force pusher, device side.
//
// FastFocusedLaser: Gaussian beam to fourth order in the diffraction angle
// (Salamin, Appl. Phys. B 86, 319 (2007)) with a pulse envelope, as in
// src/field/fast_focused_laser.rs:61-218. FastPlaneWave: analytic plane wave
// with chirp (src/field/fast_plane_wave.rs:110-185). Pusher: Vay's scheme with
// optional classical radiation reaction (src/field/lcf.rs:18-78).
#pragma once
#include "dev_common.h"

namespace ptg {

enum { K_PLANE_WAVE = 0, K_FOCUSED_LASER = 1, K_FAST_PLANE_WAVE = 2, K_FAST_FOCUSED_LASER = 3 };
enum { ENV_COS2 = 0, ENV_FLATTOP = 1, ENV_GAUSSIAN = 2 };

// All derived laser constants, computed once on the host (ptg_set_laser).
struct LaserDev {
    int kind, envelope, circular, finite_bandwidth;
    double a0, k0;            // wavevector = k0 (1, 0, 0, 1)
    double omega;             // c k0
    double n_cycles;          // effective n_cycles used by contains()/envelope
    double max_phase;         // contains(): k.r < max_phase
    double dt;                // max_timestep
    double sin_pa, cos_pa;    // polarisation angle
    double chirp_b;
    // focused beam
    double inv_waist, inv_zr, eps, E0, B0;
    double gauss_c;           // Gaussian envelope: f = exp(-gauss_c phase^2), df = -2 gauss_c phase f
    // plane wave
    double amplitude;         // m c^2 k0 a0 / e
    double a_norm;            // e / (m c omega): a = a_norm |E|
    // FocusedLaser (LMA): waist^2, Rayleigh range, (omega * duration)^2
    double waist_sqd, zr, tau_sqd;
};

// ---------------------------------------------------------------- FastFocusedLaser
struct BeamField { d3 reE, imE, reB, imB; };

// Complex field of the x-polarised Gaussian beam at (ct, x, y, z); the carrier
// phase k.r is passed in so callers can reuse it. The fourth-order expansions of
// fast_focused_laser.rs:61-111 are regrouped as polynomials in f = i/(z + i)
// with real coefficients and evaluated by Horner's rule, which keeps few
// complex temporaries alive at once.
__device__ __forceinline__ cplx horner_step(cplx acc, cplx f, double c) {
    // acc * f + c
    return {acc.re * f.re - acc.im * f.im + c, acc.re * f.im + acc.im * f.re};
}

__device__ __forceinline__ BeamField ffl_beam(const LaserDev& L, double x_m, double y_m, double z_m, double phase) {
    // rotate into the polarisation frame (rotate_around_z(-pol_angle))
    double xr = L.cos_pa * x_m + L.sin_pa * y_m;
    double yr = L.cos_pa * y_m - L.sin_pa * x_m;
    double x = xr * L.inv_waist, y = yr * L.inv_waist, z = z_m * L.inv_zr;
    double rho2 = x * x + y * y;
    double rho4 = rho2 * rho2;
    double e = L.eps, e2 = e * e, e4 = e2 * e2;
    // f = i / (z + i) = (1 + i z) / (1 + z^2)
    double den = 1.0 / (1.0 + z * z);
    cplx f{den, z * den};
    cplx f2{f.re * f.re - f.im * f.im, 2.0 * f.re * f.im};
    // prefactor = f exp(-f rho^2 + i phase)
    double mag = exp(-f.re * rho2);
    double sn, cs;
    sincos_phase(phase - f.im * rho2, &sn, &cs);
    cplx pre = f * cplx{mag * cs, mag * sn};

    auto close = [&](cplx p, cplx poly) { return p * poly; };
    double x2 = x * x;
    // Ex/(-i pre) = 1 + f^2 [e^2 x^2 + e^4/8 + f (-(e^2 + e^4) rho^4/4 ... )]: coefficients a2..a6
    double a6 = e4 * rho4 * rho4 / 32.0;
    double a5 = 0.25 * e4 * rho4 * (-x2 - 0.5 * rho2);
    double a4 = e4 * (x2 * rho2 - rho4 / 16.0);
    double a3 = -0.25 * e2 * rho4 - 0.25 * e4 * rho2;
    double a2 = e2 * x2 + 0.125 * e4;
    cplx h = horner_step(horner_step(horner_step(cplx{a6, 0.0}, f, a5), f, a4), f, a3);
    h = horner_step(h, f, a2);
    cplx px = f2 * h;
    px.re += 1.0;
    cplx ex = mul_mi(close(pre, px));
    // Ey/(-i pre) = f^2 (e^2 x y + f^2 (e^4 x y rho^2 - f e^4 x y rho^4 / 4))
    double xy = x * y;
    cplx hy = horner_step(horner_step(horner_step(cplx{-0.25 * e4 * xy * rho4, 0.0}, f, e4 * xy * rho2), f, 0.0), f, e2 * xy);
    cplx ey = mul_mi(close(pre, f2 * hy));
    // Ez/pre = f (e x + f (-e^3 x/2 + f (e^3 x rho^2 - f e^3 x rho^4/4)))
    double e3 = e2 * e;
    cplx hz = horner_step(horner_step(horner_step(cplx{-0.25 * e3 * x * rho4, 0.0}, f, e3 * x * rho2), f, -0.5 * e3 * x), f, e * x);
    cplx ez = close(pre, f * hz);
    // By/(-i pre) = 1 + f^2 (b2 + f (b3 + f (b4 + f (b5 + f b6))))
    double b6 = e4 * rho4 * rho4 / 32.0;
    double b5 = -0.25 * e4 * rho4 * rho2;
    double b4 = 0.3125 * e4 * rho4;
    double b3 = -0.25 * e2 * rho4 + 0.25 * e4 * rho2;
    double b2 = 0.5 * e2 * rho2 - 0.125 * e4;
    cplx hb = horner_step(horner_step(horner_step(cplx{b6, 0.0}, f, b5), f, b4), f, b3);
    hb = horner_step(hb, f, b2);
    cplx pby = f2 * hb;
    pby.re += 1.0;
    cplx by = mul_mi(close(pre, pby));
    // Bz/pre = f (e y + f (e^3 y/2 + f (e^3 y rho^2/2 - f e^3 y rho^4/4)))
    cplx hbz = horner_step(horner_step(horner_step(cplx{-0.25 * e3 * y * rho4, 0.0}, f, 0.5 * e3 * y * rho2), f, 0.5 * e3 * y), f, e * y);
    cplx bz = close(pre, f * hbz);

    // back to the lab frame (rotate_around_z(+pol_angle)), scale by E0, B0
    auto rz = [&](double ax, double ay, double az, double s) {
        return d3{s * (L.cos_pa * ax - L.sin_pa * ay), s * (L.cos_pa * ay + L.sin_pa * ax), s * az};
    };
    return {rz(ex.re, ey.re, ez.re, L.E0), rz(ex.im, ey.im, ez.im, L.E0), rz(0.0, by.re, bz.re, L.B0), rz(0.0, by.im, bz.im, L.B0)};
}

__device__ __forceinline__ void ffl_envelope(const LaserDev& L, double phase, double& f, double& df) {
    if (L.envelope == ENV_GAUSSIAN) {
        f = exp(-L.gauss_c * phase * phase);
        df = -2.0 * L.gauss_c * phase * f;
    } else if (L.envelope == ENV_COS2) {
        double nc = L.n_cycles;
        if (fabs(phase) < PI_D * nc) {
            double sn, cs;
            sincos_phase(phase / (2.0 * nc), &sn, &cs);
            f = cs * cs;
            df = -sn * cs / nc;  // -tan(phi/2n) cos^2(phi/2n) / n
        } else { f = 0.0; df = 0.0; }
    } else {
        double nc = L.n_cycles, ap = fabs(phase);
        if (ap > PI_D * (nc + 1.0)) { f = 0.0; df = 0.0; }
        else if (ap > PI_D * (nc - 1.0)) {
            double arg = 0.25 * (ap - (nc - 1.0) * PI_D);
            double sa, ca;
            sincos_phase(arg, &sa, &ca);
            f = ca * ca;
            df = -0.25 * copysign(1.0, phase) * (2.0 * sa * ca);
        } else { f = 1.0; df = 0.0; }
    }
}

// E, B (SI) of the pulsed focused laser at r (fast_focused_laser.rs:190-218).
// CP is a compile-time choice so that linear polarisation does not carry the
// registers of the second (y-polarised) beam evaluation.
template <bool CP>
__device__ __forceinline__ void ffl_fields_t(const LaserDev& L, d4 r, d3& E, d3& B) {
    double phase = L.k0 * (r.t - r.z);
    double f, df;
    ffl_envelope(L, phase, f, df);
    BeamField b = ffl_beam(L, r.x, r.y, r.z, phase);
    E = f * b.reE + df * b.imE;
    B = f * b.reB + df * b.imB;
    if (CP) {
        // second beam polarised along y: evaluate at r rotated by -pi/2 about z,
        // take the quadrature component, rotate back by +pi/2
        BeamField c = ffl_beam(L, r.y, -r.x, r.z, phase);
        d3 ey = f * c.imE - df * c.reE, by = f * c.imB - df * c.reB;
        E = E + d3{-ey.y, ey.x, ey.z};
        B = B + d3{-by.y, by.x, by.z};
    }
}

__device__ __forceinline__ void ffl_fields(const LaserDev& L, d4 r, d3& E, d3& B) {
    if (L.circular) ffl_fields_t<true>(L, r, E, B);
    else ffl_fields_t<false>(L, r, E, B);
}

// ---------------------------------------------------------------- FastPlaneWave
__device__ __forceinline__ void fpw_fields(const LaserDev& L, d4 r, d3& E, d3& B) {
    double phi = L.k0 * (r.t - r.z);
    double psi = phi * (1.0 + L.chirp_b * phi);
    double dpsi = 1.0 + 2.0 * L.chirp_b * phi;
    double nc = L.n_cycles, f, df;
    if (L.envelope == ENV_COS2) {
        if (fabs(phi) < nc * PI_D) {
            double sh, ch;
            sincos_phase(phi / (2.0 * nc), &sh, &ch);
            f = ch * ch;
            df = -(2.0 * sh * ch) / (2.0 * nc);
        } else { f = 0.0; df = 0.0; }
    } else if (L.envelope == ENV_FLATTOP) {
        double ap = fabs(phi);
        if (ap > PI_D * (nc + 1.0)) { f = 0.0; df = 0.0; }
        else if (ap > PI_D * (nc - 1.0)) {
            double arg = 0.25 * (ap - (nc - 1.0) * PI_D);
            double sa, ca;
            sincos_phase(arg, &sa, &ca);
            f = ca * ca;
            df = -copysign(1.0, phi) * 0.25 * (2.0 * sa * ca);
        } else { f = 1.0; df = 0.0; }
    } else {
        double q = phi / (PI_D * nc);
        f = exp2(-0.5 * q * q);
        df = -LN2_D * phi * f / ((PI_D * nc) * (PI_D * nc));
    }
    double sn, cs;
    sincos_phase(psi, &sn, &cs);
    double dax = sn * df + cs * dpsi * f;
    double day = L.circular ? (cs * df - sn * dpsi * f) : 0.0;
    double A = L.amplitude, Ab = L.amplitude / C_LIGHT;
    d3 e{-A * dax, -A * day, 0.0}, b{Ab * day, -Ab * dax, 0.0};
    E = rot_z(e, L.sin_pa, L.cos_pa);
    B = rot_z(b, L.sin_pa, L.cos_pa);
}


__device__ __forceinline__ bool field_contains(const LaserDev& L, d4 r) { return L.k0 * (r.t - r.z) < L.max_phase; }

// ---------------------------------------------------------------- Vay pusher
// One step of Vay's relativistic pusher; E, B at mid-step. Returns the work done
// by the field (in mc^2). rr_mode: 0 none, 1 Landau-Lifshitz, 2 LL with Gaunt factor
// (src/field/lcf.rs:18-78).
template <typename GauntFn>
__device__ __forceinline__ double vay_push(d3& u, d3 E, d3 B, double rqm, double dt, int rr_mode, GauntFn gaunt) {
    double gamma = sqrt(1.0 + nsq(u));
    d3 v = (C_LIGHT / gamma) * u;
    double alpha = rqm * dt / (2.0 * C_LIGHT);
    d3 u_half = u + alpha * (E + cross(v, B));
    double gamma_half = sqrt(1.0 + nsq(u_half));
    double dwork = rqm * dot(E, u_half) * dt / (gamma_half * C_LIGHT);
    d3 u_rad{0.0, 0.0, 0.0};
    if (rr_mode) {
        double mag = sqrt(nsq(u_half));
        d3 beta = (1.0 / mag) * u_half;
        double eb = dot(E, beta);
        double E_rf_sqd = nsq(E + C_LIGHT * cross(beta, B)) - eb * eb;
        double chi = E_rf_sqd > 0.0 ? gamma_half * sqrt(E_rf_sqd) / E_CRIT : 0.0;
        double power = 2.0 * ALPHA * chi * chi / (3.0 * T_COMPTON);
        double g = rr_mode == 2 ? gaunt(chi) : 1.0;
        u_rad = (g * power * dt / mag) * u_half;
    }
    d3 u_prime = u_half + alpha * E;
    double gp2 = 1.0 + nsq(u_prime);
    d3 tau = (alpha * C_LIGHT) * B;
    double u_star = dot(u_prime, tau);
    double tt = nsq(tau);
    double sigma = gp2 - tt;
    double g_new = sqrt(0.5 * sigma + sqrt(0.25 * sigma * sigma + tt + u_star * u_star));
    d3 t = (1.0 / g_new) * tau;
    double s = 1.0 / (1.0 + nsq(t));
    u = s * (u_prime + dot(u_prime, t) * t + cross(u_prime, t)) - u_rad;
    return dwork;
}

}  // namespace ptg
