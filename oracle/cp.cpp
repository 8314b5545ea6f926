// Ptarmigan CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.hpp).
// Bessel J_n(x) and circularly polarised nonlinear Compton scattering in the
// LMA: quantum (nonlinear_compton/cp/mod.rs, cp/tables/mod.rs) and classical
// (cp/classical.rs). The quantum CDF table is not shipped by the reference;
// the oracle reads the regenerated one ("nonlinear_compton/cp/tables/cdf:TABLE")
// on the grid of the total-rate table, as cp/mod.rs:529-611 writes it.
#include <cstdio>
#include <cstdlib>

#include "oracle.hpp"

namespace orc {

// ---------------------------------------------------------------- Bessel J (special_functions/bessel.rs)
static double gl_cos_sum(int nodes, double x, int order) {
    Table N = table(nodes == 16 ? "quadrature/tables:GAUSS_16_NODES" : "quadrature/tables:GAUSS_32_NODES");
    Table W = table(nodes == 16 ? "quadrature/tables:GAUSS_16_WEIGHTS" : "quadrature/tables:GAUSS_32_WEIGHTS");
    double acc = 0.0;
    for (int i = 0; i < nodes; ++i) {
        if (order == 0) {
            acc += W[i] * 0.5 * std::cos(-x * std::sin(FRAC_PI_2 * (N[i] + 1.0)));
        } else {
            double s = FRAC_PI_2 * (N[i] + 1.0);
            acc += W[i] * 0.5 * std::cos(s - x * std::sin(s));
        }
    }
    return acc;
}
// bessel.rs:37-72
static double j0(double x) {
    if (x < 0.1) return 1.0 - x * x / 4.0 + x * x * x * x / 64.0;
    if (x < 5.0) return gl_cos_sum(16, x, 0);
    if (x < 20.0) return gl_cos_sum(32, x, 0);
    double prefactor = std::sqrt(2.0 / (PI * x));
    double phase = PI / 4.0 - x;
    if (x < 60.0)
        return prefactor * ((1.0 - 9.0 / (128.0 * x * x) + 3675.0 / (32768.0 * x * x * x * x) - 0.5725014209747314 * powi(x, -6)) *
                                std::cos(phase) -
                            (1.0 / (8.0 * x) - 75.0 / (1024.0 * x * x * x) + 0.2271080017089844 * powi(x, -5)) * std::sin(phase));
    return prefactor * ((1.0 - 9.0 / (128.0 * x * x) + 3675.0 / (32768.0 * x * x * x * x)) * std::cos(phase) -
                        (1.0 / (8.0 * x) - 75.0 / (1024.0 * x * x * x)) * std::sin(phase));
}
// bessel.rs:74-118
static double j1(double x) {
    if (x < 0.1) return (x / 2.0) * (1.0 - x * x / 8.0 + x * x * x * x / 192.0);
    if (x < 5.0) return gl_cos_sum(16, x, 1);
    if (x < 20.0) return gl_cos_sum(32, x, 1);
    double prefactor = std::sqrt(2.0 / (PI * x));
    double phase = PI / 4.0 + x;
    if (x < 60.0)
        return prefactor * (-(1.0 + 15.0 / (128.0 * x * x) - 4725.0 / (32768.0 * x * x * x * x) + 0.6765925884246826 * powi(x, -6)) *
                                std::cos(phase) +
                            (3.0 / (8.0 * x) - 105.0 / (1024.0 * x * x * x) + 0.2775764465332031 * powi(x, -5)) * std::sin(phase));
    return prefactor * (-(1.0 + 15.0 / (128.0 * x * x) - 4725.0 / (32768.0 * x * x * x * x)) * std::cos(phase) +
                        (3.0 / (8.0 * x) - 105.0 / (1024.0 * x * x * x)) * std::sin(phase));
}
// bessel.rs:120-176: Miller's backward recurrence from an even start above n,
// normalised by J_0 + 2 sum J_2k = 1
static void triple_j(int n, double x, double out[3]) {
    if (x == 0.0) {
        out[0] = n == 1 ? 1.0 : 0.0;
        out[1] = n == 0 ? 1.0 : 0.0;
        out[2] = n == -1 ? 1.0 : 0.0;
        return;
    }
    double v = x / double(n);
    int nmax = 1 + n + int(5.0 / (1.0 - 0.9 * v));
    if (((nmax % 2) + 2) % 2 != 0) nmax += 1;
    int n_even = (((n % 2) + 2) % 2 == 0) ? 1 : 0;
    int k_stop = n - 1 - n_even;
    double quad[4] = {1.0, 0.0, 0.0, 0.0};
    double total = 0.0;
    auto step = [&](int k) {
        quad[3] = quad[1];
        quad[2] = quad[0];
        quad[1] = double(2 * (k + 2)) * quad[2] / x - quad[3];
        quad[0] = double(2 * (k + 1)) * quad[1] / x - quad[2];
        total += quad[0];
    };
    for (int k = nmax; k >= k_stop; k -= 2) step(k);
    double t0 = quad[n_even], t1 = quad[n_even + 1], t2 = quad[n_even + 2];
    for (int k = k_stop - 2; k >= 0; k -= 2) step(k);
    double norm = 2.0 * total - quad[0];
    if (std::isnan(norm)) { out[0] = out[1] = out[2] = 0.0; return; }
    out[0] = t0 / norm; out[1] = t1 / norm; out[2] = t2 / norm;
}
// bessel.rs:178-186
static double zeta_at(double z) {
    if (z < 1.0) {
        double tmp = std::log((1.0 + std::sqrt(1.0 - z * z)) / z) - std::sqrt(1.0 - z * z);
        return std::pow(3.0 * tmp / 2.0, 2.0 / 3.0);
    }
    double tmp = std::sqrt(z * z - 1.0) - std::acos(1.0 / z);
    return -std::pow(3.0 * tmp / 2.0, 2.0 / 3.0);
}
// bessel.rs:188-198: uniform (Debye) expansion through Ai, Ai'; Ai of a
// negative argument is None in the reference and taken as 0
static double j_asymptotic(int n, double x) {
    double m = double(n);
    double z = x / m;
    double zeta = zeta_at(z);
    double arg = std::pow(m, 2.0 / 3.0) * zeta;
    double ai = 0.0, aip = 0.0;
    if (!airy_ai(arg, &ai)) ai = 0.0;
    if (!airy_ai_prime(arg, &aip)) aip = 0.0;
    double prefactor = 4.0 * zeta / (1.0 - z * z);
    double ak = 1.0;
    double bk = 5.0 / (48.0 * zeta * zeta) - (2.0 + 3.0 * z * z) / (24.0 * std::sqrt(zeta) * std::pow(1.0 - z * z, 1.5));
    return std::pow(prefactor, 0.25) * (ai * ak / std::pow(m, 1.0 / 3.0) - aip * bk / std::pow(m, 5.0 / 3.0));
}
// bessel.rs:26-33: (J_{n-1}, J_n, J_{n+1})
void bessel_j_pm(int n, double x, double out[3]) {
    if (n < 1000) { triple_j(n, x, out); return; }
    double jn = j_asymptotic(n, x), jp = j_asymptotic(n + 1, x);
    out[0] = double(2 * n) * jn / x - jp;
    out[1] = jn;
    out[2] = jp;
}
// bessel.rs:18-24
double bessel_j(int n, double x) {
    if (n == 0) return j0(x);
    if (n == 1) return j1(x);
    double t[3];
    triple_j(n, x, t);
    return t[1];
}

// ---------------------------------------------------------------- CP nonlinear Compton, quantum (cp/mod.rs)
static const char* CPT = "nonlinear_compton/cp/tables/total:";
static const char* CPC = "nonlinear_compton/cp/tables/cdf:TABLE";
static const char* CLR = "nonlinear_compton/cp/classical/rate_table:";
static const char* CLC = "nonlinear_compton/cp/classical/cdf_table:";
static Table cpt(const char* base, const char* n) { return table((std::string(base) + n).c_str()); }
constexpr double LASER_S3 = -1.0;  // cp/mod.rs:12

// Rust's `(4 * n * n) as f64` in i32 arithmetic wraps in release builds for n > 23170
static double four_n_sqd_i32(int n) { return double(int32_t(uint32_t(4) * uint32_t(n) * uint32_t(n))); }

// cp/mod.rs:17-38
double cp_diff_partial_rate(int n, double a, double eta, double v) {
    if (v < 0.0 || v >= 1.0) return 0.0;
    double sn = 2.0 * double(n) * eta / (1.0 + a * a);
    double smax = sn / (1.0 + sn);
    double vsmax = v * smax;
    double z = std::sqrt(4.0 * powi(double(n), 2) * (a * a / (1.0 + a * a)) * (vsmax / (sn * (1.0 - vsmax))) *
                         (1.0 - vsmax / (sn * (1.0 - vsmax))));
    double j[3];
    bessel_j_pm(n, z, j);
    return -smax * (powi(j[1], 2) + 0.5 * a * a * (1.0 + 0.5 * powi(vsmax, 2) / (1.0 - vsmax)) *
                                        (2.0 * powi(j[1], 2) - powi(j[2], 2) - powi(j[0], 2)));
}
static int cp_n_switch(double a) { return int(32.3 * (1.0 + 0.476 * std::pow(a, 1.56))); }
// cp/mod.rs:42-75
double cp_ceiling(int n, double a, double eta) {
    double sn = 2.0 * double(n) * eta / (1.0 + a * a);
    double lower = 0.0, upper = 0.0;
    if (sn < 2.0 || n < cp_n_switch(a)) {
        double v_opt = (1.0 + sn) / (2.0 + sn);
        for (int i = 0; i < 16; ++i) {
            double fl = cp_diff_partial_rate(n, a, eta, double(i) * v_opt / 16.0);
            double fu = cp_diff_partial_rate(n, a, eta, v_opt + double(i) * (1.0 - v_opt) / 16.0);
            lower = i == 0 ? fl : std::fmax(lower, fl);
            upper = i == 0 ? fu : std::fmax(upper, fu);
        }
    } else {
        Table N = table("quadrature/tables:GAUSS_16_NODES");
        for (int i = 0; i < 16; ++i) {
            double ul = 0.5 * (1.0 + sn) * (N[i] + 1.0);
            double uu = (1.0 + sn) + 0.5 * 2.0 * (1.0 + sn) * (N[i] + 1.0);
            double fl = cp_diff_partial_rate(n, a, eta, ul / (1.0 + ul));
            double fu = cp_diff_partial_rate(n, a, eta, uu / (1.0 + uu));
            lower = i == 0 ? fl : std::fmax(lower, fl);
            upper = i == 0 ? fu : std::fmax(upper, fu);
        }
    }
    return 1.2 * std::fmax(lower, upper);
}
// cp/mod.rs:79-89
double cp_spectrum_low_eta(int n, double a, double v) {
    if (v < 0.0 || v >= 1.0) return 0.0;
    double z = std::sqrt(4.0 * a * a * powi(double(n), 2) * v * (1.0 - v) / (1.0 + a * a));
    double j[3];
    bessel_j_pm(n, z, j);
    return double(n) * (a * a * powi(j[0], 2) - 2.0 * (1.0 + a * a) * powi(j[1], 2) + a * a * powi(j[2], 2)) / (1.0 + a * a);
}
// cp/mod.rs:92-133
double cp_partial_rate(int n, double a, double eta) {
    Table N = table("quadrature/tables:GAUSS_32_NODES"), W = table("quadrature/tables:GAUSS_32_WEIGHTS");
    double sn = 2.0 * double(n) * eta / (1.0 + a * a);
    double lower = 0.0, upper = 0.0;
    if (sn < 2.0 || n < cp_n_switch(a)) {
        double vmid = (1.0 + sn) / (2.0 + sn);
        for (int i = 0; i < 32; ++i) {
            double v = 0.5 * vmid * (N[i] + 1.0);
            lower += 0.5 * vmid * W[i] * cp_diff_partial_rate(n, a, eta, v);
        }
        for (int i = 0; i < 32; ++i) {
            double v = vmid + 0.5 * (1.0 - vmid) * (N[i] + 1.0);
            upper += 0.5 * (1.0 - vmid) * W[i] * cp_diff_partial_rate(n, a, eta, v);
        }
    } else {
        for (int i = 0; i < 32; ++i) {
            double u = 0.5 * (1.0 + sn) * (N[i] + 1.0);
            lower += 0.5 * (1.0 + sn) * W[i] * cp_diff_partial_rate(n, a, eta, u / (1.0 + u)) / powi(1.0 + u, 2);
        }
        for (int i = 0; i < 32; ++i) {
            double u = (1.0 + sn) + 0.5 * 2.0 * (1.0 + sn) * (N[i] + 1.0);
            upper += 0.5 * 2.0 * (1.0 + sn) * W[i] * cp_diff_partial_rate(n, a, eta, u / (1.0 + u)) / powi(1.0 + u, 2);
        }
    }
    return lower + upper;
}
// cp/mod.rs:137-155
double cp_integrated_spectrum_low_eta(int n, double a) {
    Table N = table("quadrature/tables:GAUSS_32_NODES"), W = table("quadrature/tables:GAUSS_32_WEIGHTS");
    double lower = 0.0, upper = 0.0;
    for (int i = 0; i < 32; ++i) lower += 0.25 * W[i] * cp_spectrum_low_eta(n, a, 0.25 * (N[i] + 1.0));
    for (int i = 0; i < 32; ++i) upper += 0.25 * W[i] * cp_spectrum_low_eta(n, a, 0.25 * (N[i] + 3.0));
    return lower + upper;
}

// cp/classical.rs:22-44
bool cp_classical_rate(double a, double eta, double* out) {
    double x = std::log(a);
    double MIN = cpt(CLR, "MIN")[0], STEP = cpt(CLR, "STEP")[0];
    Table T = cpt(CLR, "TABLE");
    if (x < MIN) { *out = 2.0 * a * a * eta / 3.0; return true; }
    size_t ix = size_t((x - MIN) / STEP);
    if (!(ix < T.n - 1)) return false;
    double dx = (x - MIN) / STEP - double(ix);
    double f = (1.0 - dx) * T[ix] + dx * T[ix + 1];
    *out = eta * std::exp(f);
    return true;
}

// cp/tables/mod.rs:10-12
static bool cp_tables_contains(double a, double eta) {
    Table M = cpt(CPT, "MIN");
    return std::log(a) >= M[0] && a < 20.0 && std::log(eta) >= M[1] && eta < 2.0;
}
// cp/tables/mod.rs:14-34
static double cp_tables_interpolate(double a, double eta) {
    const size_t N_COLS = 61, N_ROWS = 70;
    Table M = cpt(CPT, "MIN"), S = cpt(CPT, "STEP"), T = cpt(CPT, "TABLE");
    size_t ia = size_t((std::log(a) - M[0]) / S[0]), ie = size_t((std::log(eta) - M[1]) / S[1]);
    if (!(ia < N_COLS - 1 && ie < N_ROWS - 1)) { std::fprintf(stderr, "oracle: CP NLC table index\n"); std::abort(); }
    double da = (std::log(a) - M[0]) / S[0] - double(ia), de = (std::log(eta) - M[1]) / S[1] - double(ie);
    auto at = [&](size_t r, size_t c) { return T[r * N_COLS + c]; };
    double f = (1.0 - da) * (1.0 - de) * at(ie, ia) + da * (1.0 - de) * at(ie, ia + 1) + (1.0 - da) * de * at(ie + 1, ia) +
               da * de * at(ie + 1, ia + 1);
    return std::exp(f);
}
// cp/mod.rs:162-176
bool cp_rate(double a, double eta, double* out) {
    Table M = cpt(CPT, "MIN");
    if (std::log(eta) < M[1]) return cp_classical_rate(a, eta, out);
    if (std::log(a) < M[0]) {
        *out = a * a * (2.0 + 8.0 * eta + 9.0 * eta * eta + eta * eta * eta) / (2.0 * eta * powi(1.0 + 2.0 * eta, 2)) -
               a * a * (2.0 + 2.0 * eta - eta * eta) * std::log(1.0 + 2.0 * eta) / (4.0 * eta * eta);
        return true;
    }
    if (cp_tables_contains(a, eta)) { *out = cp_tables_interpolate(a, eta); return true; }
    return false;
}

// cp/tables/mod.rs:36-44 and cp/classical.rs:71-79: the first 31 points in
// (ln n, ln(-ln(1 - F)))
static double cp_rescale(double frac, const double* t32, double out31[62]) {
    for (int i = 0; i < 31; ++i) {
        out31[2 * i] = std::log(t32[2 * i]);
        out31[2 * i + 1] = std::log(-1.0 * std::log(1.0 - t32[2 * i + 1]));
    }
    return std::log(-1.0 * std::log(1.0 - frac));
}
// one column of the harmonic inversion: 0.9 below the first tabulated CDF value
static double cp_invert_one(const double* t32, double frac) {
    if (frac <= t32[1]) return 0.9;
    double rs[62];
    double f2 = cp_rescale(frac, t32, rs);
    double x;
    if (!Interp{rs, 31, true}.invert(f2, &x)) { std::fprintf(stderr, "oracle: CP invert failed\n"); std::abort(); }
    return std::exp(x);
}
// cp/tables/mod.rs:46-121
int cp_harmonic(double a, double eta, double frac) {
    const size_t N_COLS = 61;
    Table M = cpt(CPT, "MIN"), S = cpt(CPT, "STEP"), T = table(CPC);
    if (std::log(a) <= M[0]) return 1;
    size_t ix = size_t((std::log(a) - M[0]) / S[0]);
    double dx = (std::log(a) - M[0]) / S[0] - double(ix);
    double n = 0.0;
    if (std::log(eta) <= M[1]) {
        n = cp_invert_one(T.p + 64 * ix, frac) * (1.0 - dx) + cp_invert_one(T.p + 64 * (ix + 1), frac) * dx;
    } else {
        size_t iy = size_t((std::log(eta) - M[1]) / S[1]);
        double dy = (std::log(eta) - M[1]) / S[1] - double(iy);
        size_t idx[4] = {N_COLS * iy + ix, N_COLS * iy + ix + 1, N_COLS * (iy + 1) + ix, N_COLS * (iy + 1) + ix + 1};
        double w[4] = {(1.0 - dx) * (1.0 - dy), dx * (1.0 - dy), (1.0 - dx) * dy, dx * dy};
        for (int k = 0; k < 4; ++k) n += cp_invert_one(T.p + 64 * idx[k], frac) * w[k];
    }
    return int(std::ceil(n));
}
// cp/classical.rs:83-123
int cp_classical_harmonic(double a, double frac) {
    double MIN = cpt(CLC, "MIN")[0], STEP = cpt(CLC, "STEP")[0];
    Table T = cpt(CLC, "TABLE");
    if (std::log(a) <= MIN) return 1;
    size_t ix = size_t((std::log(a) - MIN) / STEP);
    double dx = (std::log(a) - MIN) / STEP - double(ix);
    if (ix + 1 >= 61) { std::fprintf(stderr, "oracle: CP classical CDF index\n"); std::abort(); }
    double n = cp_invert_one(T.p + 64 * ix, frac) * (1.0 - dx) + cp_invert_one(T.p + 64 * (ix + 1), frac) * dx;
    return int(std::ceil(n));
}
// cp/classical.rs:46-69
double cp_classical_ceiling(int n, double a) {
    double n_switch = std::round(3.0 * (1.0 + 0.54 * std::pow(a, 2.75)));
    double mx;
    if (n == 1) {
        mx = cp_spectrum_low_eta(n, a, 0.99);
    } else if (a < 0.15 || n > int(n_switch)) {
        mx = cp_spectrum_low_eta(n, a, 0.5);
    } else {
        double b = a < 0.5 ? 2.3 * a : 0.8 * std::exp(-(a - 0.5) / 3.5) + 0.35;
        double v = 1.0 - 0.5 * std::pow(std::log(double(n)), b) / std::pow(std::log(n_switch), b);
        const int delta = 4;
        mx = 0.0;
        for (int i = -delta; i < delta; ++i) {
            double f = cp_spectrum_low_eta(n, a, std::fmin(v + 0.1 * double(i) / double(delta), 1.0));
            mx = i == -delta ? f : std::fmax(mx, f);
        }
    }
    return 1.1 * mx;
}

// Stokes parameters of the photon in the (k x k', k'_perp) basis, rotated to
// (e_1, e_2) (cp/mod.rs:203-232, cp/classical.rs:155-175)
static SV cp_stokes(int n, double a, double s, double sn, double v, double theta, bool classical) {
    double z = classical ? std::sqrt(four_n_sqd_i32(n) * (a * a / (1.0 + a * a)) * v * (1.0 - v))
                         : std::sqrt(four_n_sqd_i32(n) * (a * a / (1.0 + a * a)) * (s / (sn * (1.0 - s))) * (1.0 - s / (sn * (1.0 - s))));
    double j[3];
    bessel_j_pm(n, z, j);
    double jm = j[0], jn = j[1], jp = j[2];
    double base = jm * jm + jp * jp - 2.0 * jn * jn;
    double xi0, xi3;
    if (classical) {
        xi0 = 2.0 * base - powi(2.0 * jn / a, 2);
        xi3 = 2.0 * (1.0 - 2.0 * v) * (jm * jm - jp * jp);
    } else {
        xi0 = (1.0 - s + 1.0 / (1.0 - s)) * base - powi(2.0 * jn / a, 2);
        xi3 = LASER_S3 * (1.0 - s + 1.0 / (1.0 - s)) * (1.0 - 2.0 * s / (sn * (1.0 - s))) * (jm * jm - jp * jp);
    }
    double xi1 = 2.0 * base + 8.0 * jn * jn * (1.0 - powi(double(n) / z, 2) - 0.5 / (a * a));
    double xi2 = 0.0;
    SV sv{1.0, xi1 / xi0, xi2 / xi0, xi3 / xi0};
    return rotate_by(sv, FRAC_PI_2 + theta);
}
// cp/mod.rs:181-234
void cp_sample(double a, double eta, Xoshiro& rng, int fixed_n, int* n_out, double* s_out, double* theta_out, SV* sv_out,
               uint64_t* trials) {
    int n = fixed_n > 0 ? fixed_n : cp_harmonic(a, eta, rng.gen_f64());
    double mx = cp_ceiling(n, a, eta);
    double v;
    for (;;) {
        v = rng.gen_f64();
        double u = rng.gen_f64();
        double f = cp_diff_partial_rate(n, a, eta, v);
        if (trials) ++*trials;
        if (u <= f / mx) break;
    }
    double sn = 2.0 * double(n) * eta / (1.0 + a * a);
    double smax = sn / (1.0 + sn);
    double s = v * smax;
    double theta = 2.0 * PI * rng.gen_f64();
    *n_out = n;
    *s_out = s;
    *theta_out = theta;
    *sv_out = cp_stokes(n, a, s, sn, v, theta, false);
}
// cp/classical.rs:128-178
void cp_classical_sample(double a, double eta, Xoshiro& rng, int fixed_n, int* n_out, double* s_out, double* theta_out, SV* sv_out,
                         uint64_t* trials) {
    int n = fixed_n > 0 ? fixed_n : cp_classical_harmonic(a, rng.gen_f64());
    double mx = cp_classical_ceiling(n, a);
    double v;
    for (;;) {
        v = rng.gen_f64();
        double u = rng.gen_f64();
        double f = cp_spectrum_low_eta(n, a, v);
        if (trials) ++*trials;
        if (u <= f / mx) break;
    }
    double sn = 2.0 * double(n) * eta / (1.0 + a * a);
    double s = v * sn;
    double theta = 2.0 * PI * rng.gen_f64();
    *n_out = n;
    *s_out = s;
    *theta_out = theta;
    *sv_out = cp_stokes(n, a, s, sn, v, theta, true);
}

// The cumulative harmonic table of one (a, eta) node as cp/mod.rs:529-611 (quantum,
// partial_rate) and cp/classical.rs:220-300 (classical, integrated_spectrum_low_eta)
// build it: the partial rates of n = 1..n_max summed, then 32 points
// [n, F(n)/F(n_max)], the first up to 32 harmonics directly, else log-spaced and
// interpolated with PWMCI. Returns the total.
double cp_cdf_row(double a, double eta, bool classical, double out64[64]) {
    int n_max = int(10.0 * (1.0 + powi(a, 3)));
    std::vector<double> rates(2 * size_t(n_max));
    double cs = 0.0;
    for (int n = 1; n <= n_max; ++n) {
        cs = cs + (classical ? cp_integrated_spectrum_low_eta(n, a) : cp_partial_rate(n, a, eta));
        rates[2 * size_t(n - 1)] = double(n);
        rates[2 * size_t(n - 1) + 1] = cs;
    }
    double rate = rates[2 * size_t(n_max - 1) + 1];
    Interp I{rates.data(), n_max};
    auto eval = [&](double n) {
        double f;
        if (!I.evaluate(n, &f)) { std::fprintf(stderr, "oracle: CP CDF evaluate\n"); std::abort(); }
        return f;
    };
    out64[0] = rates[0];
    out64[1] = rates[1] / rate;
    double limit = rates[2 * size_t(n_max - 1)];
    if (n_max <= 32) {
        for (int i = 1; i <= 31; ++i) {
            if (i < n_max) { out64[2 * i] = rates[2 * i]; out64[2 * i + 1] = rates[2 * i + 1] / rate; }
            else { out64[2 * i] = double(i + 1); out64[2 * i + 1] = 1.0; }
        }
    } else if (n_max < 100) {
        for (int i = 1; i <= 7; ++i) { out64[2 * i] = rates[2 * i]; out64[2 * i + 1] = rates[2 * i + 1] / rate; }
        double delta = (std::log(double(n_max)) - std::log(9.0)) / (31.0 - 8.0);
        for (int i = 8; i <= 31; ++i) {
            double n = std::fmin(std::exp(std::log(9.0) + double(i - 8) * delta), limit);
            out64[2 * i] = n;
            out64[2 * i + 1] = eval(n) / rate;
        }
    } else {
        double delta = std::log(double(n_max)) / 31.0;
        for (int i = 1; i <= 31; ++i) {
            double n = std::fmin(std::exp(double(i) * delta), limit);
            out64[2 * i] = n;
            out64[2 * i + 1] = eval(n) / rate;
        }
    }
    return rate;
}

// ---------------------------------------------------------------- CP nonlinear Breit-Wheeler (pair_creation/cp/mod.rs)
static const char* CBT = "pair_creation/cp/total:";

// pair_creation/cp/mod.rs:35-40
static void cpbw_s_bounds(int n, double a, double eta, double* s_min, double* s_max) {
    double sn = 2.0 * double(n) * eta / (1.0 + powi(a, 2));
    *s_min = std::fmin(0.5 * (1.0 - std::sqrt(1.0 - 4.0 / sn)), 0.5);
    *s_max = 0.5;
}
// pair_creation/cp/mod.rs:45-66: unpolarised part, the S_1/S_2 and the S_3 coefficients
static void cpbw_elements(int n, double a, double eta, double s, double out3[3]) {
    double ell = double(n);
    double sn = 2.0 * ell * eta / (1.0 + powi(a, 2));
    double tmp = 1.0 / (sn * s * (1.0 - s));
    double z = std::sqrt((4.0 * ell * ell) * (powi(a, 2) / (1.0 + powi(a, 2))) * tmp * (1.0 - tmp));
    double j[3];
    bessel_j_pm(n, z, j);
    double jm = j[0], jn = j[1], jp = j[2];
    out3[0] = powi(jn, 2) + 0.25 * powi(a, 2) * ((1.0 - s) / s + s / (1.0 - s)) * (powi(jm, 2) + powi(jp, 2) - 2.0 * powi(jn, 2));
    out3[1] = (2.0 * powi(ell * a / z, 2) - (1.0 + powi(a, 2))) * powi(jn, 2) - 0.5 * powi(a, 2) * (powi(jm, 2) + powi(jp, 2));
    out3[2] = 0.25 * powi(a, 2) * ((1.0 - s) / s + s / (1.0 - s)) * (1.0 - 2.0 / (sn * s * (1.0 - s))) * (powi(jm, 2) - powi(jp, 2));
}
// pair_creation/cp/mod.rs:71-78. The reference binds (sin 2theta, cos 2theta) from
// sin_cos() to names in the opposite order; the arithmetic is kept as written.
static double cpbw_fully_resolved(int n, double a, double eta, double s, double theta, double sv1, double sv2, double sv3) {
    double e[3];
    cpbw_elements(n, a, eta, s, e);
    double first = std::sin(2.0 * theta), second = std::cos(2.0 * theta);
    return (e[0] - (sv1 * first + sv2 * second) * e[1] - LASER_S3 * sv3 * e[2]) / (2.0 * PI);
}
// pair_creation/cp/mod.rs:80-93: (S_3 = +1, S_3 = -1) integrated over theta
static C cpbw_over_theta(int n, double a, double eta, double s) {
    double e[3];
    cpbw_elements(n, a, eta, s, e);
    return C{e[0] - LASER_S3 * e[2], e[0] + LASER_S3 * e[2]};
}
static double cpbw_over_theta_s3(int n, double a, double eta, double s, double sv3) {
    C sp = cpbw_over_theta(n, a, eta, s);
    return 0.5 * (sp.re * (1.0 + sv3) + sp.im * (1.0 - sv3));
}
// pair_creation/cp/mod.rs:98-184: partial rate of harmonic n, (S_3 = +1, S_3 = -1)
C cpbw_partial_rate(int n, double a, double eta) {
    Table N = table("quadrature/tables:GAUSS_32_NODES"), W = table("quadrature/tables:GAUSS_32_WEIGHTS");
    double ell = double(n);
    double sn = 2.0 * ell * eta / (1.0 + a * a);
    if (sn <= 4.0) return C{0.0, 0.0};
    double beta = powi(a, 4) * powi(1.0 / ell + 1.0, 2) + 16.0 * powi(a * a - 2.0, 2) / (sn * sn) - 8.0 * a * a * (a * a - 2.0) / sn;
    beta = std::sqrt(beta) / (a * a - 2.0);
    double alpha = (a * a + 2.0 * ell) / (ell * (2.0 - a * a)) - 4.0 / sn;
    double s_peak = 0.5 * (1.0 - std::sqrt(alpha + beta));
    double s_min, s_max;
    cpbw_s_bounds(n, a, eta, &s_min, &s_max);
    auto seg = [&](double lo, double hi) {
        C acc{0.0, 0.0};
        for (int i = 0; i < 32; ++i) {
            double s = 0.5 * (hi - lo) * N[i] + 0.5 * (lo + hi);
            acc = acc + (0.5 * (hi - lo) * W[i]) * cpbw_over_theta(n, a, eta, s);
        }
        return acc;
    };
    if (std::isfinite(s_peak)) {
        double s_mid = 2.0 * s_peak - s_min;
        if (s_mid > s_max) {
            C lower = seg(s_min, s_peak), upper = seg(s_peak, s_max);
            return 2.0 * (upper + lower);
        }
        C lower = seg(s_min, s_peak), middle = seg(s_peak, s_mid), upper = seg(s_mid, s_max);
        return 2.0 * (lower + middle + upper);
    }
    return 2.0 * seg(s_min, s_max);
}
// pair_creation/cp/mod.rs:188-245: (s, theta) at harmonic n
static void cpbw_sample_at(int n, double a, double eta, double sv1, double sv2, double sv3, Xoshiro& rng, double* s_out,
                           double* theta_out) {
    Table N = table("quadrature/tables:GAUSS_32_NODES");
    double s_min, s_max;
    cpbw_s_bounds(n, a, eta, &s_min, &s_max);
    double mx = NAN;
    for (int i = 0; i < 32; ++i) mx = std::fmax(mx, cpbw_over_theta_s3(n, a, eta, 0.5 * (s_max - s_min) * N[i] + 0.5 * (s_min + s_max), sv3));
    mx = 1.5 * mx;
    double s;
    for (;;) {
        s = s_min + (s_max - s_min) * rng.gen_f64();
        double z = mx * rng.gen_f64();
        if (z < cpbw_over_theta_s3(n, a, eta, s, sv3)) break;
    }
    uint32_t r32 = uint32_t(rng.next_u64() >> 32);  // rand 0.7 gen_range(0, 2)
    if (int((uint64_t(r32) * 2u) >> 32) == 0) s = 1.0 - s;
    double theta_opt = 0.5 * std::atan2(sv2, sv1);
    double mt = 0.0;
    for (int i = 0; i < 4; ++i) {
        double f = cpbw_fully_resolved(n, a, eta, s, theta_opt + double(i) * FRAC_PI_2, sv1, sv2, sv3);
        mt = i == 0 ? f : std::fmax(mt, f);
    }
    mt = 1.01 * mt;
    double theta;
    for (;;) {
        theta = 2.0 * PI * rng.gen_f64();
        double z = mt * rng.gen_f64();
        if (z < cpbw_fully_resolved(n, a, eta, s, theta, sv1, sv2, sv3)) break;
    }
    *s_out = s;
    *theta_out = theta;
}
// pair_creation/cp/mod.rs:368-381
static void cpbw_sum_limits(double a, double eta, int* lo, int* hi) {
    double n_min = std::ceil(2.0 * (1.0 + a * a) / eta);
    double range = 2.0 * (0.25 + a * a) / eta + 13.0 * a * (2.0 + a);
    double test = 0.25 - (1.0 + a * a) / (2.0 * n_min * eta);
    if (test <= 2.220446049250313e-16) { *lo = int(n_min) + 1; *hi = int(n_min + 1.0 + range); }
    else { *lo = int(n_min); *hi = int(n_min + range); }
}
static bool cpbw_is_too_small(double a, double eta) { return std::log10(eta) < -1.0 - powi(std::log10(a) + 2.0, 2) / 4.5; }
// pair_creation/cp/mod.rs:398-404
static void cpbw_by_summation(double a, double eta, double out2[2]) {
    int lo, hi;
    cpbw_sum_limits(a, eta, &lo, &hi);
    C total{0.0, 0.0};
    for (int n = lo; n <= hi; ++n) total = total + cpbw_partial_rate(n, a, eta);
    out2[0] = total.re;
    out2[1] = total.im;
}
// pair_creation/cp/mod.rs:406-449
static void cpbw_by_lookup(double a, double eta, double out2[2]) {
    const size_t N_COLS = 74, N_ROWS = 121;
    double LMA_ = cpt(CBT, "LN_MIN_A")[0], LME = cpt(CBT, "LN_MIN_ETA")[0], LAS = cpt(CBT, "LN_A_STEP")[0], LES = cpt(CBT, "LN_ETA_STEP")[0];
    Table T = cpt(CBT, "TABLE");
    double x = std::log(a), y = std::log(eta);
    if (x < LMA_) { std::fprintf(stderr, "oracle: NBW [CP] rate lookup out of bounds\n"); std::abort(); }
    if (y < LME) { out2[0] = out2[1] = 0.0; return; }
    size_t ix = size_t((x - LMA_) / LAS), iy = size_t((y - LME) / LES);
    if (!(ix < N_COLS - 1 && iy < N_ROWS - 1)) { std::fprintf(stderr, "oracle: NBW [CP] rate lookup out of bounds\n"); std::abort(); }
    auto at = [&](size_t r, size_t c, int k) { return T[(r * N_COLS + c) * 2 + k]; };
    if (a * eta > 2.0) {
        double dx = (x - LMA_) / LAS - double(ix), dy = (y - LME) / LES - double(iy);
        double w[4] = {(1.0 - dx) * (1.0 - dy), dx * (1.0 - dy), (1.0 - dx) * dy, dx * dy};
        for (int k = 0; k < 2; ++k)
            out2[k] = std::exp(w[0] * at(iy, ix, k) + w[1] * at(iy, ix + 1, k) + w[2] * at(iy + 1, ix, k) + w[3] * at(iy + 1, ix + 1, k));
    } else {
        double a_min = std::exp(LMA_ + double(ix) * LAS), a_max = std::exp(LMA_ + double(ix + 1) * LAS);
        double e_min = std::exp(LME + double(iy) * LES), e_max = std::exp(LME + double(iy + 1) * LES);
        double dx = (a - a_min) / (a_max - a_min), dy = (eta - e_min) / (e_max - e_min);
        double w[4] = {(1.0 - dx) * (1.0 - dy), dx * (1.0 - dy), (1.0 - dx) * dy, dx * dy};
        for (int k = 0; k < 2; ++k) {
            double f = w[0] / at(iy, ix, k) + w[1] / at(iy, ix + 1, k) + w[2] / at(iy + 1, ix, k) + w[3] / at(iy + 1, ix + 1, k);
            out2[k] = std::exp(1.0 / f);
        }
    }
}
// (S_3 = +1, S_3 = -1) total rates (pair_creation/cp/mod.rs:271-281)
static void cpbw_total(double a, double eta, double f[2]) {
    if (a < 0.02 || cpbw_is_too_small(a, eta)) { f[0] = f[1] = 0.0; return; }
    if (a < std::exp(cpt(CBT, "LN_MIN_A")[0])) cpbw_by_summation(a, eta, f);
    else cpbw_by_lookup(a, eta, f);
}
// pair_creation/cp/mod.rs:286-322
double cpbw_probability(double a, double eta, SV sv, double dphi, SV* out) {
    double f[2];
    cpbw_total(a, eta, f);
    double prob = ALPHA_FINE * (0.5 * ((f[0] + f[1]) + sv[3] * (f[0] - f[1]))) * dphi / eta;
    double prob_avg = ALPHA_FINE * (0.5 * (f[0] + f[1])) * dphi / eta;
    double delta = ALPHA_FINE * (0.5 * (f[0] - f[1])) * dphi / eta;
    *out = SV{sv[0], sv[1] * (1.0 - prob_avg) / (1.0 - prob), sv[2] * (1.0 - prob_avg) / (1.0 - prob),
              (sv[3] * (1.0 - prob_avg) - delta) / (1.0 - prob)};
    return prob;
}
// pair_creation/cp/mod.rs:327-365: harmonic by explicit summation of partial rates
int cpbw_harmonic(double a, double eta, double sv3, double frac, bool* fell_back) {
    int lo, hi;
    cpbw_sum_limits(a, eta, &lo, &hi);
    double f[2];
    cpbw_total(a, eta, f);
    double target = 0.5 * (f[0] * (1.0 + sv3) + f[1] * (1.0 - sv3)) * frac;
    double cumsum = 0.0;
    for (int k = lo; k <= hi; ++k) {
        C pr = cpbw_partial_rate(k, a, eta);
        cumsum += 0.5 * ((1.0 + sv3) * pr.re + (1.0 - sv3) * pr.im);
        if (cumsum > target) { if (fell_back) *fell_back = false; return k; }
    }
    if (fell_back) *fell_back = true;
    return hi - 1;
}
void cpbw_sample(double a, double eta, double sv1, double sv2, double sv3, Xoshiro& rng, int* n_out, double* s_out, double* theta_out) {
    int n = cpbw_harmonic(a, eta, sv3, rng.gen_f64(), nullptr);
    cpbw_sample_at(n, a, eta, sv1, sv2, sv3, rng, s_out, theta_out);
    *n_out = n;
}

}  // namespace orc

// ---- probes
using namespace orc;
extern "C" {
void orc_bessel_j_pm(int n, double x, double* out3) { bessel_j_pm(n, x, out3); }
double orc_bessel_j(int n, double x) { return bessel_j(n, x); }
void orc_bessel_j_parts(int n, double x, double* out2) {  // recurrence and uniform expansion separately
    double t[3];
    triple_j(n, x, t);
    out2[0] = t[1];
    out2[1] = j_asymptotic(n, x);
}
double orc_cp_ddpr(int n, double a, double eta, double v) { return cp_diff_partial_rate(n, a, eta, v); }
// the grid maxima the reference's ceiling tests compare against: v = i / 10^4,
// i < 10^4 (cp/mod.rs:265-270), or v = 0.5 (1 + i / 100), i < 100 (classical.rs:204-208)
double orc_cp_grid_max(int n, double a, double eta, int classical) {
    double mx = 0.0;
    if (classical) {
        for (int i = 0; i < 100; ++i) {
            double f = cp_spectrum_low_eta(n, a, 0.5 * (1.0 + double(i) / 100.0));
            mx = i == 0 ? f : std::fmax(mx, f);
        }
    } else {
        for (int i = 0; i < 10000; ++i) {
            double f = cp_diff_partial_rate(n, a, eta, double(i) / 10000.0);
            mx = i == 0 ? f : std::fmax(mx, f);
        }
    }
    return mx;
}
double orc_cp_ceiling(int n, double a, double eta) { return cp_ceiling(n, a, eta); }
double orc_cp_partial_rate(int n, double a, double eta) { return cp_partial_rate(n, a, eta); }
double orc_cp_spectrum_low_eta(int n, double a, double v) { return cp_spectrum_low_eta(n, a, v); }
double orc_cp_integrated_spectrum_low_eta(int n, double a) { return cp_integrated_spectrum_low_eta(n, a); }
double orc_cp_classical_ceiling(int n, double a) { return cp_classical_ceiling(n, a); }
int orc_cp_rate(double a, double eta, double* out) { return cp_rate(a, eta, out) ? 1 : 0; }
int orc_cp_classical_rate(double a, double eta, double* out) { return cp_classical_rate(a, eta, out) ? 1 : 0; }
int orc_cp_harmonic(double a, double eta, double frac) { return has_table("nonlinear_compton/cp/tables/cdf:TABLE") ? cp_harmonic(a, eta, frac) : -2; }
int orc_cp_classical_harmonic(double a, double frac) { return cp_classical_harmonic(a, frac); }
double orc_cp_cdf_row(double a, double eta, int classical, double* out64) { return cp_cdf_row(a, eta, classical != 0, out64); }
// count samples of (n, s, theta, S1, S2, S3, trials) from one seeded stream
// Stokes parameters (1, S1, S2, S3) of the CP photon at harmonic n, fraction v of
// the largest s (quantum: s = v s_max; classical: s = v s_n) and ZMF azimuth theta
void orc_cp_stokes(int n, double a, double eta, double v, double theta, int classical, double* out4) {
    double sn = 2.0 * double(n) * eta / (1.0 + a * a);
    double s = classical ? v * sn : v * sn / (1.0 + sn);
    SV sv = cp_stokes(n, a, s, sn, v, theta, classical != 0);
    out4[0] = sv.i; out4[1] = sv.q; out4[2] = sv.u; out4[3] = sv.v;
}

void orc_cp_sample(double a, double eta, int classical, uint64_t seed, int fixed_n, size_t count, double* out7) {
    Xoshiro rng(seed);
    for (size_t i = 0; i < count; ++i) {
        int n;
        double s, th;
        SV sv;
        uint64_t tr = 0;
        if (classical) cp_classical_sample(a, eta, rng, fixed_n, &n, &s, &th, &sv, &tr);
        else cp_sample(a, eta, rng, fixed_n, &n, &s, &th, &sv, &tr);
        double* o = out7 + 7 * i;
        o[0] = n; o[1] = s; o[2] = th; o[3] = sv[1]; o[4] = sv[2]; o[5] = sv[3]; o[6] = double(tr);
    }
}
}

extern "C" {
void orc_cpbw_partial_rate(int n, double a, double eta, double* out2) { C r = cpbw_partial_rate(n, a, eta); out2[0] = r.re; out2[1] = r.im; }
double orc_cpbw_fully_resolved(int n, double a, double eta, double s, double theta, double sv1, double sv2, double sv3) {
    return cpbw_fully_resolved(n, a, eta, s, theta, sv1, sv2, sv3);
}
int orc_cpbw_sum_limits(double a, double eta, int* hi) { int lo; cpbw_sum_limits(a, eta, &lo, hi); return lo; }
// 0: zero (a < 0.02 or negligible), 1: by summation, 2: by lookup
int orc_cpbw_total(double a, double eta, double* out2) {
    cpbw_total(a, eta, out2);
    if (a < 0.02 || cpbw_is_too_small(a, eta)) return 0;
    return a < std::exp(cpt(CBT, "LN_MIN_A")[0]) ? 1 : 2;
}
void orc_cpbw_by_summation(double a, double eta, double* out2) { cpbw_by_summation(a, eta, out2); }
double orc_cpbw_probability(double a, double eta, const double* sv4, double dphi, double* sv_out4) {
    SV o;
    double p = cpbw_probability(a, eta, SV{sv4[0], sv4[1], sv4[2], sv4[3]}, dphi, &o);
    for (int k = 0; k < 4; ++k) sv_out4[k] = o[k];
    return p;
}
int orc_cpbw_harmonic(double a, double eta, double sv3, double frac) { return cpbw_harmonic(a, eta, sv3, frac, nullptr); }
void orc_cpbw_sample(double a, double eta, double sv1, double sv2, double sv3, uint64_t seed, int fixed_n, size_t count, double* out3) {
    Xoshiro rng(seed);
    for (size_t i = 0; i < count; ++i) {
        int n;
        double s, th;
        if (fixed_n > 0) { n = fixed_n; cpbw_sample_at(n, a, eta, sv1, sv2, sv3, rng, &s, &th); }
        else cpbw_sample(a, eta, sv1, sv2, sv3, rng, &n, &s, &th);
        out3[3 * i] = n; out3[3 * i + 1] = s; out3[3 * i + 2] = th;
    }
}
}
