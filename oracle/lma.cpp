// Ptarmigan CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.hpp).
// Locally-monochromatic (LMA) plane wave (src/field/plane_wave.rs), linearly
// polarised nonlinear Compton rates and sampler (src/nonlinear_compton/mod.rs,
// lp/mod.rs, lp/linear.rs) and the generalised Bessel function J(n, x, y)
// (src/special_functions/double_bessel.rs).
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "oracle.hpp"

namespace orc {

struct RadEvent { V4 k, u_prime; SV pol; double a_eff, chi, absorption; };
struct PairEvent { V4 u_e, u_p; double frac, a_eff, chi, absorption; };

// ---------------------------------------------------------------- complex functions (num-complex 0.4 formulas)
static C cmake(double re, double im) { return C{re, im}; }
static double carg(C a) { return std::atan2(a.im, a.re); }
static double cnorm(C a) { return std::hypot(a.re, a.im); }
static C cfrom_polar(double r, double th) { return {r * std::cos(th), r * std::sin(th)}; }
static C cln(C a) { return {std::log(cnorm(a)), carg(a)}; }
static C csqrt(C a) {
    if (a.im == 0.0) {
        if (!std::signbit(a.re)) return {std::sqrt(a.re), a.im};
        double im = std::sqrt(-a.re);
        return std::signbit(a.im) ? C{0.0, -im} : C{0.0, im};
    }
    if (a.re == 0.0) {
        double x = std::sqrt(std::fabs(a.im) / 2.0);
        return std::signbit(a.im) ? C{x, -x} : C{x, x};
    }
    return cfrom_polar(std::sqrt(cnorm(a)), carg(a) / 2.0);
}
static C cacos(C z) {  // -i ln(i sqrt(1 - z^2) + z)
    const C i{0.0, 1.0};
    C w = i * csqrt(1.0 - z * z) + z;
    return -i * cln(w);
}
static C csin(C a) { return {std::sin(a.re) * std::cosh(a.im), std::cos(a.re) * std::sinh(a.im)}; }
static C cexp2(C a) { return cfrom_polar(std::exp(a.re), a.im); }

// ---------------------------------------------------------------- DoubleBessel (double_bessel.rs)
struct Element { double xi[3], lambda[2], f, g, h; };

struct DoubleBessel {
    int n, k_min;
    double x_max, y_max;
    std::vector<Element> tbl;

    // double_bessel.rs:39-76
    DoubleBessel(int n_, double xm, double ym) : n(n_) {
        int n_minus_bound = int(std::floor(-2.0 * ym - xm));
        int n_plus_bound = xm > 4.0 * (2.0 - SQRT_2) * ym ? int(std::ceil(xm)) : int(std::ceil(2.0 * ym + xm * xm / (16.0 * ym)));
        size_t size = size_t(n_plus_bound - n_minus_bound + 2 * 100);
        tbl.assign(size, Element{{1.0, 1.0, 1.0}, {1.0, 1.0}, 0.0, 0.0, 0.0});
        k_min = n_minus_bound - 100;
        x_max = std::fabs(xm);
        y_max = std::fabs(ym);
    }

    // double_bessel.rs:142-246
    void around(double x, double y, double j[5]) {
        int n_minus = int(std::floor(-2.0 * y - x));
        int n_plus = 8.0 * y > x ? int(std::ceil(2.0 * y + x * x / (16.0 * y))) : int(std::ceil(-2.0 * y + x));
        size_t len = tbl.size();
        for (size_t i = 1; i < len; ++i) {
            int k = int(i) + k_min;
            Element& t = tbl[i];
            const Element& p = tbl[i - 1];
            t.xi[0] = -x - 4.0 * y * y / p.xi[2];
            t.xi[1] = 2.0 * double(k - 1) - 2.0 * y * p.xi[0] / p.xi[2];
            t.xi[2] = -x - 2.0 * y * p.xi[1] / p.xi[2];
            t.lambda[0] = t.xi[0] - 2.0 * y * t.xi[2] / p.lambda[1];
            t.lambda[1] = t.xi[1] - p.lambda[0] * t.xi[2] / p.lambda[1];
        }
        double initial = n < 2000 ? 1.0e-100 : 1.0e-300;
        tbl[len - 1].f = 0.0;
        tbl[len - 2].f = 1.0 * initial;
        tbl[len - 3].f = 2.0 * initial;
        tbl[len - 1].g = 0.0;
        tbl[len - 2].g = 1.0 / initial;
        tbl[len - 3].g = -(2.0 * y * tbl[len - 1].g + tbl[len - 2].lambda[0] * tbl[len - 2].g) / tbl[len - 2].lambda[1];
        int k_match = (n_plus + n_minus) / 2;
        size_t i_match = size_t(k_match - k_min);
        tbl[i_match].f = 0.0;
        tbl[i_match].g = 0.0;
        for (long i = long(len) - 4; i >= 0; --i) {
            Element& t = tbl[i];
            t.f = -(2.0 * y * tbl[i + 3].f + tbl[i + 2].xi[0] * tbl[i + 2].f + tbl[i + 2].xi[1] * tbl[i + 1].f) / tbl[i + 2].xi[2];
            t.g = -(2.0 * y * tbl[i + 2].g + tbl[i + 1].lambda[0] * tbl[i + 1].g) / tbl[i + 1].lambda[1];
            if (t.f > 1.0 && std::fabs(tbl[i_match].g) < 1.0e200) {
                for (size_t jj = size_t(i); jj < len; ++jj) {
                    tbl[jj].f *= 1.0e-100;
                    tbl[jj].g *= 1.0e100;
                }
            }
        }
        double f_match = tbl[i_match].f, g_match = tbl[i_match].g;
        double sum_linear = 0.0, rss = 0.0;
        for (size_t i = 0; i < i_match; ++i) {
            tbl[i].h = tbl[i].g;
            sum_linear += tbl[i].h;
            rss = std::hypot(rss, tbl[i].h);
        }
        for (size_t i = i_match; i < len; ++i) {
            tbl[i].h = g_match * tbl[i].f / f_match;
            sum_linear += tbl[i].h;
            rss = std::hypot(rss, tbl[i].h);
        }
        size_t it = size_t(n - k_min);
        for (int q = 0; q < 5; ++q) {
            const Element& t = tbl[it - 2 + q];
            double val = signum(t.h / sum_linear) * std::fabs(t.h) / rss;
            j[q] = std::isfinite(val) ? val : 0.0;
        }
    }

    // double_bessel.rs:272-355
    static double j_spa(int ni, double x, double y) {
        double n = double(ni);
        if (n <= -2.0 * y + x) {
            double cos_theta = y > 0.0 ? x / (8.0 * y) - std::sqrt(x * x / (64.0 * y * y) + 0.5 - n / (4.0 * y)) : n / x;
            double theta = std::acos(cos_theta);
            double f = x * std::sin(theta) - y * std::sin(2.0 * theta);
            double d2f = -x * std::sin(theta) + 4.0 * y * std::sin(2.0 * theta);
            double phi = (PI / 4.0) * signum(d2f);
            return std::cos(f - n * theta + phi) * std::sqrt(2.0 / (PI * std::fabs(d2f)));
        } else if (8.0 * y < x || n > 2.0 * y + x * x / (16.0 * y)) {
            double disc = x * x / (64.0 * y * y) + 0.5 - n / (4.0 * y);
            C cos_theta;
            if (std::isnan(disc)) cos_theta = {n / x, 0.0};
            else if (disc < 0.0) cos_theta = {x / (8.0 * y), std::sqrt(-disc)};
            else cos_theta = {x / (8.0 * y) - std::sqrt(disc), 0.0};
            C theta = cacos(cos_theta);
            const C i{0.0, 1.0};
            C f = i * (-n * theta + x * csin(theta) - y * csin(2.0 * theta));
            C d2f = i * (-x * csin(theta) + 4.0 * y * csin(2.0 * theta));
            C phi{0.0, (PI - carg(d2f)) / 2.0};
            C phase = f + phi;
            C term;
            if (f.im == 0.0 && d2f.im == 0.0) {
                double val = 0.5 * std::exp(-f.re) * std::sqrt(2.0 / (PI * cnorm(d2f)));
                term = {val, 0.0};
            } else {
                term = cexp2(phase) * std::sqrt(2.0 / (PI * cnorm(d2f)));
            }
            return disc != 0.0 ? term.re : 0.0;
        } else {
            double r = std::sqrt(x * x / (64.0 * y * y) + 0.5 - n / (4.0 * y));
            double ct[2] = {x / (8.0 * y) - r, x / (8.0 * y) + r};
            double out = 0.0;
            for (int k = 0; k < 2; ++k) {
                double theta = std::acos(ct[k]);
                double f = x * std::sin(theta) - y * std::sin(2.0 * theta);
                double d2f = -x * std::sin(theta) + 4.0 * y * std::sin(2.0 * theta);
                double phi = (PI / 4.0) * signum(d2f);
                out += std::cos(f - n * theta + phi) * std::sqrt(2.0 / (PI * std::fabs(d2f)));
            }
            return out;
        }
    }

    // double_bessel.rs:386-409
    bool use_saddle_point(double x, double y) const {
        if (n <= 10) return false;
        double ne = double(n);
        ne = ne * (1.0 - std::sqrt(10.0 / ne));
        double x_crit = y > ne / 6.0 ? 4.0 * std::sqrt(y * (ne - 2.0 * y)) : ne + 2.0 * y;
        return x < x_crit;
    }

    void evaluate(double x, double y, double j[5]) {
        if (use_saddle_point(x, y))
            for (int q = 0; q < 5; ++q) j[q] = j_spa(n - 2 + q, x, y);
        else
            around(x, y, j);
    }
};

// ---------------------------------------------------------------- LP nonlinear Compton (lp/mod.rs)
static const char* RT = "nonlinear_compton/lp/rate_table:";
static const char* CT = "nonlinear_compton/lp/cdf_table:";
static Table tb(const char* base, const char* n) { return table((std::string(base) + n).c_str()); }

// lp/linear.rs:6-45
static double linear_rate(double a, double eta) {
    double eta2 = eta * eta, eta3 = eta2 * eta;
    if (eta > 1.0e-3)
        return (a * a / 8.0) * (0.5 + 4.0 / eta - 0.5 / ((1.0 + 2.0 * eta) * (1.0 + 2.0 * eta)) +
                                (1.0 - 2.0 / eta - 2.0 / eta2) * std::log(1.0 + 2.0 * eta));
    return (a * a / 8.0) * (8.0 * eta / 3.0 - 16.0 * eta2 / 3.0 + 208.0 * eta3 / 15.0);
}

// lp/mod.rs:430-463. Returns false on an out-of-table lookup (reference prints and returns None).
static bool lp_rate(double a, double eta, double* out) {
    Table MIN = tb(RT, "MIN"), STEP = tb(RT, "STEP"), T = tb(RT, "TABLE");
    const int N_COLS = 61, N_ROWS = 68;
    double x = std::log(a), y = std::log(eta);
    if (x < MIN[0]) { *out = linear_rate(a, eta); return true; }
    if (y < MIN[1]) {
        size_t ix = size_t((x - MIN[0]) / STEP[0]);
        // beyond the last column the reference's TABLE[0][ix + 1] panics (index out
        // of bounds); report it as outside the table
        if (ix >= size_t(N_COLS - 1)) return false;
        double dx = (x - MIN[0]) / STEP[0] - double(ix);
        double f = (1.0 - dx) * T[ix] + dx * T[ix + 1];
        *out = std::exp(f - MIN[1]) * eta;
        return true;
    }
    size_t ix = size_t((x - MIN[0]) / STEP[0]);
    size_t iy = size_t((y - MIN[1]) / STEP[1]);
    if (ix < size_t(N_COLS - 1) && iy < size_t(N_ROWS - 1)) {
        double dx = (x - MIN[0]) / STEP[0] - double(ix);
        double dy = (y - MIN[1]) / STEP[1] - double(iy);
        double f = (1.0 - dx) * (1.0 - dy) * T[iy * N_COLS + ix] + dx * (1.0 - dy) * T[iy * N_COLS + ix + 1] +
                   (1.0 - dx) * dy * T[(iy + 1) * N_COLS + ix] + dx * dy * T[(iy + 1) * N_COLS + ix + 1];
        *out = std::exp(f);
        return true;
    }
    return false;
}

// lp/mod.rs:414-419
static int sum_limit(double a, double eta) {
    double m = 3.0 - 0.4 * std::pow(eta, 0.25);
    return int(std::ceil(6.0 * (1.0 + a * a) + 2.0 * std::pow(a, m)));
}

// harmonic from one 16-point CDF table (lp/mod.rs:465-553)
static double harmonic_from_table(const double* t16, double a, double frac) {
    if (frac <= t16[1]) return 0.9;
    double n;
    if (a < 5.0) {
        Interp{t16, 16}.invert(frac, &n);
        return n;
    }
    double rs[30];
    for (int i = 0; i < 15; ++i) {
        rs[2 * i] = std::log(t16[2 * i]);
        rs[2 * i + 1] = std::log(-1.0 * std::log(1.0 - t16[2 * i + 1]));
    }
    double f2 = std::log(-1.0 * std::log(1.0 - frac));
    Interp{rs, 15, true}.invert(f2, &n);
    return std::exp(n);
}

static int get_harmonic_index(double a, double eta, double frac) {
    Table MIN = tb(CT, "MIN"), STEP = tb(CT, "STEP"), T = tb(CT, "TABLE");
    const int N_COLS = 61;
    double la = std::log(a), le = std::log(eta);
    if (la <= MIN[0]) return 1;
    if (le <= MIN[1]) {
        size_t ix = size_t((la - MIN[0]) / STEP[0]);
        double dx = (la - MIN[0]) / STEP[0] - double(ix);
        double n = 0.0;
        size_t idx[2] = {ix, ix + 1};
        double w[2] = {1.0 - dx, dx};
        for (int k = 0; k < 2; ++k) n += harmonic_from_table(T.p + 32 * idx[k], a, frac) * w[k];
        return int(std::ceil(n));
    }
    size_t ix = size_t((la - MIN[0]) / STEP[0]);
    size_t iy = size_t((le - MIN[1]) / STEP[1]);
    double dx = (la - MIN[0]) / STEP[0] - double(ix);
    double dy = (le - MIN[1]) / STEP[1] - double(iy);
    size_t idx[4] = {N_COLS * iy + ix, N_COLS * iy + ix + 1, N_COLS * (iy + 1) + ix, N_COLS * (iy + 1) + ix + 1};
    double w[4] = {(1.0 - dx) * (1.0 - dy), dx * (1.0 - dy), (1.0 - dx) * dy, dx * dy};
    double n = 0.0;
    for (int k = 0; k < 4; ++k) n += harmonic_from_table(T.p + 32 * idx[k], a, frac) * w[k];
    return int(std::ceil(n));
}

// lp/mod.rs:21-50
static double ddpr(double a, double eta, double s, double theta, DoubleBessel& dj) {
    int n = dj.n;
    double sn = 2.0 * double(n) * eta / (1.0 + 0.5 * a * a);
    double wn = std::fmin(s / (sn * (1.0 - s)), 1.0);
    double alpha = 2.0 * double(n) * a * std::cos(theta) * std::sqrt(wn * (1.0 - wn)) / std::sqrt(1.0 + 0.5 * a * a);
    double beta = a * a * s / (8.0 * eta * (1.0 - s));
    double j[5];
    dj.evaluate(alpha, beta, j);
    double g0, g1, g2;
    if (n % 2 == 0) { g0 = j[2]; g1 = -0.5 * (j[1] + j[3]); g2 = 0.25 * (j[0] + 2.0 * j[2] + j[4]); }
    else { g0 = -j[2]; g1 = 0.5 * (j[1] + j[3]); g2 = -0.25 * (j[0] + 2.0 * j[2] + j[4]); }
    return (-g0 * g0 - a * a * (1.0 + 0.5 * s * s / (1.0 - s)) * (g0 * g2 - g1 * g1)) / (2.0 * PI);
}

// lp/mod.rs:53-140
static double ceiling_ddpr(double a, double eta, DoubleBessel& dj) {
    int n = dj.n;
    double sn = 2.0 * double(n) * eta / (1.0 + 0.5 * a * a);
    double smax = sn / (1.0 + sn);
    double mx;
    auto scan = [&](int lo, int hi, auto sfun, double theta) {
        double m = -INFINITY;
        for (int i = lo; i <= hi; ++i) m = std::fmax(m, ddpr(a, eta, sfun(i), theta, dj));
        return m;
    };
    if (n == 1) {
        double theta = FRAC_PI_2;
        mx = std::fmax(ddpr(a, eta, 0.01 * smax, theta, dj), ddpr(a, eta, 0.99 * smax, theta, dj));
    } else if (a <= 2.0) {
        mx = scan(1, 32, [&](int i) { return smax * (0.4 + 0.6 * double(i) / 32.0); }, 0.0);
    } else if (a <= 5.0) {
        int n_switch = int(std::ceil(1.3 * std::pow(a, 1.2)));
        if (n < n_switch) {
            double theta = FRAC_PI_2 * (1.0 - std::log(double(n)) / std::log(double(n_switch)));
            mx = scan(0, 32, [&](int i) { return smax * double(i) / 32.0; }, theta);
        } else {
            double x_min = std::fmax(4.0 * double(n) * eta / (3.0 * a * a + 4.0 * double(n) * eta), 0.4);
            int n_max = sum_limit(a, eta);
            int m = n < n_max / 2 ? 32 : 64;
            mx = scan(1, m, [&](int i) { return smax * (x_min + (1.0 - x_min) * double(i) / double(m)); }, 0.0);
        }
    } else {
        int n_switch = int(std::ceil(a * a));
        if (n < n_switch) {
            double nn = double(n), ns = double(n_switch);
            double theta = FRAC_PI_2 * (std::pow(nn, -0.7) - std::pow(ns, -0.7)) / (1.0 - std::pow(ns, -0.7));
            mx = scan(1, 64, [&](int i) { return smax * (0.3 + 0.7 * double(i) / 64.0); }, theta);
        } else {
            double x_min = std::fmax(4.0 * double(n) * eta / (3.0 * a * a + 4.0 * double(n) * eta), 0.3);
            double x_peak = (1.0 + sn) / (2.0 + sn);
            int len = n > 10 * n_switch ? 64 : (n > 3 * n_switch ? 42 : 32);
            mx = scan(1, len, [&](int i) { return smax * (x_min + (x_peak - x_min) * double(i) / double(len)); }, 0.0);
        }
    }
    return 1.15 * mx;
}

// lp/mod.rs:143-301
struct ThetaBound {
    double s[16], f[16];
    static double ln_double_bessel_x_zero(int ni, double y) {
        double n = double(ni);
        double z = std::sqrt(n / (4.0 * y) - 0.5);
        C theta{FRAC_PI_2, std::log(std::sqrt(1.0 + z * z) - z)};
        const C i{0.0, 1.0};
        C f = i * (-n * theta - y * csin(2.0 * theta));
        C f2 = i * 4.0 * y * csin(2.0 * theta);
        double e = 0.5 * (PI - carg(f2));
        C phase = f + i * e;
        return std::log(std::sqrt(2.0 / (PI * cnorm(f2)))) + phase.re + std::log(std::fabs(std::cos(phase.im)));
    }
    static ThetaBound for_harmonic(int n, double a, double eta) {
        ThetaBound tb_;
        double sn = 2.0 * double(n) * eta / (1.0 + 0.5 * a * a);
        for (int i = 0; i < 16; ++i) {
            double z = std::cos(PI * double(i) / 15.0);
            tb_.s[i] = 0.5 * (z + 1.0) * sn / (1.0 + sn);
            double wn = std::fmin(tb_.s[i] / (sn * (1.0 - tb_.s[i])), 1.0);
            double x = 2.0 * double(n) * a * std::sqrt(wn * (1.0 - wn)) / std::sqrt(1.0 + 0.5 * a * a);
            double y = a * a * tb_.s[i] / (8.0 * eta * (1.0 - tb_.s[i]));
            double x_crit = y > double(n) / 6.0 ? 4.0 * std::sqrt(y) * std::sqrt(double(n) - 2.0 * y) : double(n) + 2.0 * y;
            double ln_j_crit = std::log(0.443) - std::log(double(n)) / 3.0;
            double ln_j_bdy = n % 2 == 0 ? ln_double_bessel_x_zero(n, y) : ln_double_bessel_x_zero(n + 1, y);
            ln_j_bdy = std::fmin(ln_j_bdy, ln_j_crit);
            double cos_theta = 1.0 - x_crit * std::log(0.01) / (x * (ln_j_bdy - ln_j_crit));
            cos_theta = std::fmax(cos_theta, -1.0);
            tb_.f[i] = std::acos(cos_theta);
            if (std::isnan(tb_.f[i])) tb_.f[i] = PI;
        }
        double f_min = tb_.f[0];
        for (int i = 1; i < 16; ++i) {
            if (tb_.f[i] < f_min) f_min = tb_.f[i];
            else tb_.f[i] = f_min;
        }
        return tb_;
    }
    // lp/mod.rs:211-284: the same construction at fixed v = s / s_n in the classical limit
    static ThetaBound for_harmonic_low_eta(int n, double a) {
        ThetaBound tb_;
        for (int i = 0; i < 16; ++i) {
            double z = std::cos(PI * double(i) / 15.0);
            double v = 0.5 * (z + 1.0);
            tb_.s[i] = v;
            double x = 2.0 * double(n) * a * std::sqrt(v * (1.0 - v) / (1.0 + 0.5 * a * a));
            double y = 0.25 * double(n) * a * a * v / (1.0 + 0.5 * a * a);
            double x_crit = y > double(n) / 6.0 ? 4.0 * std::sqrt(y) * std::sqrt(double(n) - 2.0 * y) : double(n) + 2.0 * y;
            double ln_j_crit = std::log(0.443) - std::log(double(n)) / 3.0;
            double ln_j_bdy = n % 2 == 0 ? ln_double_bessel_x_zero(n, y) : ln_double_bessel_x_zero(n + 1, y);
            ln_j_bdy = std::fmin(ln_j_bdy, ln_j_crit);
            double cos_theta = 1.0 - x_crit * std::log(0.01) / (x * (ln_j_bdy - ln_j_crit));
            cos_theta = std::fmax(cos_theta, -1.0);
            tb_.f[i] = std::acos(cos_theta);
            if (std::isnan(tb_.f[i])) tb_.f[i] = PI;
        }
        double f_min = tb_.f[0];
        for (int i = 1; i < 16; ++i) {
            if (tb_.f[i] < f_min) f_min = tb_.f[i];
            else tb_.f[i] = f_min;
        }
        return tb_;
    }
    double at(double arg) const {
        double val = f[15];
        for (int i = 1; i < 16; ++i) {
            if (arg > s[i]) {
                double weight = (arg - s[i - 1]) / (s[i] - s[i - 1]);
                val = weight * f[i] + (1.0 - weight) * f[i - 1];
                break;
            }
        }
        return std::fmin(val, FRAC_PI_2);
    }
};

// lp/mod.rs:304-410: Gauss-Legendre integral of ddpr over 0 < s < s_max, 0 < theta < 2 pi,
// and the largest integrand value met. Used only to check the rate and ceiling tables.
static double partial_rate(int n, double a, double eta, double* max_out) {
    Table G16N = table("quadrature/tables:GAUSS_16_NODES"), G16W = table("quadrature/tables:GAUSS_16_WEIGHTS");
    Table G32N = table("quadrature/tables:GAUSS_32_NODES"), G32W = table("quadrature/tables:GAUSS_32_WEIGHTS");
    double sn = 2.0 * double(n) * eta / (1.0 + 0.5 * a * a);
    double smax = sn / (1.0 + sn);
    double s_peak = sn / (2.0 + sn);
    DoubleBessel dj(n, double(n) * SQRT_2, double(n) * 0.5);
    ThetaBound tmax = ThetaBound::for_harmonic(n, a, eta);
    double mx = 0.0;
    auto single = [&](double s) {  // lp/mod.rs:306-319
        double th_max = tmax.at(s), sum = 0.0;
        for (int i = 0; i < 32; ++i) {
            double th = 0.5 * (G32N[i] + 1.0) * th_max;
            double r = ddpr(a, eta, s, th, dj);
            sum += 4.0 * (0.5 * th_max) * G32W[i] * r;
            mx = std::fmax(mx, r);
        }
        return sum;
    };
    auto segment = [&](double s0, double s1, const Table& N, const Table& W, int m) {
        double sum = 0.0;
        for (int i = 0; i < m; ++i) sum += W[i] * 0.5 * (s1 - s0) * single(s0 + 0.5 * (s1 - s0) * (N[i] + 1.0));
        return sum;
    };
    double total;
    if (sn < 1.0) {
        total = segment(0.0, s_peak, G32N, G32W, 32);
        total += segment(s_peak, smax, G32N, G32W, 32);
    } else {
        double s1 = smax - 2.0 * (smax - s_peak);
        total = segment(0.0, s1, G16N, G16W, 16);
        total += segment(s1, s_peak, G32N, G32W, 32);
        total += segment(s_peak, smax, G32N, G32W, 32);
    }
    if (max_out) *max_out = mx;
    return total;
}

// lp/mod.rs:586-676
// Stokes parameters of the photon at harmonic n, (s, theta), theta in [0, 2 pi)
// (lp/mod.rs:622-671); *xi0_out (if given) receives the unnormalised xi_0
static SV lp_stokes(int n, double a, double eta, double s, double theta, DoubleBessel& dj, double* xi0_out) {
    double sn = 2.0 * double(n) * eta / (1.0 + 0.5 * a * a);
    double wn = std::fmin(s / (sn * (1.0 - s)), 1.0);
    double x = -2.0 * double(n) * a * std::cos(theta) * std::sqrt(wn * (1.0 - wn)) / std::sqrt(1.0 + 0.5 * a * a);
    double y = a * a * s / (8.0 * eta * (1.0 - s));
    double j[5];
    dj.evaluate(std::fabs(x), y, j);
    double A0, A1, A2;
    if (x > 0.0) { A0 = j[2]; A1 = 0.5 * (j[1] + j[3]); A2 = 0.25 * (j[0] + 2.0 * j[2] + j[4]); }
    else if (n % 2 == 0) { A0 = j[2]; A1 = -0.5 * (j[1] + j[3]); A2 = 0.25 * (j[0] + 2.0 * j[2] + j[4]); }
    else { A0 = -j[2]; A1 = 0.5 * (j[1] + j[3]); A2 = -0.25 * (j[0] + 2.0 * j[2] + j[4]); }
    double u = std::sqrt(2.0 * double(n) * eta * (1.0 - s) / s - (1.0 + 0.5 * a * a));
    double xi0 = (1.0 - s + 1.0 / (1.0 - s)) * (powi(A1, 2) - A0 * A2) - 2.0 * powi(A0 / a, 2);
    double xi1 = 2.0 * (powi(A1, 2) - A0 * A2) - 2.0 * (1.0 + 2.0 * u * u * powi(std::sin(theta), 2)) * powi(A0 / a, 2);
    double xi2 = -1.0 * (2.0 * powi(u * A0 / a, 2) * std::sin(2.0 * theta) + 4.0 * u * A0 * A1 * std::sin(theta) / a);
    if (xi0_out) *xi0_out = xi0;
    return {1.0, xi1 / xi0, xi2 / xi0, 0.0};
}

static void lp_sample(double a, double eta, Xoshiro& rng, int fixed_n, int* n_out, double* s_out, double* theta_out, SV* pol_out,
                      uint64_t* trials) {
    int n = fixed_n > 0 ? fixed_n : get_harmonic_index(a, eta, rng.gen_f64());
    DoubleBessel dj(n, double(n) * SQRT_2, double(n) * 0.5);
    ThetaBound tmax = ThetaBound::for_harmonic(n, a, eta);
    double sn = 2.0 * double(n) * eta / (1.0 + 0.5 * a * a);
    double smax = sn / (1.0 + sn);
    double mx = ceiling_ddpr(a, eta, dj);
    double s, theta;
    for (;;) {
        s = smax * rng.gen_f64();
        theta = FRAC_PI_2 * rng.gen_f64();
        if (trials) ++*trials;
        if (theta > tmax.at(s)) continue;
        double z = mx * rng.gen_f64();
        double f = ddpr(a, eta, s, theta, dj);
        if (z < f) break;
    }
    // rand 0.7 gen_range(0, 4) for u32: widening multiply of one 32-bit draw
    uint32_t r32 = uint32_t(rng.next_u64() >> 32);
    int quadrant = int((uint64_t(r32) * 4u) >> 32);
    switch (quadrant) {
    case 1: theta = PI - theta; break;
    case 2: theta = PI + theta; break;
    case 3: theta = 2.0 * PI - theta; break;
    default: break;
    }
    *n_out = n;
    *s_out = s;
    *theta_out = theta;
    *pol_out = lp_stokes(n, a, eta, s, theta, dj, nullptr);
}

// ---------------------------------------------------------------- LP nonlinear Compton, classical (lp/classical.rs)
static const char* LCR = "nonlinear_compton/lp/classical/rate_table:";
static const char* LCC = "nonlinear_compton/lp/classical/cdf_table:";

// lp/classical.rs:17-40: rate differential in v = s / s_n and theta, 0 < theta < pi/2
static double lp_cl_ddpr(double a, double v, double theta, DoubleBessel& dj) {
    int n = dj.n;
    double x = 2.0 * double(n) * a * std::cos(theta) * std::sqrt(v * (1.0 - v) / (1.0 + 0.5 * a * a));
    double y = 0.25 * double(n) * a * a * v / (1.0 + 0.5 * a * a);
    double j[5];
    dj.evaluate(x, y, j);
    double g0, g1, g2;
    if (n % 2 == 0) { g0 = j[2]; g1 = -0.5 * (j[1] + j[3]); g2 = 0.25 * (j[0] + 2.0 * j[2] + j[4]); }
    else { g0 = -j[2]; g1 = 0.5 * (j[1] + j[3]); g2 = -0.25 * (j[0] + 2.0 * j[2] + j[4]); }
    return (-g0 * g0 - a * a * (g0 * g2 - g1 * g1)) / (2.0 * PI);
}
// lp/classical.rs:43-110
static double lp_cl_ceiling(double a, DoubleBessel& dj) {
    int n = dj.n;
    double mx = 0.0;
    if (n == 1) {
        mx = lp_cl_ddpr(a, 0.01, FRAC_PI_2, dj);
    } else if (a <= 1.5) {
        double v_min = 0.1 + (9.0 / std::sqrt(double(n)) - 1.0) / 10.0;
        double v_max = std::fmin(v_min + 0.4, 1.0);
        for (int i = 1; i <= 16; ++i) {
            double f = lp_cl_ddpr(a, v_min + (v_max - v_min) * double(i) / 16.0, 0.0, dj);
            mx = i == 1 ? f : std::fmax(mx, f);
        }
    } else if (a < 5.0) {
        int n_switch = int(3.0 * std::pow(a / 2.0, 4.0 / 3.0));
        double theta = n > n_switch ? 0.0 : FRAC_PI_2 * (1.0 - std::log(double(n)) / std::log(double(n_switch)));
        double v_min = std::fmin(0.35 + (7.0 / std::sqrt(double(n)) - 1.0) / 9.0, 0.55);
        double v_max = 0.55 + (7.0 / std::sqrt(double(n)) - 1.0) / 9.0;
        int num = a > 4.0 ? 32 : 16;
        for (int i = 0; i < num; ++i) {
            double f = lp_cl_ddpr(a, v_min + (v_max - v_min) * double(i) / double(num), theta, dj);
            mx = i == 0 ? f : std::fmax(mx, f);
        }
    } else {
        double n_switch = 10.0 * powi(a / 5.0, 2);
        double theta = 0.0;
        if (!(n > int(n_switch))) {
            double pw = 0.3 + (0.7 - 0.3) * (std::log(a) - std::log(5.0)) / (std::log(20.0) - std::log(5.0));
            theta = FRAC_PI_2 * (std::pow(double(n), -pw) - std::pow(n_switch, -pw)) / (1.0 - std::pow(n_switch, -pw));
        }
        double v_min = 0.3 + 0.4 * std::log(double(n)) * std::pow(double(n), -0.7);
        double v_max = v_min + 0.2;
        int num = a > 8.0 ? 64 : 32;
        for (int i = 0; i < num; ++i) {
            double f = lp_cl_ddpr(a, v_min + (v_max - v_min) * double(i) / double(num), theta, dj);
            mx = i == 0 ? f : std::fmax(mx, f);
        }
    }
    return 1.1 * mx;
}
// lp/classical.rs:112-181: integral over 0 < v < 1 (three 32-point pieces) and
// 0 < theta < 2 pi, and the largest integrand value met
static double lp_cl_partial_rate(int n, double a, double* max_out) {
    Table N = table("quadrature/tables:GAUSS_32_NODES"), W = table("quadrature/tables:GAUSS_32_WEIGHTS");
    DoubleBessel dj(n, double(n) * SQRT_2, double(n) * 0.5);
    ThetaBound tm = ThetaBound::for_harmonic_low_eta(n, a);
    auto single = [&](double v, double th_max, double* mx) {
        double sum = 0.0;
        for (int i = 0; i < 32; ++i) {
            double r = lp_cl_ddpr(a, v, 0.5 * (N[i] + 1.0) * th_max, dj);
            sum += 4.0 * (0.5 * th_max) * W[i] * r;
            *mx = std::fmax(*mx, r);
        }
        return sum;
    };
    double mxl = 0.0, mxm = 0.0, mxu = 0.0, lo = 0.0, mid = 0.0, up = 0.0;
    for (int i = 0; i < 32; ++i) {
        double v = 0.5 * (N[i] + 1.0) * 0.5;
        double m = 0.0;
        lo += W[i] * (0.5 * 0.5) * single(v, tm.at(v), &m);
        mxl = std::fmax(mxl, m);
    }
    for (int i = 0; i < 32; ++i) {
        double v = 0.5 + 0.5 * (N[i] + 1.0) * 0.25;
        double m = 0.0;
        mid += W[i] * (0.5 * 0.25) * single(v, tm.at(v), &m);
        mxm = std::fmax(mxm, m);
    }
    for (int i = 0; i < 32; ++i) {
        double v = 0.75 + 0.5 * (N[i] + 1.0) * 0.25;
        double m = 0.0;
        up += W[i] * (0.5 * 0.25) * single(v, tm.at(v), &m);
        mxu = std::fmax(mxu, m);
    }
    if (max_out) *max_out = std::fmax(std::fmax(mxu, mxm), mxl);
    return lo + mid + up;
}
// lp/classical.rs:189-213
static bool lp_cl_rate(double a, double eta, double* out) {
    double x = std::log(a);
    double MIN = tb(LCR, "MIN")[0], STEP = tb(LCR, "STEP")[0];
    Table T = tb(LCR, "TABLE");
    if (x < MIN) { *out = a * a * eta / 3.0; return true; }
    size_t ix = size_t((x - MIN) / STEP);
    if (!(ix < T.n - 1)) return false;
    double dx = (x - MIN) / STEP - double(ix);
    double f = (1.0 - dx) * T[ix] + dx * T[ix + 1];
    double prefactor = 2.0 * eta / (1.0 + 0.5 * a * a);
    *out = prefactor * std::exp(f);
    return true;
}
// lp/classical.rs:215-252
static int lp_cl_harmonic(double a, double frac) {
    double MIN = tb(LCC, "MIN")[0], STEP = tb(LCC, "STEP")[0];
    Table T = tb(LCC, "TABLE");
    if (std::log(a) <= MIN) return 1;
    size_t ix = size_t((std::log(a) - MIN) / STEP);
    double dx = (std::log(a) - MIN) / STEP - double(ix);
    if (ix + 1 >= 61) { std::fprintf(stderr, "oracle: LP classical CDF index\n"); std::abort(); }
    double n = 0.0;
    for (int k = 0; k < 2; ++k) {
        const double* t = T.p + 32 * (ix + k);
        double nk;
        if (frac <= t[1]) nk = 0.9;
        else if (!Interp{t, 16}.invert(frac, &nk)) { std::fprintf(stderr, "oracle: LP classical invert failed\n"); std::abort(); }
        n += nk * (k == 0 ? 1.0 - dx : dx);
    }
    return int(std::ceil(n));
}
// lp/classical.rs:254-333
// Stokes parameters of the classical LP photon at harmonic n, (v, theta), theta in
// [0, 2 pi) (lp/classical.rs:294-330; J(n, x, y) at the signed x with the parity
// applied on top, as the reference does)
static SV lp_cl_stokes(int n, double a, double v, double theta, DoubleBessel& dj) {
    double x = 2.0 * double(n) * a * std::cos(theta) * std::sqrt(v * (1.0 - v) / (1.0 + 0.5 * a * a));
    double y = 0.25 * double(n) * a * a * v / (1.0 + 0.5 * a * a);
    double j[5];
    dj.evaluate(x, y, j);
    double A0, A1, A2;
    if (n % 2 == 0) { A0 = j[2]; A1 = -0.5 * (j[1] + j[3]); A2 = 0.25 * (j[0] + 2.0 * j[2] + j[4]); }
    else { A0 = -j[2]; A1 = 0.5 * (j[1] + j[3]); A2 = -0.25 * (j[0] + 2.0 * j[2] + j[4]); }
    double u = std::sqrt((1.0 + 0.5 * a * a) * (1.0 - v) / v);
    double xi0 = 2.0 * (powi(A1, 2) - A0 * A2) - 2.0 * powi(A0 / a, 2);
    double xi1 = 2.0 * (powi(A1, 2) - A0 * A2) - 2.0 * (1.0 + 2.0 * u * u * powi(std::sin(theta), 2)) * powi(A0 / a, 2);
    double xi2 = 2.0 * powi(u * A0 / a, 2) * std::sin(2.0 * theta) + 4.0 * u * A0 * A1 * std::sin(theta) / a;
    return {1.0, xi1 / xi0, xi2 / xi0, 0.0};
}

static void lp_cl_sample(double a, double eta, Xoshiro& rng, int fixed_n, int* n_out, double* s_out, double* theta_out, SV* pol_out,
                         uint64_t* trials) {
    int n = fixed_n > 0 ? fixed_n : lp_cl_harmonic(a, rng.gen_f64());
    DoubleBessel dj(n, double(n) * SQRT_2, double(n) * 0.5);
    ThetaBound tm = ThetaBound::for_harmonic_low_eta(n, a);
    double smax = 2.0 * double(n) * eta / (1.0 + 0.5 * a * a);
    double mx = lp_cl_ceiling(a, dj);
    double v, s, theta;
    for (;;) {
        v = rng.gen_f64();
        s = v * smax;
        theta = FRAC_PI_2 * rng.gen_f64();
        if (theta > tm.at(v)) continue;
        double z = mx * rng.gen_f64();
        if (trials) ++*trials;
        if (z < lp_cl_ddpr(a, v, theta, dj)) break;
    }
    uint32_t r32 = uint32_t(rng.next_u64() >> 32);
    int quadrant = int((uint64_t(r32) * 4u) >> 32);
    switch (quadrant) {
    case 1: theta = PI - theta; break;
    case 2: theta = PI + theta; break;
    case 3: theta = 2.0 * PI - theta; break;
    default: break;
    }
    *n_out = n;
    *s_out = s;
    *theta_out = theta;
    *pol_out = lp_cl_stokes(n, a, v, theta, dj);
}

// nonlinear_compton/mod.rs:19-47
static bool nlc_probability(V4 k, V4 q, double dt, Pol pol, bool classical, double* out) {
    double a_sqd = q * q - 1.0;
    double a = a_sqd >= 0.0 ? std::sqrt(a_sqd) : 0.0;
    double eta = k * q;
    double dphi = dt * eta / (COMPTON_TIME * q[0]);
    double f;
    bool ok;
    if (pol == Pol::Circular) ok = classical ? cp_classical_rate(a, eta, &f) : cp_rate(a, eta, &f);
    else ok = classical ? lp_cl_rate(a * SQRT_2, eta, &f) : lp_rate(a * SQRT_2, eta, &f);
    if (!ok) return false;
    *out = ALPHA_FINE * f * dphi / eta;
    return true;
}

// nonlinear_compton/mod.rs:58-138
static void nlc_generate(V4 k, V4 q, Pol pol, double pol_angle, bool classical, Xoshiro& rng, int* n_out, V4* k_out, SV* sv_out,
                         uint64_t* trials) {
    double a = std::sqrt(q * q - 1.0);
    double eta = k * q;
    int n;
    double s, cphi;
    SV sv;
    if (pol == Pol::Circular) {
        if (classical) cp_classical_sample(a, eta, rng, 0, &n, &s, &cphi, &sv, trials);
        else cp_sample(a, eta, rng, 0, &n, &s, &cphi, &sv, trials);
    } else {
        if (classical) lp_cl_sample(a * SQRT_2, eta, rng, 0, &n, &s, &cphi, &sv, trials);
        else lp_sample(a * SQRT_2, eta, rng, 0, &n, &s, &cphi, &sv, trials);
    }
    double ell = double(n);
    double p_zmf = classical ? ell * eta / std::sqrt(1.0 + a * a) : ell * eta / std::sqrt(1.0 + a * a + 2.0 * ell * eta);
    double cos_theta_zmf = classical ? 1.0 - s * (1.0 + a * a) / (ell * eta) : 1.0 - s * (1.0 + a * a + 2.0 * ell * eta) / (ell * eta);
    V4 u_zmf;
    if (classical) u_zmf = q / std::sqrt(q * q);
    else { V4 qk = q + ell * k; u_zmf = qk / std::sqrt(qk * qk); }
    V3 along = classical ? normalize(spatial(boost_by(ell * k, u_zmf))) : -normalize(spatial(boost_by(q, u_zmf)));
    V4 eps4{0.0, std::cos(pol_angle), std::sin(pol_angle), 0.0};
    V3 eps = normalize(spatial(boost_by(eps4, u_zmf)));
    eps = normalize(cross(along, cross(eps, along)));
    V3 perp = rotate_around(eps, along, cphi);
    V3 kp = p_zmf * (cos_theta_zmf * along + std::sqrt(1.0 - powi(cos_theta_zmf, 2)) * perp);
    V4 k4 = lightlike(kp.x, kp.y, kp.z);
    V4 rev{u_zmf.t, -u_zmf.x, -u_zmf.y, -u_zmf.z};
    k4 = boost_by(k4, rev);
    sv = rotate_by(sv, -pol_angle);
    sv = from_lma_basis(sv, k4);
    *n_out = n;
    *k_out = k4;
    *sv_out = sv;
}

// ---------------------------------------------------------------- PlaneWave (plane_wave.rs)
static double pw_norm(const Laser& L) { return L.pol == Pol::Linear ? 0.5 : 1.0; }

// plane_wave.rs:68-106
double pw_a_sqd(const Laser& L, V4 r) {
    double norm = pw_norm(L);
    double phase = L.wavevector * r;
    switch (L.env) {
    case Env::CosSquared:
        return std::fabs(phase) < PI * L.n_cycles ? norm * powi(L.a0, 2) * powi(std::cos(phase / (2.0 * L.n_cycles)), 4) : 0.0;
    case Env::Flattop:
        if (std::fabs(phase) > PI * (L.n_cycles + 1.0)) return 0.0;
        if (std::fabs(phase) > PI * (L.n_cycles - 1.0)) {
            double arg = 0.25 * (std::fabs(phase) - (L.n_cycles - 1.0) * PI);
            return norm * powi(L.a0, 2) * powi(std::cos(arg), 4);
        }
        return norm * powi(L.a0, 2);
    default: {
        double arg = -powi(phase / (PI * L.n_cycles), 2);
        return norm * powi(L.a0, 2) * std::exp2(arg);
    }
    }
}

// plane_wave.rs:111-150 (time and z components of the four-gradient, both -grad)
static double pw_grad(const Laser& L, V4 r) {
    double norm = pw_norm(L);
    double phase = L.wavevector * r;
    switch (L.env) {
    case Env::CosSquared:
        return std::fabs(phase) < PI * L.n_cycles
                   ? norm * L.wavevector[0] * powi(L.a0, 2) * std::sin(phase / L.n_cycles) * powi(std::cos(phase / (2.0 * L.n_cycles)), 2) / L.n_cycles
                   : 0.0;
    case Env::Flattop: {
        if (std::fabs(phase) > PI * (L.n_cycles + 1.0) || std::fabs(phase) < PI * (L.n_cycles - 1.0)) return 0.0;
        double arg = 0.25 * (std::fabs(phase) - (L.n_cycles - 1.0) * PI);
        return norm * L.wavevector[0] * powi(L.a0, 2) * signum(phase) * std::sin(arg) * powi(std::cos(arg), 3);
    }
    default: {
        double arg = -powi(phase / (PI * L.n_cycles), 2);
        return norm * L.wavevector[0] * powi(L.a0, 2) * 2.0 * LN_2 * phase * std::exp2(arg) / powi(PI * L.n_cycles, 2);
    }
    }
}
static V4 pw_grad_a_sqd(const Laser& L, V4 r) { double g = pw_grad(L, r); return {-g, 0.0, 0.0, -g}; }

// ---------------------------------------------------------------- FocusedLaser (focused_laser.rs)
// n_cycles recomputed from the duration (focused_laser.rs:51-53)
static double fl_n_cycles(const Laser& L) { return SPEED_OF_LIGHT * L.duration * L.wavevector[0] / (2.0 * PI); }
static double fl_rayleigh(const Laser& L) { return 0.5 * L.wavevector[0] * powi(L.waist, 2); }
// focused_laser.rs:72-101
static void fl_envelope_and_grad(const Laser& L, double phase, double* env, double* grad) {
    double nc = fl_n_cycles(L);
    switch (L.env) {
    case Env::CosSquared:
        if (std::fabs(phase) < PI * nc) {
            *env = powi(std::cos(phase / (2.0 * nc)), 4);
            *grad = 2.0 * L.wavevector[0] * std::tan(phase / (2.0 * nc)) * *env / nc;
        } else {
            *env = 0.0; *grad = 0.0;
        }
        return;
    case Env::Flattop:
        if (std::fabs(phase) > PI * (nc + 1.0)) { *env = 0.0; *grad = 0.0; }
        else if (std::fabs(phase) > PI * (nc - 1.0)) {
            double arg = 0.25 * (std::fabs(phase) - (nc - 1.0) * PI);
            *env = powi(std::cos(arg), 4);
            *grad = L.wavevector[0] * signum(phase) * std::sin(arg) * powi(std::cos(arg), 3);
        } else { *env = 1.0; *grad = 0.0; }
        return;
    default: {
        double tau = SPEED_OF_LIGHT * L.wavevector[0] * L.duration;
        *env = std::exp(-4.0 * LN_2 * powi(phase, 2) / powi(tau, 2));
        *grad = 8.0 * LN_2 * L.wavevector[0] * phase * *env / powi(tau, 2);
    }
    }
}
// focused_laser.rs:103-117
double fl_a_sqd(const Laser& L, V4 r) {
    double z_r = fl_rayleigh(L);
    double width_sqd = 1.0 + powi(r[3] / z_r, 2);
    double rho_sqd = (powi(r[1], 2) + powi(r[2], 2)) / powi(L.waist, 2);
    double beam = pw_norm(L) * powi(L.a0, 2) * std::exp(-2.0 * rho_sqd / width_sqd) / width_sqd;
    double env, grad;
    fl_envelope_and_grad(L, L.wavevector * r, &env, &grad);
    return beam * env;
}
// focused_laser.rs:119-146
static V4 fl_grad_a_sqd(const Laser& L, V4 r) {
    double z_r = fl_rayleigh(L);
    double width_sqd = 1.0 + powi(r[3] / z_r, 2);
    double rho_sqd = (powi(r[1], 2) + powi(r[2], 2)) / powi(L.waist, 2);
    double beam = pw_norm(L) * powi(L.a0, 2) * std::exp(-2.0 * rho_sqd / width_sqd) / width_sqd;
    double gb0 = -4.0 * beam * r[1] / (powi(L.waist, 2) * width_sqd);
    double gb1 = -4.0 * beam * r[2] / (powi(L.waist, 2) * width_sqd);
    double gb2 = (2.0 * beam * r[3] / (powi(z_r, 2) * width_sqd)) * (2.0 * rho_sqd / width_sqd - 1.0);
    double env, ge;
    fl_envelope_and_grad(L, L.wavevector * r, &env, &ge);
    double tz = beam * ge + gb2 * env;
    return V4{-tz, -(gb0 * env), -(gb1 * env), -tz};
}
// the LMA fields' <a^2> and its four-gradient, by kind
static double lma_a_sqd(const Laser& L, V4 r) { return L.kind == Kind::FocusedLaser ? fl_a_sqd(L, r) : pw_a_sqd(L, r); }
static V4 lma_grad_a_sqd(const Laser& L, V4 r) { return L.kind == Kind::FocusedLaser ? fl_grad_a_sqd(L, r) : pw_grad_a_sqd(L, r); }
static V4 lma_ll_force(const Laser& L, V4 r, V4 u) {  // plane_wave.rs:152-160, focused_laser.rs:148-153
    double eta = SPEED_OF_LIGHT * COMPTON_TIME * (L.wavevector * u);
    return -2.0 * ALPHA_FINE * lma_a_sqd(L, r) * powi(eta, 2) * u / (3.0 * COMPTON_TIME);
}
static double lma_ll_work(const Laser& L, V4 r, V4 u) {  // plane_wave.rs:162-175, focused_laser.rs:155-163
    double eta = SPEED_OF_LIGHT * COMPTON_TIME * (L.wavevector * u);
    double omega = SPEED_OF_LIGHT * L.wavevector[0];
    double delta = L.pol == Pol::Circular ? 1.0 : 0.75;
    return 2.0 * ALPHA_FINE * omega * eta * delta * powi(lma_a_sqd(L, r), 2) / 3.0;
}

static double pw_bandwidth(const Laser& L) {  // plane_wave.rs:49-63, focused_laser.rs:45-58
    if (!L.finite_bandwidth) return 0.0;
    double nc = L.kind == Kind::FocusedLaser ? fl_n_cycles(L) : L.n_cycles;
    double n_fwhm = L.env == Env::CosSquared ? 0.36405666377387671305 * nc : nc;
    return std::sqrt(0.5 * LN_2) / (PI * n_fwhm);
}

double lma_max_timestep(const Laser& L) {  // plane_wave.rs:176-182, focused_laser.rs:167-173
    return L.env == Env::Flattop ? 0.2 / (SPEED_OF_LIGHT * L.wavevector[0]) : 1.0 / (SPEED_OF_LIGHT * L.wavevector[0]);
}
bool lma_contains(const Laser& L, V4 r) {  // plane_wave.rs:184-192, focused_laser.rs:175-183
    double nc = L.kind == Kind::FocusedLaser ? fl_n_cycles(L) : L.n_cycles;
    double phase = L.wavevector * r;
    double mp = L.env == Env::CosSquared ? PI * nc : (L.env == Env::Flattop ? PI * (nc + 1.0) : 6.0 * PI * nc);
    return phase < mp;
}
double lma_ideal_initial_z(const Laser& L) {  // plane_wave.rs:318-325, focused_laser.rs:274-281
    double nc = L.kind == Kind::FocusedLaser ? fl_n_cycles(L) : L.n_cycles;
    double wl = 2.0 * PI / L.wavevector[0];
    return L.env == Env::CosSquared ? 0.5 * wl * nc : (L.env == Env::Flattop ? 0.5 * wl * (nc + 1.0) : 2.0 * wl * nc);
}

// plane_wave.rs:196-249, focused_laser.rs:185-225
PushOut lma_push(const Laser& L, V4 r, V4 u, double rqm, double dt, Eqn eqn) {
    double dtau = dt / u[0];
    double scale = powi(rqm / (ELECTRON_CHARGE / ELECTRON_MASS), 2);
    V4 u_mid{0, 0, 0, 0};
    if (eqn == Eqn::LandauLifshitz) {
        V4 f = 0.5 * SPEED_OF_LIGHT * scale * lma_grad_a_sqd(L, r);
        V4 g = powi(scale, 2) * lma_ll_force(L, r, u);
        u_mid = u + 0.5 * (f + g) * dtau;
    } else if (eqn == Eqn::ModifiedLandauLifshitz) {
        std::fprintf(stderr, "oracle: Gaunt factor correction is unavailable in LMA mode\n");
        std::abort();
    }
    r = r + 0.5 * SPEED_OF_LIGHT * u * dtau;
    double dt_actual = 0.5 * u[0] * dtau;
    V4 f = 0.5 * SPEED_OF_LIGHT * scale * lma_grad_a_sqd(L, r);
    V4 g = eqn == Eqn::LandauLifshitz ? powi(scale, 2) * lma_ll_force(L, r, u_mid) : V4{0, 0, 0, 0};
    u = u + (f + g) * dtau;
    double dwork = eqn == Eqn::LandauLifshitz ? (f[0] + lma_ll_work(L, r, u_mid)) * dtau : f[0] * dtau;
    r = r + 0.5 * SPEED_OF_LIGHT * u * dtau;
    dt_actual = dt_actual + 0.5 * u[0] * dtau;
    u = with_sqr(u, 1.0 + lma_a_sqd(L, r));
    return {r, u, dt_actual, dwork};
}

// plane_wave.rs:251-282
bool lma_radiate(const Laser& L, V4 r, V4 u, double dt, Xoshiro& rng, bool classical, RadEvent* ev) {
    // plane_wave.rs:251-282; focused_laser.rs:227-247 (no chirp)
    double a = std::sqrt(lma_a_sqd(L, r));
    double phase = L.wavevector * r;
    double chirp = L.kind == Kind::FocusedLaser ? 1.0 : 1.0 + 2.0 * L.chirp_b * phase;
    double width = 1.0 + pw_bandwidth(L) * rng.normal();
    V4 kappa = L.kind == Kind::FocusedLaser ? SPEED_OF_LIGHT * COMPTON_TIME * L.wavevector * width
                                            : SPEED_OF_LIGHT * COMPTON_TIME * L.wavevector * chirp * width;
    double prob = 0.0;
    if (!nlc_probability(kappa, u, dt, L.pol, classical, &prob)) prob = 0.0;
    if (!(rng.gen_f64() < prob)) return false;
    int n;
    V4 k;
    SV pol;
    nlc_generate(kappa, u, L.pol, L.pol_angle, classical, rng, &n, &k, &pol, nullptr);
    *ev = RadEvent{k, u + double(n) * kappa - k, pol, a, a * (u * kappa), double(n) * kappa[0]};
    return true;
}

// ================================================================ LP nonlinear Breit-Wheeler (pair_creation/lp)
static const char* BT = "pair_creation/lp/tables/total:";
static const char* BC = "pair_creation/lp/tables/cdf:";
static const char* MT = "pair_creation/lp/tables/mid_range/total:";
static const char* MC = "pair_creation/lp/tables/mid_range/cdf:TABLE";

// pair_creation/lp/theta_bound.rs:10-100
struct BwThetaBound {
    double s[16], f[16];
    static BwThetaBound for_harmonic(int n, double a, double eta) {
        BwThetaBound t;
        double s_min = 0.5;
        double s_max = 0.5 + std::sqrt(0.25 - (1.0 + 0.5 * a * a) / (2.0 * double(n) * eta));
        for (int i = 0; i < 16; ++i) {
            double z = std::cos(PI * double(i) / 15.0);
            t.s[i] = s_min + 0.5 * (z + 1.0) * (s_max - s_min);
            double r_n = std::sqrt(2.0 * double(n) * eta * t.s[i] * (1.0 - t.s[i]) - (1.0 + 0.5 * a * a));
            double x = a * r_n / (eta * t.s[i] * (1.0 - t.s[i]));
            double y = a * a / (8.0 * eta * t.s[i] * (1.0 - t.s[i]));
            double x_crit = y > double(n) / 6.0 ? 4.0 * std::sqrt(y) * std::sqrt(double(n) - 2.0 * y) : double(n) + 2.0 * y;
            double ln_j_crit = std::log(0.443) - std::log(double(n)) / 3.0;
            double ln_j_bdy = n % 2 == 0 ? ThetaBound::ln_double_bessel_x_zero(n, y) : ThetaBound::ln_double_bessel_x_zero(n + 1, y);
            ln_j_bdy = std::fmin(ln_j_bdy, ln_j_crit);
            double cos_theta = 1.0 - x_crit * std::log(0.01) / (x * (ln_j_bdy - ln_j_crit));
            cos_theta = std::fmax(cos_theta, -1.0);
            t.f[i] = std::acos(cos_theta);
            if (std::isnan(t.f[i])) t.f[i] = PI;
        }
        double f_min = t.f[0];
        for (int i = 1; i < 16; ++i) {
            if (t.f[i] < f_min) f_min = t.f[i];
            else t.f[i] = f_min;
        }
        return t;
    }
    double at(double sv) const {
        double val = f[15];
        for (int i = 1; i < 16; ++i) {
            if (sv > s[i]) {
                double w = (sv - s[i - 1]) / (s[i] - s[i - 1]);
                val = w * f[i] + (1.0 - w) * f[i - 1];
                break;
            }
        }
        return std::fmin(val, FRAC_PI_2);
    }
};

// quadrature/adaptive.rs: 2-D adaptive Clenshaw-Curtis integration of a complex
// function over a rectangle, splitting the region with the largest error.
struct CcRegion { C integral, error; double x[2], y[2]; C cache[3][3]; };
static const double CC5[5][3] = {{0.0, 1.0 / 30.0, 1.0 / 6.0}, {0.1464466094067262, 4.0 / 15.0, 0.0}, {0.5, 2.0 / 5.0, 2.0 / 3.0},
                                 {0.8535533905932738, 4.0 / 15.0, 0.0}, {1.0, 1.0 / 30.0, 1.0 / 6.0}};
static const double CC5X[9] = {0.0, 0.1464466094067262, 0.5, 0.8535533905932738, 1.0, 1.1464466094067262, 1.5, 1.8535533905932738, 2.0};
static const double CC9[9][3] = {{0.0000000000000000, 0.0079365079365079, 1.0 / 30.0}, {0.0380602337443566, 0.0731093246080091, 0.0},
                                 {0.1464466094067262, 0.1396825396825397, 4.0 / 15.0}, {0.3086582838174551, 0.1808589293602449, 0.0},
                                 {0.5000000000000000, 0.1968253968253968, 2.0 / 5.0}, {0.6913417161825449, 0.1808589293602449, 0.0},
                                 {0.8535533905932738, 0.1396825396825397, 4.0 / 15.0}, {0.9619397662556434, 0.0731093246080091, 0.0},
                                 {1.0000000000000000, 0.0079365079365079, 1.0 / 30.0}};
static const double CC9X[17] = {0.0000000000000000, 0.0380602337443566, 0.1464466094067262, 0.3086582838174551, 0.5000000000000000,
                                0.6913417161825449, 0.8535533905932738, 0.9619397662556434, 1.0000000000000000, 1.0380602337443566,
                                1.1464466094067262, 1.3086582838174551, 1.5000000000000000, 1.6913417161825449, 1.8535533905932738,
                                1.9619397662556434, 2.0000000000000000};

template <class F, class R>
static CcRegion cc_new(F& f, R& region, double x0, double x1, double y0, double y1) {  // adaptive.rs:276-313
    CcRegion g{};
    C result{0.0, 0.0}, error{0.0, 0.0};
    for (int i = 0; i < 5; ++i) {
        double y = y0 + CC5[i][0] * (y1 - y0);
        for (int j = 0; j < 5; ++j) {
            double x = x0 + CC5[j][0] * (x1 - x0);
            C z = region(x, y) ? f(x, y) : C{0.0, 0.0};
            result = result + (CC5[i][1] * CC5[j][1] * (y1 - y0) * (x1 - x0)) * z;
            error = error + (CC5[i][2] * CC5[j][2] * (y1 - y0) * (x1 - x0)) * z;
            if (i % 2 == 0 && j % 2 == 0) g.cache[i / 2][j / 2] = z;
        }
    }
    g.integral = result;
    g.error = result - error;
    g.x[0] = x0; g.x[1] = x1; g.y[0] = y0; g.y[1] = y1;
    return g;
}

template <class F, class R>
static void cc_custom_partition(const CcRegion& g, F& f, R& region, const double* xs, const double* ys, CcRegion out[4]) {
    // adaptive.rs:32-179
    double x0 = g.x[0], x1 = g.x[1], y0 = g.y[0], y1 = g.y[1];
    double x_mid = xs ? *xs : 0.5 * (x0 + x1);
    double y_mid = ys ? *ys : 0.5 * (y0 + y1);
    C z[9][9];
    for (int i = 0; i < 9; ++i) {
        double ty = CC5X[i];
        double y = i < 4 ? y0 + ty * (y_mid - y0) : y_mid + (ty - 1.0) * (y1 - y_mid);
        for (int j = 0; j < 9; ++j) {
            double tx = CC5X[j];
            double x = j < 4 ? x0 + tx * (x_mid - x0) : x_mid + (tx - 1.0) * (x1 - x_mid);
            if (i % 8 == 0 && j % 8 == 0) z[i][j] = g.cache[i / 4][j / 4];
            else if (i % 4 == 0 && j % 4 == 0) {
                if (!xs && !ys) z[i][j] = g.cache[i / 4][j / 4];
                else z[i][j] = region(x, y) ? f(x, y) : C{0.0, 0.0};
            } else z[i][j] = region(x, y) ? f(x, y) : C{0.0, 0.0};
        }
    }
    C result[4] = {}, error[4] = {};
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 2; ++c)
            for (int i = 0; i < 5; ++i) {
                double wy = CC5[i][1], ey = CC5[i][2];
                for (int j = 0; j < 5; ++j) {
                    double wx = CC5[j][1], ex = CC5[j][2];
                    C val = z[i + 4 * r][j + 4 * c];
                    double dy = r < 1 ? y_mid - y0 : y1 - y_mid;
                    double dx = c < 1 ? x_mid - x0 : x1 - x_mid;
                    result[2 * r + c] = result[2 * r + c] + (wx * wy * dx * dy) * val;
                    error[2 * r + c] = error[2 * r + c] + (ex * ey * dx * dy) * val;
                }
            }
    const double xr[2][2] = {{x0, x_mid}, {x_mid, x1}}, yr[2][2] = {{y0, y_mid}, {y_mid, y1}};
    for (int q = 0; q < 4; ++q) {
        int r = q / 2, c = q % 2;
        CcRegion& o = out[q];
        o.integral = result[q];
        o.error = result[q] - error[q];
        o.x[0] = xr[c][0]; o.x[1] = xr[c][1]; o.y[0] = yr[r][0]; o.y[1] = yr[r][1];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) o.cache[a][b] = z[4 * r + 2 * a][4 * c + 2 * b];
    }
}

template <class F, class R>
static void cc_partition(const CcRegion& g, F& f, R& region, CcRegion out[4]) {  // adaptive.rs:181-274
    double x0 = g.x[0], x1 = g.x[1], y0 = g.y[0], y1 = g.y[1];
    double x_mid = 0.5 * (x0 + x1), y_mid = 0.5 * (y0 + y1);
    static thread_local C z[17][17];
    for (int i = 0; i < 17; ++i) {
        double y = y0 + 0.5 * CC9X[i] * (y1 - y0);
        for (int j = 0; j < 17; ++j) {
            double x = x0 + 0.5 * CC9X[j] * (x1 - x0);
            if (i % 8 == 0 && j % 8 == 0) z[i][j] = g.cache[i / 8][j / 8];
            else z[i][j] = region(x, y) ? f(x, y) : C{0.0, 0.0};
        }
    }
    C result[4] = {}, error[4] = {};
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 2; ++c)
            for (int i = 0; i < 9; ++i) {
                double wy = CC9[i][1], ey = CC9[i][2];
                for (int j = 0; j < 9; ++j) {
                    double wx = CC9[j][1], ex = CC9[j][2];
                    C val = z[i + 8 * r][j + 8 * c];
                    double dy = 0.5 * (y1 - y0);
                    double dx = 0.5 * (x1 - x0);
                    result[2 * r + c] = result[2 * r + c] + (wx * wy * dx * dy) * val;
                    error[2 * r + c] = error[2 * r + c] + (ex * ey * dx * dy) * val;
                }
            }
    const double xr[2][2] = {{x0, x_mid}, {x_mid, x1}}, yr[2][2] = {{y0, y_mid}, {y_mid, y1}};
    for (int q = 0; q < 4; ++q) {
        int r = q / 2, c = q % 2;
        CcRegion& o = out[q];
        o.integral = result[q];
        o.error = result[q] - error[q];
        o.x[0] = xr[c][0]; o.x[1] = xr[c][1]; o.y[0] = yr[r][0]; o.y[1] = yr[r][1];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) o.cache[a][b] = z[8 * r + 4 * a][8 * c + 4 * b];
    }
}

template <class F, class R>
static C cc_integrate_2d(F f, double x0, const double* x_div, double x1, double y0, const double* y_div, double y1, R region,
                         double tol, int max_recursion) {  // adaptive.rs:316-380
    std::vector<CcRegion> regions;
    regions.reserve(16 + 4 * size_t(max_recursion));
    CcRegion first = cc_new(f, region, x0, x1, y0, y1);
    if (!x_div && !y_div) regions.push_back(first);
    else {
        CcRegion q[4];
        cc_custom_partition(first, f, region, x_div, y_div, q);
        for (auto& r : q) regions.push_back(r);
    }
    C integral{0.0, 0.0};
    for (int it = 0; it < max_recursion; ++it) {
        C err{0.0, 0.0};
        integral = C{0.0, 0.0};
        for (auto& r : regions) { integral = integral + r.integral; err = err + r.error; }
        if (std::fabs(err.re) <= tol * std::fabs(integral.re) && std::fabs(err.im) <= tol * std::fabs(integral.im)) break;
        CcRegion g = regions.back();
        regions.pop_back();
        CcRegion q[4];
        cc_partition(g, f, region, q);
        for (auto& r : q) regions.push_back(r);
        std::stable_sort(regions.begin(), regions.end(), [](const CcRegion& a, const CcRegion& b) {
            return (a.error.re * a.error.re + a.error.im * a.error.im) < (b.error.re * b.error.re + b.error.im * b.error.im);
        });
    }
    return integral;
}

// pair_creation/lp/mod.rs:31-302
struct BwDdpr {
    int n;
    double a, eta;
    DoubleBessel dj;
    BwThetaBound tb;
    BwDdpr(int n_, double a_, double eta_)
        : n(n_), a(a_), eta(eta_), dj(n_, double(n_) * SQRT_2, double(n_) * 0.5), tb(BwThetaBound::for_harmonic(n_, a_, eta_)) {}
    void s_bounds(double* s_min, double* s_max) const {  // :50-55
        *s_min = 0.5;
        *s_max = std::fmax(0.5 + std::sqrt(0.25 - (1.0 + 0.5 * a * a) / (2.0 * double(n) * eta)), 0.5);
    }
    double max_theta(double s) const { return tb.at(s); }
    void elements(double s, double theta, double* unpol, double* d1, double* d2) {  // :63-95
        double r_n_sqd = 2.0 * double(n) * eta * s * (1.0 - s) - (1.0 + 0.5 * a * a);
        if (!(r_n_sqd > 0.0)) { *unpol = *d1 = *d2 = 0.0; return; }
        double x = a * std::sqrt(r_n_sqd) * std::cos(theta) / (eta * s * (1.0 - s));
        double y = a * a / (8.0 * eta * s * (1.0 - s));
        double j[5];
        dj.evaluate(x, y, j);
        double g0 = j[2], g1 = 0.5 * (j[1] + j[3]), g2 = 0.25 * (j[0] + 2.0 * j[2] + j[4]);
        double h_s = 0.5 / (s * (1.0 - s)) - 1.0;
        *unpol = powi(g0, 2) + a * a * h_s * (powi(g1, 2) - g0 * g2);
        *d1 = powi(g0 * std::sqrt(r_n_sqd) * std::cos(theta) - g1 * a, 2) - powi(g0 * std::sin(theta), 2) * r_n_sqd;
        *d2 = 2.0 * g0 * std::sqrt(r_n_sqd) * std::sin(theta) * (g0 * std::sqrt(r_n_sqd) * std::cos(theta) - g1 * a);
    }
    C at(double s, double theta) {  // :97-106
        double u, d1, d2;
        elements(s, theta, &u, &d1, &d2);
        return {(u - d1) / (2.0 * PI), (u + d1) / (2.0 * PI)};
    }
    double fully_resolved(double s, double theta, double sv1, double sv2) {  // :109-112
        double u, d1, d2;
        elements(s, theta, &u, &d1, &d2);
        return (u - sv1 * d1 - sv2 * d2) / (2.0 * PI);
    }
    C integrate_direct() {  // :125-185
        Table N16 = table("quadrature/tables:GAUSS_16_NODES"), W16 = table("quadrature/tables:GAUSS_16_WEIGHTS");
        Table N32 = table("quadrature/tables:GAUSS_32_NODES"), W32 = table("quadrature/tables:GAUSS_32_WEIGHTS");
        double s_min, s_max;
        s_bounds(&s_min, &s_max);
        double s_peak = std::fmin(0.5 + std::sqrt(0.25 - (1.0 + a * a) / (2.0 * double(n) * eta)), 0.5);
        auto seg = [&](double s0, double s1, const Table& Nn, const Table& Ww, int m) {
            C acc{0.0, 0.0};
            for (int i = 0; i < m; ++i) {
                double sv = s0 + 0.5 * (Nn[i] + 1.0) * (s1 - s0);
                C sd = theta_integral(sv, max_theta(sv));
                acc = acc + (0.5 * Ww[i] * (s1 - s0)) * sd;
            }
            return acc;
        };
        C integral;
        if (s_peak < s_min + 2.0 * (s_max - s_min) / 3.0) {
            s_peak = std::fmin(s_peak, s_min + (s_max - s_min) / 3.0);
            integral = seg(s_min, s_peak, N32, W32, 32) + seg(s_peak, s_max, N32, W32, 32);
        } else {
            double delta = s_max - s_peak;
            integral = seg(s_min, s_max - 2.0 * delta, N16, W16, 16) + seg(s_max - 2.0 * delta, s_peak, N32, W32, 32) +
                       seg(s_peak, s_max, N32, W32, 32);
        }
        return (2.0 * 4.0) * integral;
    }
    C theta_integral(double sv, double max_th) {  // integrate_over_theta, :114-123
        Table N = table("quadrature/tables:GAUSS_32_NODES"), W = table("quadrature/tables:GAUSS_32_WEIGHTS");
        C sum{0.0, 0.0};
        for (int i = 0; i < 32; ++i) {
            double th = 0.5 * (N[i] + 1.0) * max_th;
            sum = sum + (0.5 * W[i] * max_th) * at(sv, th);
        }
        return sum;
    }
    C integrate_adaptive() {  // :187-217
        double s_min, s_max;
        s_bounds(&s_min, &s_max);
        double s_peak = 0.5 + std::sqrt(0.25 - (1.0 + a * a) / (2.0 * double(n) * eta));
        bool have_peak = s_peak > s_min + 2.0 * (s_max - s_min) / 3.0;
        double theta_split = tb.at(have_peak ? s_peak : 0.5);
        auto f = [&](double sv, double th) { return at(sv, th); };
        auto region = [&](double sv, double th) { return th < tb.at(sv); };
        C integral = cc_integrate_2d(f, s_min, have_peak ? &s_peak : nullptr, s_max, 0.0, &theta_split, FRAC_PI_2, region, 1.0e-3, 16);
        return (4.0 * 2.0) * integral;
    }
    C integrate() {  // :219-228
        if (max_theta(0.5) < 0.75 * FRAC_PI_2) return integrate_adaptive();
        return integrate_direct();
    }
    double ceiling(double sv1, double sv2) {  // :230-258
        double s_min, s_max;
        s_bounds(&s_min, &s_max);
        double centre = 0.5 + std::sqrt(0.25 - (1.0 + a * a) / (2.0 * double(n) * eta));
        double lower = s_min, upper = s_max;
        if (std::isfinite(centre)) {
            lower = std::fmax(centre - 0.2, s_min);
            upper = 0.5 * (centre + s_max);
        }
        double mx = -INFINITY;
        for (int i = 0; i < 50; ++i) mx = std::fmax(mx, fully_resolved(lower + (upper - lower) * double(i) / 50.0, 0.0, sv1, sv2));
        return 1.2 * mx;
    }
    void sample(double sv1, double sv2, Xoshiro& rng, double* s_out, double* th_out, uint64_t* trials) {  // :261-301
        double s_min, s_max;
        s_bounds(&s_min, &s_max);
        double mx = ceiling(sv1, sv2);
        double s, theta;
        for (;;) {
            s = s_min + (s_max - s_min) * rng.gen_f64();
            theta = FRAC_PI_2 * rng.gen_f64();
            if (theta > max_theta(s)) continue;
            double z = mx * rng.gen_f64();
            double f = fully_resolved(s, theta, sv1, sv2);
            if (trials) ++*trials;
            if (z < f) break;
        }
        // rand 0.7 gen_range(0, 2) and gen_range(0, 4): widening multiply of one 32-bit draw
        uint32_t r1 = uint32_t(rng.next_u64() >> 32);
        if (int((uint64_t(r1) * 2u) >> 32) == 0) s = 1.0 - s;
        uint32_t r2 = uint32_t(rng.next_u64() >> 32);
        switch (int((uint64_t(r2) * 4u) >> 32)) {
        case 1: theta = PI - theta; break;
        case 2: theta = PI + theta; break;
        case 3: theta = 2.0 * PI - theta; break;
        default: break;
        }
        *s_out = s;
        *th_out = theta;
    }
};

// pair_creation/lp/mod.rs:345-362
static void bw_sum_limits(double a, double eta, int* n_lo, int* n_hi) {
    double n_min = std::ceil(2.0 * (1.0 + 0.5 * a * a) / eta);
    double range = a < 1.0 ? 2.0 + (2.0 + 20.0 * a * a) * std::exp(-std::sqrt(0.5 * eta))
                           : 3.0 + 2.8 * std::pow(a, 8.0 / 3.0) / std::cbrt(eta) + 0.25 * powi(a, 3) * std::sqrt(eta);
    double test = 0.25 - (1.0 + 0.5 * a * a) / (2.0 * n_min * eta);
    if (test <= 2.220446049250313e-16) { *n_lo = int(n_min) + 1; *n_hi = int(n_min + 1.0 + range); }
    else { *n_lo = int(n_min); *n_hi = int(n_min + range); }
}
static bool bw_is_negligible(double a, double eta) {  // :371-374
    return std::log10(eta) < -1.0 - powi(std::log10(a) + 2.0, 2) / 4.5;
}

// tables/mod.rs:18-60 (high range)
static bool bw_hi_contains(double a, double eta) {
    return std::log(a) >= tb(BT, "LN_MIN_A")[0] && a < 25.0 && std::log(eta) >= tb(BT, "LN_MIN_ETA")[0] && eta < 2.0;
}
static void bw_hi_interpolate(double a, double eta, double out[2]) {
    const int N_COLS = 25, N_ROWS = 61;
    double LMA_ = tb(BT, "LN_MIN_A")[0], LME = tb(BT, "LN_MIN_ETA")[0], LAS = tb(BT, "LN_A_STEP")[0], LES = tb(BT, "LN_ETA_STEP")[0];
    Table T = tb(BT, "TABLE");
    size_t ia = size_t((std::log(a) - LMA_) / LAS), ie = size_t((std::log(eta) - LME) / LES);
    if (!(ia < size_t(N_COLS - 1) && ie < size_t(N_ROWS - 1))) { std::fprintf(stderr, "oracle: NBW table index\n"); std::abort(); }
    auto at = [&](size_t r, size_t c, int k) { return T[(r * N_COLS + c) * 2 + k]; };
    for (int k = 0; k < 2; ++k) {
        double f;
        if (a * eta > 2.0) {
            double da = (std::log(a) - LMA_) / LAS - double(ia), de = (std::log(eta) - LME) / LES - double(ie);
            f = (1.0 - da) * (1.0 - de) * at(ie, ia, k) + da * (1.0 - de) * at(ie, ia + 1, k) + (1.0 - da) * de * at(ie + 1, ia, k) +
                da * de * at(ie + 1, ia + 1, k);
        } else {
            double a_min = std::exp(LMA_ + double(ia) * LAS), a_max = std::exp(LMA_ + double(ia + 1) * LAS);
            double e_min = std::exp(LME + double(ie) * LES), e_max = std::exp(LME + double(ie + 1) * LES);
            double da = ((a_min * a_max / a) - a_min) / (a_max - a_min);
            double de = ((e_min * e_max / eta) - e_min) / (e_max - e_min);
            f = at(ie, ia, k) * da * de + at(ie, ia + 1, k) * (1.0 - da) * de + at(ie + 1, ia, k) * da * (1.0 - de) +
                at(ie + 1, ia + 1, k) * (1.0 - da) * (1.0 - de);
        }
        out[k] = std::exp(f);
    }
}
static double poldep_interp(double par, double perp, double sv1) { return 0.5 * ((par + perp) + sv1 * (par - perp)); }

// tables/mod.rs:62-120
static int bw_hi_invert(double a, double eta, double sv1, double frac) {
    const int N_COLS = 25;
    Table MIN = tb(BC, "MIN"), STEP = tb(BC, "STEP"), T = tb(BC, "TABLE");
    double iaf = (std::log(a) - MIN[0]) / STEP[0], da = fract(iaf);
    size_t ia = size_t(iaf);
    double ief = (std::log(eta) - MIN[1]) / STEP[1], de = fract(ief);
    size_t ie = size_t(ief);
    size_t idx[4] = {N_COLS * ie + ia, N_COLS * ie + ia + 1, N_COLS * (ie + 1) + ia, N_COLS * (ie + 1) + ia + 1};
    double w[4] = {(1.0 - da) * (1.0 - de), da * (1.0 - de), (1.0 - da) * de, da * de};
    double u = 0.0;
    for (int k = 0; k < 4; ++k) {
        const double* src = T.p + idx[k] * 48;
        double n_min = src[0];
        double f_max = poldep_interp(src[15 * 3 + 1], src[15 * 3 + 2], sv1);
        double tbl[30];
        for (int i = 0; i < 15; ++i) {
            tbl[2 * i] = std::log(src[3 * i] - n_min + 1.0);
            double f = poldep_interp(src[3 * i + 1], src[3 * i + 2], sv1);
            tbl[2 * i + 1] = std::log(-1.0 * std::log(1.0 - f / f_max));
        }
        double fr = std::log(-1.0 * std::log(1.0 - frac));
        double uk;
        if (!Interp{tbl, 15, true}.invert(fr, &uk)) { std::fprintf(stderr, "oracle: NBW invert failed\n"); std::abort(); }
        u += uk * w[k];
    }
    int n_lo, n_hi;
    bw_sum_limits(a, eta, &n_lo, &n_hi);
    double n = double(n_lo) + std::expm1(u);
    return int(n) + 1;
}

// tables/mid_range/mod.rs:9-16
static bool bw_mid_contains(double a, double eta) {
    double ln_a = std::log(a), ln_ep = std::log(eta / (1.0 + 0.5 * a * a));
    return ln_a > tb(MT, "LN_MIN_A")[0] && a < 2.0 && ln_ep > tb(MT, "LN_MIN_ETA_PRIME")[0] && ln_ep < tb(MT, "LN_MAX_ETA_PRIME")[0];
}
// mid_range/mod.rs:18-82: three-point power-law fit in eta' along one a column
static void bw_mid_interp_eta_prime(double ep, size_t ia, double out[2]) {
    const int N_COLS = 36;
    double D = tb(MT, "ETA_PRIME_DENSITY")[0], LMX = tb(MT, "LN_MAX_ETA_PRIME")[0];
    Table T = tb(MT, "TABLE");
    size_t ie = size_t(D * (std::exp(LMX) / ep - 1.0));
    if (ie % size_t(D) == size_t(D) - 1) ie = ie - 1;
    auto row = [&](size_t r, int k) { return T[(r * N_COLS + ia) * 2 + k]; };
    double lo0 = 2.0 / (1.0 + double(ie + 2) / D), mid0 = 2.0 / (1.0 + double(ie + 1) / D), up0 = 2.0 / (1.0 + double(ie) / D);
    for (int k = 0; k < 2; ++k) {
        double lo1 = row(ie + 2, k), mid1 = row(ie + 1, k), up1 = row(ie, k);
        double m = std::log((up1 - lo1) / (mid1 - lo1)) / std::log((up0 - lo0) / (mid0 - lo0));
        double alpha = (mid1 - lo1) / std::pow(mid0 - lo0, m);
        double v = alpha * std::pow(ep - lo0, m) + lo1;
        if (!std::isfinite(v)) {
            if (ep < mid0) {
                double w = (std::log(ep) - std::log(lo0)) / (std::log(mid0) - std::log(lo0));
                v = (1.0 - w) * lo1 + w * mid1;
            } else {
                double w = (std::log(ep) - std::log(mid0)) / (std::log(up0) - std::log(mid0));
                v = (1.0 - w) * mid1 + w * up1;
            }
        }
        out[k] = v;
    }
}
static void bw_mid_interpolate(double a, double eta, double out[2]) {  // mid_range/mod.rs:84-100
    double LMA_ = tb(MT, "LN_MIN_A")[0], LAS = tb(MT, "LN_A_STEP")[0];
    size_t ia = size_t((std::log(a) - LMA_) / LAS);
    double ep = eta / (1.0 + 0.5 * a * a);
    double f0[2], f1[2];
    bw_mid_interp_eta_prime(ep, ia, f0);
    bw_mid_interp_eta_prime(ep, ia + 1, f1);
    for (int k = 0; k < 2; ++k) {
        double f;
        if (a < 0.5) {
            double da = (std::log(a) - LMA_) / LAS - double(ia);
            f = (1.0 - da) * f0[k] + da * f1[k];
        } else {
            double a_min = std::exp(LMA_ + double(ia) * LAS), a_max = std::exp(LMA_ + double(ia + 1) * LAS);
            double da = (a - a_min) / (a_max - a_min);
            f = ((1.0 - da) * a_min * f0[k] + da * a_max * f1[k]) / a;
        }
        out[k] = std::exp(f);
    }
}
static double bw_mid_project(double a, double eta, double sv1) {
    double f[2];
    bw_mid_interpolate(a, eta, f);
    return poldep_interp(f[0], f[1], sv1);
}
// mid_range/mod.rs:109-162
static void bw_mid_cdf_setup(double a, double eta, size_t* idx, double* w, bool* next_to_bdy) {
    const int N_COLS = 36;
    double LMA_ = tb(MT, "LN_MIN_A")[0], LAS = tb(MT, "LN_A_STEP")[0], LMX = tb(MT, "LN_MAX_ETA_PRIME")[0];
    double D = tb(MT, "ETA_PRIME_DENSITY")[0];
    size_t ia = size_t((std::log(a) - LMA_) / LAS);
    double da = (std::log(a) - LMA_) / LAS - double(ia);
    double ep = eta / (1.0 + 0.5 * a * a);
    size_t ie = size_t(D * (std::exp(LMX) / ep - 1.0));
    if (next_to_bdy) {
        *next_to_bdy = ie % size_t(D) == size_t(D) - 1;
        if (*next_to_bdy) ie = ie - 1;
    }
    double lower = std::exp(LMX) / (1.0 + double(ie + 1) / D), upper = std::exp(LMX) / (1.0 + double(ie) / D);
    double de = (std::log(ep) - std::log(lower)) / (std::log(upper) - std::log(lower));
    idx[0] = N_COLS * ie + ia; idx[1] = N_COLS * ie + ia + 1; idx[2] = N_COLS * (ie + 1) + ia; idx[3] = N_COLS * (ie + 1) + ia + 1;
    w[0] = (1.0 - da) * de; w[1] = da * de; w[2] = (1.0 - da) * (1.0 - de); w[3] = da * (1.0 - de);
}
static void bw_mid_lowest(double a, double eta, double sv1, int* n_min_out, double* rate_out) {
    size_t idx[4];
    double w[4];
    bw_mid_cdf_setup(a, eta, idx, w, nullptr);
    Table T = table(MC);
    int n_lo, n_hi;
    bw_sum_limits(a, eta, &n_lo, &n_hi);
    double total = bw_mid_project(a, eta, sv1);
    double acc = 0.0;
    for (int k = 0; k < 4; ++k) {
        const double* t = T.p + idx[k] * 48;
        double frac = 0.0;
        if (!(t[0] > double(n_lo))) frac = poldep_interp(t[1], t[2], sv1) / poldep_interp(t[45 + 1], t[45 + 2], sv1);
        acc += w[k] * powi(frac * total, 3);
    }
    *n_min_out = n_lo;
    *rate_out = std::cbrt(acc);
}
static int bw_mid_invert(double a, double eta, double sv1, double frac) {  // mid_range/mod.rs:164-272
    size_t idx[4];
    double w[4];
    bool bdy = false;
    bw_mid_cdf_setup(a, eta, idx, w, &bdy);
    Table T = table(MC);
    double n = 0.0;
    if (bdy) {
        int n_min;
        double rate_lowest;
        bw_mid_lowest(a, eta, sv1, &n_min, &rate_lowest);
        double p_lowest = rate_lowest / bw_mid_project(a, eta, sv1);
        if (frac <= p_lowest) return n_min;
        for (int k = 0; k < 4; ++k) {
            const double* src = T.p + idx[k] * 48;
            double f_max = poldep_interp(src[46], src[47], sv1);
            double tbl[32];
            for (int i = 0; i < 16; ++i) { tbl[2 * i] = src[3 * i]; tbl[2 * i + 1] = poldep_interp(src[3 * i + 1], src[3 * i + 2], sv1) / f_max; }
            if (tbl[0] == double(n_min)) {
                double t01 = tbl[1];
                for (int i = 0; i < 16; ++i) tbl[2 * i + 1] = (t01 - tbl[2 * i + 1] + p_lowest * (tbl[2 * i + 1] - 1.0)) / (t01 - 1.0);
            }
            double nk;
            if (!Interp{tbl, 16}.invert(frac, &nk)) { std::fprintf(stderr, "oracle: NBW mid invert failed\n"); std::abort(); }
            n += nk * w[k];
        }
    } else {
        for (int k = 0; k < 4; ++k) {
            const double* src = T.p + idx[k] * 48;
            double f_max = poldep_interp(src[46], src[47], sv1);
            double tbl[32];
            for (int i = 0; i < 16; ++i) { tbl[2 * i] = src[3 * i]; tbl[2 * i + 1] = poldep_interp(src[3 * i + 1], src[3 * i + 2], sv1) / f_max; }
            double nk;
            if (frac <= tbl[1]) nk = tbl[0] - 0.1;
            else if (!Interp{tbl, 16}.invert(frac, &nk)) { std::fprintf(stderr, "oracle: NBW mid invert failed\n"); std::abort(); }
            n += nk * w[k];
        }
    }
    return int(std::ceil(n));
}

// pair_creation/lp/mod.rs:445-512 (TotalRate::by_summation): partial rates of every
// harmonic in sum_limits, cumulated, sampled at 16 log-spaced harmonics; out48 is
// the [n, par, perp] x 16 row of a CDF table. Used to check the regenerated table.
static void bw_by_summation(double a, double eta, double out48[48]) {
    int n_min, n_max;
    bw_sum_limits(a, eta, &n_min, &n_max);
    std::vector<C> cum;
    C total{0.0, 0.0};
    for (int n = n_min; n <= n_max; ++n) {
        BwDdpr d(n, a, eta);
        total = total + d.integrate();
        cum.push_back(total);
    }
    double delta = double(n_max - n_min);
    double u_max = std::log(1.0 + delta);
    double nn[16];
    nn[0] = double(n_min);
    double u = 0.0, du = u_max / 15.0;
    for (int i = 1; i <= 15; ++i) {
        u = u + du;
        double n = double(n_min) + std::expm1(u);
        if (std::round(n) <= nn[i - 1]) {
            n = nn[i - 1] + 1.0;
            u = std::log(n - double(n_min) + 1.0);
            du = (u_max - u) / (15.0 - double(i));
            nn[i] = n;
        } else nn[i] = std::round(n);
    }
    for (int i = 0; i < 16; ++i) {
        size_t index = size_t(nn[i]) - size_t(n_min);
        C c = index < cum.size() ? cum[index] : cum.back();
        out48[3 * i] = nn[i];
        out48[3 * i + 1] = c.re;
        out48[3 * i + 2] = c.im;
    }
}

// pair_creation/lp/mod.rs:385-422
static double bw_probability(double a, double eta, SV sv, double dphi, SV* sv_out) {
    double f0 = 0.0, f1 = 0.0;
    if (bw_is_negligible(a, eta)) {}
    else if (bw_mid_contains(a, eta)) { double f[2]; bw_mid_interpolate(a, eta, f); f0 = f[0]; f1 = f[1]; }
    else if (bw_hi_contains(a, eta)) { double f[2]; bw_hi_interpolate(a, eta, f); f0 = f[0]; f1 = f[1]; }
    double prob = ALPHA_FINE * (0.5 * ((f0 + f1) + sv[1] * (f0 - f1))) * dphi / eta;
    double prob_avg = ALPHA_FINE * (0.5 * (f0 + f1)) * dphi / eta;
    double delta = ALPHA_FINE * (0.5 * (f0 - f1)) * dphi / eta;
    double sv1 = (sv[1] * (1.0 - prob_avg) - delta) / (1.0 - prob);
    double sv2 = sv[2] * (1.0 - prob_avg) / (1.0 - prob);
    double sv3 = sv[3] * (1.0 - prob_avg) / (1.0 - prob);
    *sv_out = SV{sv[0], sv1, sv2, sv3};
    return prob;
}
// pair_creation/lp/mod.rs:424-442
static void bw_sample(double a, double eta, double sv1, double sv2, Xoshiro& rng, int* n_out, double* s_out, double* th_out,
                      uint64_t* trials) {
    double frac = rng.gen_f64();
    int n;
    if (bw_mid_contains(a, eta)) n = bw_mid_invert(a, eta, sv1, frac);
    else if (bw_hi_contains(a, eta)) n = bw_hi_invert(a, eta, sv1, frac);
    else { std::fprintf(stderr, "oracle: NBW [LP] rate lookup out of bounds\n"); std::abort(); }
    BwDdpr sp(n, a, eta);
    sp.sample(sv1, sv2, rng, s_out, th_out, trials);
    *n_out = n;
}

// pair_creation/mod.rs:21-41 (linear polarisation)
static double nbw_probability(V4 ell, SV sv, V4 k, double a, double dt, Pol pol, double pol_angle, SV* sv_out) {
    double eta = k * ell;
    SV s = in_lma_basis(sv, ell);
    s = rotate_by(s, pol_angle);
    double dphi = eta * dt / (COMPTON_TIME * ell[0]);
    SV s2;
    double prob = pol == Pol::Circular ? cpbw_probability(a, eta, s, dphi, &s2) : bw_probability(a * SQRT_2, eta, s, dphi, &s2);
    s2 = rotate_by(s2, -pol_angle);
    *sv_out = from_lma_basis(s2, ell);
    return prob;
}
// pair_creation/mod.rs:47-107 (linear polarisation)
static V4 nbw_generate(V4 ell, SV sv, V4 k, double a, Pol pol, double pol_angle, Xoshiro& rng, int* n_out) {
    double eta = k * ell;
    SV s = rotate_by(in_lma_basis(sv, ell), pol_angle);
    int n;
    double sfrac, cphi;
    if (pol == Pol::Circular) cpbw_sample(a, eta, s[1], s[2], s[3], rng, &n, &sfrac, &cphi);
    else bw_sample(a * SQRT_2, eta, s[1], s[2], rng, &n, &sfrac, &cphi, nullptr);
    double j = double(n);
    double e_zmf = std::sqrt(0.5 * j * eta);
    double p_zmf = std::sqrt(0.5 * j * eta - 1.0 - a * a);
    double cos_theta_zmf = (1.0 - 2.0 * sfrac) * e_zmf / p_zmf;
    V4 ljk = ell + j * k;
    V4 u_zmf = ljk / std::sqrt(ljk * ljk);
    V3 along = -normalize(spatial(boost_by(j * k, u_zmf)));
    V4 eps4{0.0, std::cos(pol_angle), std::sin(pol_angle), 0.0};
    V3 eps = normalize(spatial(boost_by(eps4, u_zmf)));
    eps = normalize(cross(along, cross(eps, along)));
    V3 perp = rotate_around(eps, along, cphi);
    V3 q3 = p_zmf * (cos_theta_zmf * along + std::sqrt(1.0 - powi(cos_theta_zmf, 2)) * perp);
    V4 q = with_sqr(lightlike(q3.x, q3.y, q3.z), 1.0 + a * a);
    q = boost_by(q, V4{u_zmf.t, -u_zmf.x, -u_zmf.y, -u_zmf.z});
    double s_new = 1.0 - (k * q) / (k * ell);
    double error = std::fabs(sfrac - s_new) / sfrac;
    if (!(error < 1.0e-3)) { std::fprintf(stderr, "oracle: pair_creation::generate sanity check failed\n"); std::abort(); }
    *n_out = n;
    return q;
}

// plane_wave.rs:284-316
double lma_pair_create(const Laser& L, V4 r, V4 ell, SV pol, double dt, Xoshiro& rng, double rate_increase, SV* pol_new,
                       bool* happened, PairEvent* ev) {
    // plane_wave.rs:284-316; focused_laser.rs:249-272 (no chirp)
    double a = std::sqrt(lma_a_sqd(L, r));
    double phase = L.wavevector * r;
    double chirp = 1.0 + 2.0 * L.chirp_b * phase;
    V4 kappa = L.kind == Kind::FocusedLaser ? SPEED_OF_LIGHT * COMPTON_TIME * L.wavevector
                                            : SPEED_OF_LIGHT * COMPTON_TIME * L.wavevector * chirp;
    double prob = nbw_probability(ell, pol, kappa, a, dt, L.pol, L.pol_angle, pol_new);
    double ri = prob * rate_increase > 0.1 ? 0.1 / prob : rate_increase;
    *happened = false;
    if (rng.gen_f64() < prob * ri) {
        int n;
        V4 q_p = nbw_generate(ell, pol, kappa, a, L.pol, L.pol_angle, rng, &n);
        *ev = PairEvent{ell + double(n) * kappa - q_p, q_p, 1.0 / ri, a, a * (ell * kappa), double(n) * kappa[0]};
        *happened = true;
    }
    return prob;
}

}  // namespace orc

// ---- probes
using namespace orc;
extern "C" {
int orc_lp_rate(double a, double eta, double* out) { return lp_rate(a, eta, out) ? 1 : 0; }
int orc_lp_harmonic(double a, double eta, double frac) { return get_harmonic_index(a, eta, frac); }
void orc_double_bessel(int n, double x, double y, double* j5) {
    DoubleBessel dj(n, double(n) * SQRT_2, double(n) * 0.5);
    dj.evaluate(x, y, j5);
}
void orc_double_bessel_recurrence(int n, double x, double y, double* j5) {
    DoubleBessel dj(n, std::fmax(x, double(n) * SQRT_2), std::fmax(y, double(n) * 0.5));
    dj.around(x, y, j5);
}
double orc_lp_ddpr(int n, double a, double eta, double s, double theta) {
    DoubleBessel dj(n, double(n) * SQRT_2, double(n) * 0.5);
    return ddpr(a, eta, s, theta, dj);
}
double orc_lp_partial_rate(int n, double a, double eta, double* max_out) { return partial_rate(n, a, eta, max_out); }
// (xi0, S1, S2, S3) of the quantum LP photon at harmonic n, (s, theta)
void orc_lp_stokes(int n, double a, double eta, double s, double theta, double* out4) {
    DoubleBessel dj(n, double(n) * SQRT_2, double(n) * 0.5);
    double xi0 = 0.0;
    SV sv = lp_stokes(n, a, eta, s, theta, dj, &xi0);
    out4[0] = xi0; out4[1] = sv.q; out4[2] = sv.u; out4[3] = sv.v;
}
// (1, S1, S2, S3) of the classical LP photon at harmonic n, (v, theta)
void orc_lp_cl_stokes(int n, double a, double v, double theta, double* out4) {
    DoubleBessel dj(n, double(n) * SQRT_2, double(n) * 0.5);
    SV sv = lp_cl_stokes(n, a, v, theta, dj);
    out4[0] = sv.i; out4[1] = sv.q; out4[2] = sv.u; out4[3] = sv.v;
}
double orc_lp_cl_partial_rate(int n, double a, double* max_out) { return lp_cl_partial_rate(n, a, max_out); }
double orc_lp_cl_ceiling(int n, double a) {
    DoubleBessel dj(n, double(n) * SQRT_2, double(n) * 0.5);
    return lp_cl_ceiling(a, dj);
}
double orc_lp_cl_ddpr(int n, double a, double v, double theta) {
    DoubleBessel dj(n, double(n) * SQRT_2, double(n) * 0.5);
    return lp_cl_ddpr(a, v, theta, dj);
}
double orc_lp_theta_bound_low_eta(int n, double a, double v) { return ThetaBound::for_harmonic_low_eta(n, a).at(v); }
int orc_lp_cl_rate(double a, double eta, double* out) { return lp_cl_rate(a, eta, out) ? 1 : 0; }
int orc_lp_cl_harmonic(double a, double frac) { return lp_cl_harmonic(a, frac); }
// v-marginal of the classical double-differential rate at harmonic n: the midpoint
// sum over nth theta cells of [0, pi/2] below the theta bound, at nv midpoints in v
void orc_lp_cl_v_marginal(int n, double a, int nv, int nth, double* out) {
    DoubleBessel dj(n, double(n) * SQRT_2, double(n) * 0.5);
    ThetaBound tm = ThetaBound::for_harmonic_low_eta(n, a);
    for (int i = 0; i < nv; ++i) {
        double v = (double(i) + 0.5) / double(nv), acc = 0.0, tb_v = tm.at(v);
        for (int k = 0; k < nth; ++k) {
            double th = FRAC_PI_2 * (double(k) + 0.5) / double(nth);
            if (th <= tb_v) acc += lp_cl_ddpr(a, v, th, dj);
        }
        out[i] = acc;
    }
}
void orc_lp_cl_harmonic_batch(double a, const double* frac, size_t count, int* out) {
    for (size_t i = 0; i < count; ++i) out[i] = lp_cl_harmonic(a, frac[i]);
}
void orc_lp_cl_sample(double a, double eta, uint64_t seed, int fixed_n, size_t count, double* out5) {
    Xoshiro rng(seed);
    for (size_t i = 0; i < count; ++i) {
        int n;
        double s, th;
        SV sv;
        uint64_t tr = 0;
        lp_cl_sample(a, eta, rng, fixed_n, &n, &s, &th, &sv, &tr);
        double* o = out5 + 5 * i;
        o[0] = n; o[1] = s; o[2] = th; o[3] = sv[1]; o[4] = sv[2];
    }
}
int orc_lp_sum_limit(double a, double eta) { return sum_limit(a, eta); }
double orc_lp_linear_rate(double a, double eta) { return linear_rate(a, eta); }
double orc_lp_theta_bound(int n, double a, double eta, double s) { return ThetaBound::for_harmonic(n, a, eta).at(s); }
double orc_lp_ceiling(int n, double a, double eta) {
    DoubleBessel dj(n, double(n) * SQRT_2, double(n) * 0.5);
    return ceiling_ddpr(a, eta, dj);
}
void orc_bw_partial_rate(int n, double a, double eta, int method, double* out2) {  // 0 integrate, 1 direct, 2 adaptive
    BwDdpr d(n, a, eta);
    C v = method == 1 ? d.integrate_direct() : (method == 2 ? d.integrate_adaptive() : d.integrate());
    out2[0] = v.re; out2[1] = v.im;
}
double orc_bw_ddpr(int n, double a, double eta, double s, double theta, double sv1, double sv2) {
    BwDdpr d(n, a, eta);
    return d.fully_resolved(s, theta, sv1, sv2);
}
double orc_bw_ceiling(int n, double a, double eta, double sv1, double sv2) { BwDdpr d(n, a, eta); return d.ceiling(sv1, sv2); }
double orc_bw_theta_bound(int n, double a, double eta, double s) { return BwThetaBound::for_harmonic(n, a, eta).at(s); }
int orc_bw_sum_limits(double a, double eta, int* hi) { int lo; bw_sum_limits(a, eta, &lo, hi); return lo; }
/* total rate (par, perp) -> out2; returns 0 negligible, 1 mid-range table, 2 high-range table, -1 outside */
int orc_bw_rate(double a, double eta, double* out2) {
    out2[0] = out2[1] = 0.0;
    if (bw_is_negligible(a, eta)) return 0;
    if (bw_mid_contains(a, eta)) { bw_mid_interpolate(a, eta, out2); return 1; }
    if (bw_hi_contains(a, eta)) { bw_hi_interpolate(a, eta, out2); return 2; }
    return -1;
}
double orc_bw_probability(double a, double eta, const double* sv4, double dphi, double* sv_out4) {
    SV o;
    double p = bw_probability(a, eta, SV{sv4[0], sv4[1], sv4[2], sv4[3]}, dphi, &o);
    sv_out4[0] = o.i; sv_out4[1] = o.q; sv_out4[2] = o.u; sv_out4[3] = o.v;
    return p;
}
void orc_bw_by_summation(double a, double eta, double* out48) { bw_by_summation(a, eta, out48); }
int orc_bw_harmonic(double a, double eta, double sv1, double frac) {
    if (bw_mid_contains(a, eta)) return has_table(MC) ? bw_mid_invert(a, eta, sv1, frac) : -2;
    if (bw_hi_contains(a, eta)) return bw_hi_invert(a, eta, sv1, frac);
    return -1;
}
/* count samples of (n, s, theta) */
void orc_bw_sample(double a, double eta, double sv1, double sv2, uint64_t seed, int fixed_n, size_t count, double* out3) {
    Xoshiro rng(seed);
    for (size_t i = 0; i < count; ++i) {
        int n;
        double sf, th;
        if (fixed_n > 0) {
            BwDdpr d(fixed_n, a, eta);
            d.sample(sv1, sv2, rng, &sf, &th, nullptr);
            n = fixed_n;
        } else bw_sample(a, eta, sv1, sv2, rng, &n, &sf, &th, nullptr);
        out3[3 * i] = n; out3[3 * i + 1] = sf; out3[3 * i + 2] = th;
    }
}
/* n samples of (n, s, theta, S1, S2) at (a, eta); fixed_n > 0 fixes the harmonic. */
void orc_lp_sample(double a, double eta, uint64_t seed, int fixed_n, size_t count, double* out5) {
    Xoshiro rng(seed);
    for (size_t i = 0; i < count; ++i) {
        int n;
        double s, th;
        SV sv;
        lp_sample(a, eta, rng, fixed_n, &n, &s, &th, &sv, nullptr);
        double* o = out5 + 5 * i;
        o[0] = n; o[1] = s; o[2] = th; o[3] = sv.q; o[4] = sv.u;
    }
}
}
