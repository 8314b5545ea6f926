// Ptarmigan CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.hpp).
// Locally-monochromatic (LMA) plane wave (src/field/plane_wave.rs), linearly
// polarised nonlinear Compton rates and sampler (src/nonlinear_compton/mod.rs,
// lp/mod.rs, lp/linear.rs) and the generalised Bessel function J(n, x, y)
// (src/special_functions/double_bessel.rs).
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
    *n_out = n;
    *s_out = s;
    *theta_out = theta;
    *pol_out = {1.0, xi1 / xi0, xi2 / xi0, 0.0};
}

// nonlinear_compton/mod.rs:19-47 (linear polarisation, quantum)
static bool nlc_probability(V4 k, V4 q, double dt, Pol pol, double* out) {
    double a_sqd = q * q - 1.0;
    double a = a_sqd >= 0.0 ? std::sqrt(a_sqd) : 0.0;
    double eta = k * q;
    double dphi = dt * eta / (COMPTON_TIME * q[0]);
    if (pol != Pol::Linear) { std::fprintf(stderr, "oracle: circular LMA emission not restated\n"); std::abort(); }
    double f;
    if (!lp_rate(a * SQRT_2, eta, &f)) return false;
    *out = ALPHA_FINE * f * dphi / eta;
    return true;
}

// nonlinear_compton/mod.rs:58-138 (linear polarisation, quantum)
static void nlc_generate(V4 k, V4 q, double pol_angle, Xoshiro& rng, int* n_out, V4* k_out, SV* sv_out, uint64_t* trials) {
    double a = std::sqrt(q * q - 1.0);
    double eta = k * q;
    int n;
    double s, cphi;
    SV sv;
    lp_sample(a * SQRT_2, eta, rng, 0, &n, &s, &cphi, &sv, trials);
    double ell = double(n);
    double p_zmf = ell * eta / std::sqrt(1.0 + a * a + 2.0 * ell * eta);
    double cos_theta_zmf = 1.0 - s * (1.0 + a * a + 2.0 * ell * eta) / (ell * eta);
    V4 qk = q + ell * k;
    V4 u_zmf = qk / std::sqrt(qk * qk);
    V3 along = -normalize(spatial(boost_by(q, u_zmf)));
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

// plane_wave.rs:152-175
static V4 pw_ll_force(const Laser& L, V4 r, V4 u) {
    double eta = SPEED_OF_LIGHT * COMPTON_TIME * (L.wavevector * u);
    return -2.0 * ALPHA_FINE * pw_a_sqd(L, r) * powi(eta, 2) * u / (3.0 * COMPTON_TIME);
}
static double pw_ll_work(const Laser& L, V4 r, V4 u) {
    double eta = SPEED_OF_LIGHT * COMPTON_TIME * (L.wavevector * u);
    double omega = SPEED_OF_LIGHT * L.wavevector[0];
    double delta = L.pol == Pol::Circular ? 1.0 : 0.75;
    return 2.0 * ALPHA_FINE * omega * eta * delta * powi(pw_a_sqd(L, r), 2) / 3.0;
}

static double pw_bandwidth(const Laser& L) {  // plane_wave.rs:49-63
    if (!L.finite_bandwidth) return 0.0;
    double n_fwhm = L.env == Env::CosSquared ? 0.36405666377387671305 * L.n_cycles : L.n_cycles;
    return std::sqrt(0.5 * LN_2) / (PI * n_fwhm);
}

double lma_max_timestep(const Laser& L) {  // plane_wave.rs:176-182
    if (L.kind != Kind::PlaneWave) { std::fprintf(stderr, "oracle: FocusedLaser not restated\n"); std::abort(); }
    return L.env == Env::Flattop ? 0.2 / (SPEED_OF_LIGHT * L.wavevector[0]) : 1.0 / (SPEED_OF_LIGHT * L.wavevector[0]);
}
bool lma_contains(const Laser& L, V4 r) {  // plane_wave.rs:184-192
    double phase = L.wavevector * r;
    double mp = L.env == Env::CosSquared ? PI * L.n_cycles : (L.env == Env::Flattop ? PI * (L.n_cycles + 1.0) : 6.0 * PI * L.n_cycles);
    return phase < mp;
}
double lma_ideal_initial_z(const Laser& L) {  // plane_wave.rs:318-325
    double wl = 2.0 * PI / L.wavevector[0];
    return L.env == Env::CosSquared ? 0.5 * wl * L.n_cycles : (L.env == Env::Flattop ? 0.5 * wl * (L.n_cycles + 1.0) : 2.0 * wl * L.n_cycles);
}

// plane_wave.rs:196-249
PushOut lma_push(const Laser& L, V4 r, V4 u, double rqm, double dt, Eqn eqn) {
    double dtau = dt / u[0];
    double scale = powi(rqm / (ELECTRON_CHARGE / ELECTRON_MASS), 2);
    V4 u_mid{0, 0, 0, 0};
    if (eqn == Eqn::LandauLifshitz) {
        V4 f = 0.5 * SPEED_OF_LIGHT * scale * pw_grad_a_sqd(L, r);
        V4 g = powi(scale, 2) * pw_ll_force(L, r, u);
        u_mid = u + 0.5 * (f + g) * dtau;
    } else if (eqn == Eqn::ModifiedLandauLifshitz) {
        std::fprintf(stderr, "oracle: Gaunt factor correction is unavailable in LMA mode\n");
        std::abort();
    }
    r = r + 0.5 * SPEED_OF_LIGHT * u * dtau;
    double dt_actual = 0.5 * u[0] * dtau;
    V4 f = 0.5 * SPEED_OF_LIGHT * scale * pw_grad_a_sqd(L, r);
    V4 g = eqn == Eqn::LandauLifshitz ? powi(scale, 2) * pw_ll_force(L, r, u_mid) : V4{0, 0, 0, 0};
    u = u + (f + g) * dtau;
    double dwork = eqn == Eqn::LandauLifshitz ? (f[0] + pw_ll_work(L, r, u_mid)) * dtau : f[0] * dtau;
    r = r + 0.5 * SPEED_OF_LIGHT * u * dtau;
    dt_actual = dt_actual + 0.5 * u[0] * dtau;
    u = with_sqr(u, 1.0 + pw_a_sqd(L, r));
    return {r, u, dt_actual, dwork};
}

// plane_wave.rs:251-282
bool lma_radiate(const Laser& L, V4 r, V4 u, double dt, Xoshiro& rng, bool classical, RadEvent* ev) {
    if (classical) { std::fprintf(stderr, "oracle: classical LMA emission not restated\n"); std::abort(); }
    double a = std::sqrt(pw_a_sqd(L, r));
    double phase = L.wavevector * r;
    double chirp = 1.0 + 2.0 * L.chirp_b * phase;
    double width = 1.0 + pw_bandwidth(L) * rng.normal();
    V4 kappa = SPEED_OF_LIGHT * COMPTON_TIME * L.wavevector * chirp * width;
    double prob = 0.0;
    if (!nlc_probability(kappa, u, dt, L.pol, &prob)) prob = 0.0;
    if (!(rng.gen_f64() < prob)) return false;
    int n;
    V4 k;
    SV pol;
    nlc_generate(kappa, u, L.pol_angle, rng, &n, &k, &pol, nullptr);
    *ev = RadEvent{k, u + double(n) * kappa - k, pol, a, a * (u * kappa), double(n) * kappa[0]};
    return true;
}

double lma_pair_create(const Laser&, V4, V4, SV, double, Xoshiro&, double, SV*, bool*, PairEvent*) {
    std::fprintf(stderr, "oracle: LMA pair creation not restated\n");
    std::abort();
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
int orc_lp_sum_limit(double a, double eta) { return sum_limit(a, eta); }
double orc_lp_linear_rate(double a, double eta) { return linear_rate(a, eta); }
double orc_lp_theta_bound(int n, double a, double eta, double s) { return ThetaBound::for_harmonic(n, a, eta).at(s); }
double orc_lp_ceiling(int n, double a, double eta) {
    DoubleBessel dj(n, double(n) * SQRT_2, double(n) * 0.5);
    return ceiling_ddpr(a, eta, dj);
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
