// Linearly polarised nonlinear Breit-Wheeler pair creation in the LMA plane wave.
//
// A photon with energy parameter eta in a wave of (rms x sqrt 2) amplitude a
// decays with the probability tabulated as a function of (a, eta) for photons
// polarised parallel and perpendicular to the wave's polarisation; the pair's
// harmonic comes from a CDF table, its lightfront fraction s and azimuth by
// rejection sampling on the double-differential rate built from J(n, x, y)
// (src/pair_creation/lp/mod.rs, lp/theta_bound.rs, lp/tables/mod.rs,
// lp/tables/mid_range/mod.rs).
//
// The rate, theta bound, ceiling and harmonic limits are __host__ __device__:
// the host-side table generator (tables/gen_bw_tables.hip) integrates the same
// functions.
#pragma once
#include "dev_lma.h"

namespace ptg {

// ---------------------------------------------------------------- table layout
// The BW pack (doubles): high-range total rate [61][25][2] (ln par, ln perp),
// high-range CDF [1525][16][3] (n, par, perp), mid-range total [157][36][2],
// mid-range CDF [5652][16][3].
namespace bw_pack {
constexpr int HI_COLS = 25, HI_ROWS = 61;
constexpr int HI_TOTAL = 0;                                   // pair_creation/lp/tables/total.rs:8
constexpr int HI_CDF = HI_TOTAL + HI_ROWS * HI_COLS * 2;       // lp/tables/cdf.rs:4
constexpr int HI_CDF_N = 1525;
constexpr int MID_COLS = 36, MID_ROWS = 157;
constexpr int MID_TOTAL = HI_CDF + HI_CDF_N * 48;              // lp/tables/mid_range/total.rs:7
constexpr int MID_CDF = MID_TOTAL + MID_ROWS * MID_COLS * 2;   // regenerated (tables/gen_bw_tables.hip)
constexpr int MID_CDF_N = MID_ROWS * MID_COLS;
constexpr int SIZE = MID_CDF + MID_CDF_N * 48;
}  // namespace bw_pack

struct BwScalars {
    double hi_ln_min_a, hi_ln_min_eta, hi_ln_a_step, hi_ln_eta_step;  // total.rs:4-7
    double hi_cdf_min[2], hi_cdf_step[2];                             // cdf.rs:2-3
    double mid_ln_min_a, mid_ln_max_eta_prime, mid_ln_min_eta_prime;  // mid_range/total.rs:1-5
    double mid_ln_a_step, mid_eta_prime_density;
};

// ---------------------------------------------------------------- harmonic range (lp/mod.rs:345-374)
PTG_HD inline void bw_sum_limits(double a, double eta, int& n_lo, int& n_hi) {
    double n_min = ceil(2.0 * (1.0 + 0.5 * a * a) / eta);
    double range = a < 1.0 ? 2.0 + (2.0 + 20.0 * a * a) * exp(-sqrt(0.5 * eta))
                           : 3.0 + 2.8 * pow(a, 8.0 / 3.0) / cbrt(eta) + 0.25 * (a * a * a) * sqrt(eta);
    double test = 0.25 - (1.0 + 0.5 * a * a) / (2.0 * n_min * eta);
    if (test <= 2.220446049250313e-16) {
        n_lo = int(n_min) + 1;
        n_hi = int(n_min + 1.0 + range);
    } else {
        n_lo = int(n_min);
        n_hi = int(n_min + range);
    }
}
PTG_HD inline bool bw_is_negligible(double a, double eta) {
    double l = log10(a) + 2.0;
    return log10(eta) < -1.0 - l * l / 4.5;
}

// ---------------------------------------------------------------- theta bound (lp/theta_bound.rs:10-100)
struct BwThetaBound { double s[16], f[16]; };
// point i of 16 (Chebyshev nodes in s over [1/2, s_max]) and its bound on theta
PTG_HD inline void bw_theta_bound_point(int n, double a, double eta, int i, double& s_i, double& f_i) {
    double nn = double(n);
    double s_min = 0.5;
    double s_max = 0.5 + sqrt(0.25 - (1.0 + 0.5 * a * a) / (2.0 * nn * eta));
    double ln_j_crit = log(0.443) - log(nn) / 3.0;
    double z = cos(PI_D * double(i) / 15.0);
    double s = s_min + 0.5 * (z + 1.0) * (s_max - s_min);
    s_i = s;
    double r_n = sqrt(2.0 * nn * eta * s * (1.0 - s) - (1.0 + 0.5 * a * a));
    double x = a * r_n / (eta * s * (1.0 - s));
    double y = a * a / (8.0 * eta * s * (1.0 - s));
    double x_crit = y > nn / 6.0 ? 4.0 * sqrt(y) * sqrt(nn - 2.0 * y) : nn + 2.0 * y;
    double ln_j_bdy = ln_dbessel_x_zero(n % 2 == 0 ? n : n + 1, y);
    ln_j_bdy = fmin(ln_j_bdy, ln_j_crit);
    double cos_theta = fmax(1.0 - x_crit * log(0.01) / (x * (ln_j_bdy - ln_j_crit)), -1.0);
    double f = acos(cos_theta);
    f_i = __builtin_isnan(f) ? PI_D : f;
}
// the bound is made non-increasing from the first point (s = s_max) on
PTG_HD inline void bw_theta_bound_running_min(double* f) {
    double f_min = f[0];
    for (int i = 1; i < 16; ++i) {
        if (f[i] < f_min) f_min = f[i];
        else f[i] = f_min;
    }
}
PTG_HD inline BwThetaBound bw_theta_bound(int n, double a, double eta) {
    BwThetaBound t;
    for (int i = 0; i < 16; ++i) bw_theta_bound_point(n, a, eta, i, t.s[i], t.f[i]);
    bw_theta_bound_running_min(t.f);
    return t;
}
PTG_HD inline double bw_theta_at(const double* ts, const double* tf, double s) {
    double val = tf[15];
    for (int i = 1; i < 16; ++i) {
        if (s > ts[i]) {
            double w = (s - ts[i - 1]) / (ts[i] - ts[i - 1]);
            val = w * tf[i] + (1.0 - w) * tf[i - 1];
            break;
        }
    }
    return fmin(val, 0.5 * PI_D);
}
PTG_HD inline double bw_theta_at(const BwThetaBound& t, double s) { return bw_theta_at(t.s, t.f, s); }

// ---------------------------------------------------------------- double-differential rate (lp/mod.rs:50-112)
PTG_HD inline void bw_s_bounds(int n, double a, double eta, double& s_min, double& s_max) {
    s_min = 0.5;
    s_max = fmax(0.5 + sqrt(0.25 - (1.0 + 0.5 * a * a) / (2.0 * double(n) * eta)), 0.5);
}
// unpolarised part and the S1, S2 coefficients: d^2W/(ds dtheta) = (u - S1 d1 - S2 d2) / 2 pi
template <class DB>
PTG_HD inline void bw_elements(const DB& D, double a, double eta, double s, double theta, double& u, double& d1,
                               double& d2) {
    double nn = double(D.n);
    double r_n_sqd = 2.0 * nn * eta * s * (1.0 - s) - (1.0 + 0.5 * a * a);
    if (!(r_n_sqd > 0.0)) { u = d1 = d2 = 0.0; return; }
    double rn = sqrt(r_n_sqd);
    double ct = cos(theta), st = sin(theta);
    double x = a * rn * ct / (eta * s * (1.0 - s));
    double y = a * a / (8.0 * eta * s * (1.0 - s));
    double j[5];
    dbessel_eval(D, x, y, j);
    double g0 = j[2], g1 = 0.5 * (j[1] + j[3]), g2 = 0.25 * (j[0] + 2.0 * j[2] + j[4]);
    double h_s = 0.5 / (s * (1.0 - s)) - 1.0;
    u = g0 * g0 + a * a * h_s * (g1 * g1 - g0 * g2);
    double c = g0 * rn * ct - g1 * a;
    d1 = c * c - (g0 * st) * (g0 * st) * r_n_sqd;
    d2 = 2.0 * g0 * rn * st * c;
}
template <class DB>
PTG_HD inline double bw_fully_resolved(const DB& D, double a, double eta, double s, double theta, double sv1, double sv2) {
    double u, d1, d2;
    bw_elements(D, a, eta, s, theta, u, d1, d2);
    return (u - sv1 * d1 - sv2 * d2) / (2.0 * PI_D);
}
// (lp/mod.rs:230-258) largest value along theta = 0 over 50 points around the peak, x 1.2
constexpr int BW_CEILING_POINTS = 50;
PTG_HD inline void bw_ceiling_range(int n, double a, double eta, double& lower, double& upper) {
    double s_min, s_max;
    bw_s_bounds(n, a, eta, s_min, s_max);
    double centre = 0.5 + sqrt(0.25 - (1.0 + a * a) / (2.0 * double(n) * eta));
    lower = s_min;
    upper = s_max;
    if (__builtin_isfinite(centre)) {
        lower = fmax(centre - 0.2, s_min);
        upper = 0.5 * (centre + s_max);
    }
}
template <class DB>
PTG_HD inline double bw_ceiling(const DB& D, double a, double eta, double sv1, double sv2) {
    double lower, upper;
    bw_ceiling_range(D.n, a, eta, lower, upper);
    double mx = -1.0e300;
    for (int i = 0; i < BW_CEILING_POINTS; ++i)
        mx = fmax(mx, bw_fully_resolved(D, a, eta, lower + (upper - lower) * double(i) / double(BW_CEILING_POINTS), 0.0, sv1, sv2));
    return 1.2 * mx;
}

// ---------------------------------------------------------------- total rate tables
__device__ __forceinline__ double bw_interp(double par, double perp, double sv1) { return 0.5 * ((par + perp) + sv1 * (par - perp)); }

// high range (lp/tables/mod.rs:18-60)
__device__ inline bool bw_hi_contains(const BwScalars& c, double a, double eta) {
    return log(a) >= c.hi_ln_min_a && a < 25.0 && log(eta) >= c.hi_ln_min_eta && eta < 2.0;
}
__device__ inline bool bw_hi_interpolate(const double* T, const BwScalars& c, double a, double eta, double f[2]) {
    using namespace bw_pack;
    double fa = (log(a) - c.hi_ln_min_a) / c.hi_ln_a_step, fe = (log(eta) - c.hi_ln_min_eta) / c.hi_ln_eta_step;
    if (!(fa >= 0.0 && fa < double(HI_COLS - 1) && fe >= 0.0 && fe < double(HI_ROWS - 1))) return false;
    int ia = int(fa), ie = int(fe);
    const double* t00 = T + (ie * HI_COLS + ia) * 2;
    const double* t10 = t00 + HI_COLS * 2;
    if (a * eta > 2.0) {  // log y against log x
        double da = fa - double(ia), de = fe - double(ie);
        for (int k = 0; k < 2; ++k)
            f[k] = exp((1.0 - da) * (1.0 - de) * t00[k] + da * (1.0 - de) * t00[2 + k] + (1.0 - da) * de * t10[k] +
                       da * de * t10[2 + k]);
    } else {  // log y against 1/x
        double a_min = exp(c.hi_ln_min_a + double(ia) * c.hi_ln_a_step), a_max = exp(c.hi_ln_min_a + double(ia + 1) * c.hi_ln_a_step);
        double e_min = exp(c.hi_ln_min_eta + double(ie) * c.hi_ln_eta_step);
        double e_max = exp(c.hi_ln_min_eta + double(ie + 1) * c.hi_ln_eta_step);
        double da = ((a_min * a_max / a) - a_min) / (a_max - a_min);
        double de = ((e_min * e_max / eta) - e_min) / (e_max - e_min);
        for (int k = 0; k < 2; ++k)
            f[k] = exp(t00[k] * da * de + t00[2 + k] * (1.0 - da) * de + t10[k] * da * (1.0 - de) + t10[2 + k] * (1.0 - da) * (1.0 - de));
    }
    return true;
}
// (lp/tables/mod.rs:62-120): inverse of the CDF in ln(n - n_min + 1) against ln(-ln(1 - F/F_max))
__device__ inline int bw_hi_invert(const double* C, const BwScalars& c, double a, double eta, double sv1, double frac) {
    using namespace bw_pack;
    double fa = (log(a) - c.hi_cdf_min[0]) / c.hi_cdf_step[0], fe = (log(eta) - c.hi_cdf_min[1]) / c.hi_cdf_step[1];
    if (!(fa >= 0.0 && fa < double(HI_COLS - 1) && fe >= 0.0 && fe < double(HI_CDF_N / HI_COLS - 1))) return 0;
    int ia = int(fa), ie = int(fe);
    double da = fa - double(ia), de = fe - double(ie);
    const int idx[4] = {HI_COLS * ie + ia, HI_COLS * ie + ia + 1, HI_COLS * (ie + 1) + ia, HI_COLS * (ie + 1) + ia + 1};
    const double w[4] = {(1.0 - da) * (1.0 - de), da * (1.0 - de), (1.0 - da) * de, da * de};
    double fr = log(-1.0 * log(1.0 - frac));
    double u = 0.0;
    for (int q = 0; q < 4; ++q) {
        const double* src = C + 48 * idx[q];
        double n_min = src[0];
        double f_max = bw_interp(src[46], src[47], sv1);
        double tbl[30];
        for (int i = 0; i < 15; ++i) {
            tbl[2 * i] = log(src[3 * i] - n_min + 1.0);
            tbl[2 * i + 1] = log(-1.0 * log(1.0 - bw_interp(src[3 * i + 1], src[3 * i + 2], sv1) / f_max));
        }
        double uq = 0.0;
        if (!pwmci_invert(tbl, 15, fr, true, uq)) return 0;
        u += uq * w[q];
    }
    int n_lo, n_hi;
    bw_sum_limits(a, eta, n_lo, n_hi);
    double n = double(n_lo) + expm1(u);
    return (n >= 0.0 && n < 1.0e7) ? int(n) + 1 : 0;
}

// mid range (lp/tables/mid_range/mod.rs)
__device__ inline bool bw_mid_contains(const BwScalars& c, double a, double eta) {
    double ln_a = log(a), ln_ep = log(eta / (1.0 + 0.5 * a * a));
    return ln_a > c.mid_ln_min_a && a < 2.0 && ln_ep > c.mid_ln_min_eta_prime && ln_ep < c.mid_ln_max_eta_prime;
}
// three-point power-law fit in eta' along column ia (mid_range/mod.rs:18-82)
__device__ inline void bw_mid_column(const double* T, const BwScalars& c, double ep, int ia, double out[2]) {
    using namespace bw_pack;
    const double D = c.mid_eta_prime_density;
    int ie = int(D * (exp(c.mid_ln_max_eta_prime) / ep - 1.0));
    if (ie % int(D) == int(D) - 1) ie = ie - 1;
    ie = ie < 0 ? 0 : (ie > MID_ROWS - 3 ? MID_ROWS - 3 : ie);
    double lo0 = 2.0 / (1.0 + double(ie + 2) / D), mid0 = 2.0 / (1.0 + double(ie + 1) / D), up0 = 2.0 / (1.0 + double(ie) / D);
    for (int k = 0; k < 2; ++k) {
        double lo1 = T[((ie + 2) * MID_COLS + ia) * 2 + k], mid1 = T[((ie + 1) * MID_COLS + ia) * 2 + k];
        double up1 = T[(ie * MID_COLS + ia) * 2 + k];
        double m = log((up1 - lo1) / (mid1 - lo1)) / log((up0 - lo0) / (mid0 - lo0));
        double alpha = (mid1 - lo1) / pow(mid0 - lo0, m);
        double v = alpha * pow(ep - lo0, m) + lo1;
        if (!__builtin_isfinite(v)) {
            if (ep < mid0) {
                double w = (log(ep) - log(lo0)) / (log(mid0) - log(lo0));
                v = (1.0 - w) * lo1 + w * mid1;
            } else {
                double w = (log(ep) - log(mid0)) / (log(up0) - log(mid0));
                v = (1.0 - w) * mid1 + w * up1;
            }
        }
        out[k] = v;
    }
}
__device__ inline void bw_mid_interpolate(const double* T, const BwScalars& c, double a, double eta, double f[2]) {
    using namespace bw_pack;
    double fa = (log(a) - c.mid_ln_min_a) / c.mid_ln_a_step;
    int ia = int(fa);
    ia = ia > MID_COLS - 2 ? MID_COLS - 2 : ia;
    double ep = eta / (1.0 + 0.5 * a * a);
    double f0[2], f1[2];
    bw_mid_column(T, c, ep, ia, f0);
    bw_mid_column(T, c, ep, ia + 1, f1);
    for (int k = 0; k < 2; ++k) {
        double v;
        if (a < 0.5) {
            double da = fa - double(ia);
            v = (1.0 - da) * f0[k] + da * f1[k];
        } else {
            double a_min = exp(c.mid_ln_min_a + double(ia) * c.mid_ln_a_step);
            double a_max = exp(c.mid_ln_min_a + double(ia + 1) * c.mid_ln_a_step);
            double da = (a - a_min) / (a_max - a_min);
            v = ((1.0 - da) * a_min * f0[k] + da * a_max * f1[k]) / a;
        }
        f[k] = exp(v);
    }
}
// bilinear CDF cell in (ln a, ln eta') and whether eta' sits next to a harmonic
// boundary; with `shift` the cell moves one row down there (invert does, the
// lowest-harmonic estimate does not: mid_range/mod.rs:109-131 against :164-179)
__device__ inline void bw_mid_cell(const BwScalars& c, double a, double eta, bool shift, int idx[4], double w[4], bool& bdy) {
    using namespace bw_pack;
    const double D = c.mid_eta_prime_density;
    double fa = (log(a) - c.mid_ln_min_a) / c.mid_ln_a_step;
    int ia = int(fa);
    ia = ia > MID_COLS - 2 ? MID_COLS - 2 : ia;
    double da = fa - double(ia);
    double ep = eta / (1.0 + 0.5 * a * a);
    int ie = int(D * (exp(c.mid_ln_max_eta_prime) / ep - 1.0));
    bdy = ie % int(D) == int(D) - 1;
    if (bdy && shift) ie = ie - 1;
    ie = ie < 0 ? 0 : (ie > MID_ROWS - 2 ? MID_ROWS - 2 : ie);
    double lower = exp(c.mid_ln_max_eta_prime) / (1.0 + double(ie + 1) / D);
    double upper = exp(c.mid_ln_max_eta_prime) / (1.0 + double(ie) / D);
    double de = (log(ep) - log(lower)) / (log(upper) - log(lower));
    idx[0] = MID_COLS * ie + ia; idx[1] = MID_COLS * ie + ia + 1;
    idx[2] = MID_COLS * (ie + 1) + ia; idx[3] = MID_COLS * (ie + 1) + ia + 1;
    w[0] = (1.0 - da) * de; w[1] = da * de; w[2] = (1.0 - da) * (1.0 - de); w[3] = da * (1.0 - de);
}
// (mid_range/mod.rs:109-272)
__device__ inline int bw_mid_invert(const double* T, const double* C, const BwScalars& c, double a, double eta, double sv1,
                                    double frac) {
    int idx[4];
    double w[4];
    bool bdy;
    bw_mid_cell(c, a, eta, true, idx, w, bdy);
    double tot[2];
    bw_mid_interpolate(T, c, a, eta, tot);
    double total = bw_interp(tot[0], tot[1], sv1);
    int n_lo, n_hi;
    bw_sum_limits(a, eta, n_lo, n_hi);
    double n = 0.0;
    if (bdy) {
        int idl[4];  // rate of the lowest harmonic (mid_range/mod.rs:109-162)
        double wl[4];
        bool unused;
        bw_mid_cell(c, a, eta, false, idl, wl, unused);
        double acc = 0.0;
        for (int q = 0; q < 4; ++q) {
            const double* t = C + 48 * idl[q];
            double fr = 0.0;
            if (!(t[0] > double(n_lo))) fr = bw_interp(t[1], t[2], sv1) / bw_interp(t[46], t[47], sv1);
            double v = fr * total;
            acc += wl[q] * v * v * v;
        }
        double p_lowest = cbrt(acc) / total;
        if (frac <= p_lowest) return n_lo;
        for (int q = 0; q < 4; ++q) {
            const double* src = C + 48 * idx[q];
            double f_max = bw_interp(src[46], src[47], sv1);
            double tbl[32];
            for (int i = 0; i < 16; ++i) {
                tbl[2 * i] = src[3 * i];
                tbl[2 * i + 1] = bw_interp(src[3 * i + 1], src[3 * i + 2], sv1) / f_max;
            }
            if (tbl[0] == double(n_lo)) {
                double t01 = tbl[1];
                for (int i = 0; i < 16; ++i) tbl[2 * i + 1] = (t01 - tbl[2 * i + 1] + p_lowest * (tbl[2 * i + 1] - 1.0)) / (t01 - 1.0);
            }
            double nq = 0.0;
            if (!pwmci_invert(tbl, 16, frac, false, nq)) return 0;
            n += nq * w[q];
        }
    } else {
        for (int q = 0; q < 4; ++q) {
            const double* src = C + 48 * idx[q];
            double f_max = bw_interp(src[46], src[47], sv1);
            double tbl[32];
            for (int i = 0; i < 16; ++i) {
                tbl[2 * i] = src[3 * i];
                tbl[2 * i + 1] = bw_interp(src[3 * i + 1], src[3 * i + 2], sv1) / f_max;
            }
            double nq = 0.0;
            if (frac <= tbl[1]) nq = tbl[0] - 0.1;
            else if (!pwmci_invert(tbl, 16, frac, false, nq)) return 0;
            n += nq * w[q];
        }
    }
    n = ceil(n);
    return (n >= 1.0 && n < 1.0e7) ? int(n) : 0;
}

// (par, perp) total rates at (a, eta); false when outside both tables (probability 0)
__device__ inline bool bw_total(const double* P, const BwScalars& c, double a, double eta, double f[2]) {
    using namespace bw_pack;
    f[0] = f[1] = 0.0;
    if (bw_is_negligible(a, eta)) return true;
    if (bw_mid_contains(c, a, eta)) { bw_mid_interpolate(P + MID_TOTAL, c, a, eta, f); return true; }
    if (bw_hi_contains(c, a, eta)) return bw_hi_interpolate(P + HI_TOTAL, c, a, eta, f);
    return false;
}
// TotalRate::probability (lp/mod.rs:385-422): probability over dphi and the Stokes
// parameters (LMA basis) after the step, whether or not the photon decays
__device__ inline double bw_probability(const double* P, const BwScalars& c, double a, double eta, sv4& sv, double dphi,
                                        bool& in_table) {
    double f[2];
    in_table = bw_total(P, c, a, eta, f);
    double prob = ALPHA * (0.5 * ((f[0] + f[1]) + sv.q * (f[0] - f[1]))) * dphi / eta;
    double prob_avg = ALPHA * (0.5 * (f[0] + f[1])) * dphi / eta;
    double delta = ALPHA * (0.5 * (f[0] - f[1])) * dphi / eta;
    double inv = 1.0 / (1.0 - prob);
    sv = {sv.i, (sv.q * (1.0 - prob_avg) - delta) * inv, sv.u * (1.0 - prob_avg) * inv, sv.v * (1.0 - prob_avg) * inv};
    return prob;
}
// harmonic index for a decay (lp/mod.rs:424-442); 0 if outside the tables
__device__ inline int bw_harmonic(const double* P, const BwScalars& c, double a, double eta, double sv1, double frac) {
    using namespace bw_pack;
    if (bw_mid_contains(c, a, eta)) return bw_mid_invert(P + MID_TOTAL, P + MID_CDF, c, a, eta, sv1, frac);
    if (bw_hi_contains(c, a, eta)) return bw_hi_invert(P + HI_CDF, c, a, eta, sv1, frac);
    return 0;
}
}  // namespace ptg
