// Circularly polarised nonlinear Compton scattering in the LMA plane wave, and
// the Bessel functions J_{n-1}, J_n, J_{n+1} it is built on.
//
// An electron of rms normalised amplitude a and energy parameter eta emits at
// harmonic n with a lightfront fraction s whose spectrum is a combination of
// J_{n-1}^2, J_n^2 and J_{n+1}^2 of one argument z(n, a, s): the harmonic comes
// from an inverse CDF table, s by 1-D rejection, the azimuth is uniform
// (src/nonlinear_compton/cp/mod.rs, cp/tables/mod.rs, cp/classical.rs,
// src/special_functions/bessel.rs).
//
// Everything here is __host__ __device__: the host-side table generator
// (tables/gen_cp_tables.hip) sums the same partial rates the device samples.
#pragma once
#include "dev_lcfa.h"
#include "dev_lma.h"

namespace ptg {

// ---------------------------------------------------------------- table layout
// The CP pack (doubles): quantum total rate [70][61] ln values, rows in ln eta,
// columns in ln a (cp/tables/total.rs:1-5); quantum CDF [70 * 61][32][2]
// (n, F(n)/F_total) on the same grid (regenerated: tables/gen_cp_tables.py);
// classical rate [61] and CDF [61][32][2] in ln a (cp/classical/rate_table.rs,
// cdf_table.rs); Gauss-Legendre nodes (16) and nodes + weights (32) on [-1, 1]
// (quadrature/tables.rs).
namespace cp_pack {
constexpr int COLS = 61, ROWS = 70;
constexpr int RATE = 0;
constexpr int CDF = RATE + COLS * ROWS;
constexpr int CDF_N = COLS * ROWS;
constexpr int CL_N = 61;
constexpr int CL_RATE = CDF + CDF_N * 64;
constexpr int CL_CDF = CL_RATE + CL_N;
constexpr int G16N = CL_CDF + CL_N * 64;
constexpr int G32N = G16N + 16;
constexpr int G32W = G32N + 32;
constexpr int BW_COLS = 74, BW_ROWS = 121;     // CP Breit-Wheeler total rate [121][74][2]
constexpr int BW_TOTAL = G32W + 32;            //   ln (S_3 = +1, S_3 = -1) (pair_creation/cp/total.rs:2-8)
// the first 31 points of every CDF row above as (ln n, ln(-ln(1 - F))), the
// coordinates its inversion interpolates in; filled on the device from the rows
// (launch_cp_cdf_logs) with the same functions the inversion used to apply per call
constexpr int CDF_LN = BW_TOTAL + BW_COLS * BW_ROWS * 2;   // [CDF_N][31][2]
constexpr int CL_CDF_LN = CDF_LN + CDF_N * 62;             // [CL_N][31][2]
constexpr int SIZE = CL_CDF_LN + CL_N * 62;
}  // namespace cp_pack

struct CpScalars {
    double min[2], step[2];               // ln a, ln eta of the quantum tables (total.rs:3-4)
    double cl_rate_min, cl_rate_step;     // classical/rate_table.rs:2-3
    double cl_cdf_min, cl_cdf_step;       // classical/cdf_table.rs:1-2
    double bw_ln_min_a, bw_ln_min_eta;    // pair_creation/cp/total.rs:4-7
    double bw_ln_a_step, bw_ln_eta_step;
};

// ---------------------------------------------------------------- Bessel J (bessel.rs)
// (J_{n-1}(x), J_n(x), J_{n+1}(x)) by Miller's backward recurrence from an even
// order a little above n, normalised by J_0 + 2 sum_k J_2k = 1 (bessel.rs:120-176).
// The reference divides (2(k+2) f) / x in every step; here the product is
// multiplied by 1/x, formed once: two FP64 divisions per step were most of a
// CP emission's work below n = 1000, and the minimal solution the backward
// recurrence follows damps the difference (the J stay within ~1e-15 of the
// divided form; x < n in the rates, where the J have no zeros).
PTG_HD inline void bessel_triple_j(int n, double x, double out[3]) {
    if (x == 0.0) {
        out[0] = n == 1 ? 1.0 : 0.0;
        out[1] = n == 0 ? 1.0 : 0.0;
        out[2] = n == -1 ? 1.0 : 0.0;
        return;
    }
    double v = x / double(n);
    int nmax = 1 + n + int(5.0 / (1.0 - 0.9 * v));
    if (((nmax % 2) + 2) % 2 != 0) nmax += 1;
    const int n_even = (((n % 2) + 2) % 2 == 0) ? 1 : 0;
    const int k_stop = n - 1 - n_even;
    const double inv_x = 1.0 / x;
    double q0 = 1.0, q1 = 0.0, q2 = 0.0, q3 = 0.0;
    double total = 0.0;
    // k runs down by 2 from nmax; the factors 2 (k + 2) and 2 (k + 1) are carried
    // as doubles, and the loops end on f2 = 2 (k + 2) (integers below 2^53:
    // stepping by 4 is exact and gives the converted values)
    double f2 = double(2 * (nmax + 2)), f1 = double(2 * (nmax + 1));
    for (const double end = double(2 * (k_stop + 2)); f2 >= end; f2 -= 4.0, f1 -= 4.0) {
        q3 = q1;
        q2 = q0;
        q1 = f2 * q2 * inv_x - q3;
        q0 = f1 * q1 * inv_x - q2;
        total += q0;
    }
    double t0 = n_even ? q1 : q0, t1 = n_even ? q2 : q1, t2 = n_even ? q3 : q2;
    for (; f2 >= 4.0; f2 -= 4.0, f1 -= 4.0) {  // k >= 0
        q3 = q1;
        q2 = q0;
        q1 = f2 * q2 * inv_x - q3;
        q0 = f1 * q1 * inv_x - q2;
        total += q0;
    }
    double norm = 2.0 * total - q0;
    if (__builtin_isnan(norm)) { out[0] = out[1] = out[2] = 0.0; return; }
    out[0] = t0 / norm; out[1] = t1 / norm; out[2] = t2 / norm;
}
// bessel.rs:178-186; t^(2/3) as cbrt(t^2) (the reference's powf, to an ulp or two)
PTG_HD inline double bessel_zeta(double z) {
    if (z < 1.0) {
        double tmp = log((1.0 + sqrt(1.0 - z * z)) / z) - sqrt(1.0 - z * z);
        double t = 3.0 * tmp / 2.0;
        return cbrt(t * t);
    }
    double tmp = sqrt(z * z - 1.0) - acos(1.0 / z);
    double t = 3.0 * tmp / 2.0;
    return -cbrt(t * t);
}
// Debye's uniform expansion through Ai and Ai' (bessel.rs:188-198); Ai of a
// negative argument is None in the reference and taken as 0. `tb` is the LCFA pack.
// The reference's powers of m (1/3, 2/3, 5/3), of (1 - z^2) (3/2) and of the
// prefactor (1/4) are formed from one cbrt and square roots (to an ulp or two).
PTG_HD inline double bessel_j_asymptotic(TP tb, int n, double x) {
    double m = double(n);
    double z = x / m;
    double zeta = bessel_zeta(z);
    double m13 = cbrt(m), m23 = m13 * m13;
    double arg = m23 * zeta;
    double ai = 0.0, aip = 0.0;
    if (!(arg < 0.0)) airy_pair(tb, arg, ai, aip);
    double w = 1.0 - z * z;
    double prefactor = 4.0 * zeta / w;
    double bk = 5.0 / (48.0 * zeta * zeta) - (2.0 + 3.0 * z * z) / (24.0 * sqrt(zeta) * (w * sqrt(w)));
    return sqrt(sqrt(prefactor)) * (ai / m13 - aip * bk / (m * m23));
}
// (J_{n-1}, J_n, J_{n+1}): recurrence below n = 1000, uniform expansion above (bessel.rs:26-33)
PTG_HD inline void bessel_j_pm(TP tb, int n, double x, double out[3]) {
    if (n < 1000) { bessel_triple_j(n, x, out); return; }
    double jn = bessel_j_asymptotic(tb, n, x), jp = bessel_j_asymptotic(tb, n + 1, x);
    out[0] = double(2 * n) * jn / x - jp;
    out[1] = jn;
    out[2] = jp;
}

// ---------------------------------------------------------------- CP spectra (cp/mod.rs)
// d(rate)/dv at harmonic n, v = s / s_max in [0, 1) (cp/mod.rs:17-38)
PTG_HD inline double cp_ddpr(TP tb, int n, double a, double eta, double v) {
    if (v < 0.0 || v >= 1.0) return 0.0;
    double dn = double(n);
    double sn = 2.0 * dn * eta / (1.0 + a * a);
    double smax = sn / (1.0 + sn);
    double vsmax = v * smax;
    double w = vsmax / (sn * (1.0 - vsmax));
    double z = sqrt(4.0 * (dn * dn) * (a * a / (1.0 + a * a)) * w * (1.0 - w));
    double j[3];
    bessel_j_pm(tb, n, z, j);
    return -smax * (j[1] * j[1] + 0.5 * a * a * (1.0 + 0.5 * (vsmax * vsmax) / (1.0 - vsmax)) *
                                      (2.0 * (j[1] * j[1]) - j[2] * j[2] - j[0] * j[0]));
}
PTG_HD __forceinline__ int cp_n_switch(double a) { return int(32.3 * (1.0 + 0.476 * pow(a, 1.56))); }
// 1.2 x the largest of 32 samples of cp_ddpr (cp/mod.rs:42-75); g16 = 16 Gauss nodes.
// Sample k < 16 lies below v_opt, k >= 16 above (point k - 16).
PTG_HD inline double cp_ceiling_arg(const double* g16, int n, double a, double eta, int k) {
    double sn = 2.0 * double(n) * eta / (1.0 + a * a);
    bool direct = sn < 2.0 || n < cp_n_switch(a);
    double v_opt = (1.0 + sn) / (2.0 + sn);
    const int i = k & 15;
    if (direct) return k < 16 ? double(i) * v_opt / 16.0 : v_opt + double(i) * (1.0 - v_opt) / 16.0;
    double u = k < 16 ? 0.5 * (1.0 + sn) * (g16[i] + 1.0) : (1.0 + sn) + 0.5 * 2.0 * (1.0 + sn) * (g16[i] + 1.0);
    return u / (1.0 + u);
}
PTG_HD inline double cp_ceiling(TP tb, const double* g16, int n, double a, double eta) {
    double lower = 0.0, upper = 0.0;
    for (int i = 0; i < 16; ++i) {
        double fl = cp_ddpr(tb, n, a, eta, cp_ceiling_arg(g16, n, a, eta, i));
        double fu = cp_ddpr(tb, n, a, eta, cp_ceiling_arg(g16, n, a, eta, 16 + i));
        lower = i == 0 ? fl : fmax(lower, fl);
        upper = i == 0 ? fu : fmax(upper, fu);
    }
    return 1.2 * fmax(lower, upper);
}
// classical (eta -> 0) spectrum in v = s / s_n (cp/mod.rs:79-89)
PTG_HD inline double cp_spectrum_low_eta(TP tb, int n, double a, double v) {
    if (v < 0.0 || v >= 1.0) return 0.0;
    double dn = double(n);
    double z = sqrt(4.0 * a * a * (dn * dn) * v * (1.0 - v) / (1.0 + a * a));
    double j[3];
    bessel_j_pm(tb, n, z, j);
    return dn * (a * a * (j[0] * j[0]) - 2.0 * (1.0 + a * a) * (j[1] * j[1]) + a * a * (j[2] * j[2])) / (1.0 + a * a);
}
// partial rate of harmonic n by 2 x 32-point Gauss-Legendre (cp/mod.rs:92-133);
// g32n, g32w = nodes and weights on [-1, 1]
PTG_HD inline double cp_partial_rate(TP tb, const double* g32n, const double* g32w, int n, double a, double eta) {
    double sn = 2.0 * double(n) * eta / (1.0 + a * a);
    double lower = 0.0, upper = 0.0;
    if (sn < 2.0 || n < cp_n_switch(a)) {
        double vmid = (1.0 + sn) / (2.0 + sn);
        for (int i = 0; i < 32; ++i) lower += 0.5 * vmid * g32w[i] * cp_ddpr(tb, n, a, eta, 0.5 * vmid * (g32n[i] + 1.0));
        for (int i = 0; i < 32; ++i)
            upper += 0.5 * (1.0 - vmid) * g32w[i] * cp_ddpr(tb, n, a, eta, vmid + 0.5 * (1.0 - vmid) * (g32n[i] + 1.0));
    } else {
        for (int i = 0; i < 32; ++i) {
            double u = 0.5 * (1.0 + sn) * (g32n[i] + 1.0);
            lower += 0.5 * (1.0 + sn) * g32w[i] * cp_ddpr(tb, n, a, eta, u / (1.0 + u)) / ((1.0 + u) * (1.0 + u));
        }
        for (int i = 0; i < 32; ++i) {
            double u = (1.0 + sn) + 0.5 * 2.0 * (1.0 + sn) * (g32n[i] + 1.0);
            upper += 0.5 * 2.0 * (1.0 + sn) * g32w[i] * cp_ddpr(tb, n, a, eta, u / (1.0 + u)) / ((1.0 + u) * (1.0 + u));
        }
    }
    return lower + upper;
}
// classical partial rate per unit eta (cp/mod.rs:137-155)
PTG_HD inline double cp_integrated_spectrum_low_eta(TP tb, const double* g32n, const double* g32w, int n, double a) {
    double lower = 0.0, upper = 0.0;
    for (int i = 0; i < 32; ++i) lower += 0.25 * g32w[i] * cp_spectrum_low_eta(tb, n, a, 0.25 * (g32n[i] + 1.0));
    for (int i = 0; i < 32; ++i) upper += 0.25 * g32w[i] * cp_spectrum_low_eta(tb, n, a, 0.25 * (g32n[i] + 3.0));
    return lower + upper;
}
// 1.1 x the largest of at most 8 samples of the classical spectrum (cp/classical.rs:46-69)
// the classical ceiling's sample points in v (cp/classical.rs:92-126): how many, and the k-th
PTG_HD inline int cp_classical_ceiling_count(int n, double a) {
    double n_switch = round(3.0 * (1.0 + 0.54 * pow(a, 2.75)));
    return (n == 1 || a < 0.15 || n > int(n_switch)) ? 1 : 8;
}
PTG_HD inline double cp_classical_ceiling_arg(int n, double a, int k) {
    double n_switch = round(3.0 * (1.0 + 0.54 * pow(a, 2.75)));
    if (n == 1) return 0.99;
    if (a < 0.15 || n > int(n_switch)) return 0.5;
    double b = a < 0.5 ? 2.3 * a : 0.8 * exp(-(a - 0.5) / 3.5) + 0.35;
    double v = 1.0 - 0.5 * pow(log(double(n)), b) / pow(log(n_switch), b);
    return fmin(v + 0.1 * double(k - 4) / 4.0, 1.0);
}
PTG_HD inline double cp_classical_ceiling(TP tb, int n, double a) {
    const int m = cp_classical_ceiling_count(n, a);
    double mx = 0.0;
    for (int k = 0; k < m; ++k) {
        double f = cp_spectrum_low_eta(tb, n, a, cp_classical_ceiling_arg(n, a, k));
        mx = k == 0 ? f : fmax(mx, f);
    }
    return 1.1 * mx;
}

// ---------------------------------------------------------------- total rates
// f(a, eta) with dP/dphase = alpha f / eta, classical limit (cp/classical.rs:22-44);
// false outside the table (the reference reports and the probability is 0)
PTG_HD inline bool cp_classical_rate(const double* P, const CpScalars& c, double a, double eta, double& out) {
    double x = log(a);
    if (x < c.cl_rate_min) { out = 2.0 * a * a * eta / 3.0; return true; }
    double fx = (x - c.cl_rate_min) / c.cl_rate_step;
    if (!(fx < double(cp_pack::CL_N - 1))) return false;
    int ix = int(fx);
    double dx = fx - double(ix);
    const double* T = P + cp_pack::CL_RATE;
    out = eta * exp((1.0 - dx) * T[ix] + dx * T[ix + 1]);
    return true;
}
// quantum (cp/mod.rs:162-176, cp/tables/mod.rs:10-34); `rt` = the [70][61] ln-rate table
PTG_HD inline bool cp_rate(const double* rt, const double* P, const CpScalars& c, double a, double eta, double& out) {
    double la = log(a), le = log(eta);
    if (le < c.min[1]) return cp_classical_rate(P, c, a, eta, out);
    if (la < c.min[0]) {
        out = a * a * (2.0 + 8.0 * eta + 9.0 * eta * eta + eta * eta * eta) / (2.0 * eta * ((1.0 + 2.0 * eta) * (1.0 + 2.0 * eta))) -
              a * a * (2.0 + 2.0 * eta - eta * eta) * log(1.0 + 2.0 * eta) / (4.0 * eta * eta);
        return true;
    }
    if (!(a < 20.0 && eta < 2.0)) return false;
    double fa = (la - c.min[0]) / c.step[0], fe = (le - c.min[1]) / c.step[1];
    int ia = int(fa), ie = int(fe);
    ia = ia > cp_pack::COLS - 2 ? cp_pack::COLS - 2 : ia;
    ie = ie > cp_pack::ROWS - 2 ? cp_pack::ROWS - 2 : ie;
    double da = fa - double(ia), de = fe - double(ie);
    const double* t0 = rt + ie * cp_pack::COLS + ia;
    const double* t1 = t0 + cp_pack::COLS;
    out = exp((1.0 - da) * (1.0 - de) * t0[0] + da * (1.0 - de) * t0[1] + (1.0 - da) * de * t1[0] + da * de * t1[1]);
    return true;
}

// cp_rate on an FP32 copy of the quantum table, in FP32; below the table in eta
// (classical rate) and a (closed form) it stays FP64
// FP32 table and interpolation; the branch and cell decisions are taken in FP64,
// as cp_rate takes them, so that both paths agree on which (a, eta) use the
// table, the explicit sum or neither.
__device__ inline bool cp_rate_f(const float* rt, const double* P, const CpScalars& c, double a, double eta, float& out) {
    const double la = log(a), le = log(eta);
    if (le < c.min[1] || la < c.min[0]) {
        double o = 0.0;
        bool ok = cp_rate(P + cp_pack::RATE, P, c, a, eta, o);
        out = float(o);
        return ok;
    }
    if (!(a < 20.0 && eta < 2.0)) return false;
    const double fa = (la - c.min[0]) / c.step[0], fe = (le - c.min[1]) / c.step[1];
    int ia = int(fa), ie = int(fe);
    ia = ia > cp_pack::COLS - 2 ? cp_pack::COLS - 2 : ia;
    ie = ie > cp_pack::ROWS - 2 ? cp_pack::ROWS - 2 : ie;
    const float da = float(fa - double(ia)), de = float(fe - double(ie));
    const float* t0 = rt + ie * cp_pack::COLS + ia;
    const float* t1 = t0 + cp_pack::COLS;
    out = expf((1.0f - da) * (1.0f - de) * t0[0] + da * (1.0f - de) * t0[1] + (1.0f - da) * de * t1[0] + da * de * t1[1]);
    return true;
}

// ---------------------------------------------------------------- harmonic index
// The k-th point of a CDF row [32][2] in the inversion's coordinates (ln n,
// ln(-ln(1 - F))) (cp/tables/mod.rs:64-76).
PTG_HD __forceinline__ void cp_cdf_log_point(const double* t32, int k, double* out2) {
    out2[0] = log(t32[2 * k]);
    out2[1] = log(-1.0 * log(1.0 - t32[2 * k + 1]));
}
// One tabulated CDF [32][2] inverted at frac: 0.9 below its first value, else the
// first 31 points in (ln n, ln(-ln(1 - F))) with linear extrapolation
// (cp/tables/mod.rs:36-44,64-76; cp/classical.rs:71-79,100-112). `l62` is the row
// in those coordinates (CDF_LN / CL_CDF_LN): 93 logarithms per row, four rows
// per event, were a quarter of a CP emission's work when formed per call.
PTG_HD inline double cp_invert_one(const double* t32, const double* l62, double frac) {
    if (frac <= t32[1]) return 0.9;
    double x = 0.0;
    pwmci_invert(l62, 31, log(-1.0 * log(1.0 - frac)), true, x);
    return exp(x);
}
// quantum harmonic (cp/tables/mod.rs:46-121); 0 if the lookup leaves the table
PTG_HD inline double cp_harmonic_cont(const double* P, const CpScalars& c, double a, double eta, double frac) {
    using namespace cp_pack;
    double la = log(a), le = log(eta);
    if (la <= c.min[0]) return 1.0;
    double fx = (la - c.min[0]) / c.step[0];
    if (!(fx < double(COLS - 1))) return 0;
    int ix = int(fx);
    double dx = fx - double(ix);
    const double* T = P + CDF;
    const double* L = P + CDF_LN;
    auto inv = [&](int row) { return cp_invert_one(T + 64 * row, L + 62 * row, frac); };
    double n;
    if (le <= c.min[1]) {
        n = inv(ix) * (1.0 - dx) + inv(ix + 1) * dx;
    } else {
        double fy = (le - c.min[1]) / c.step[1];
        if (!(fy < double(ROWS - 1))) return 0;
        int iy = int(fy);
        double dy = fy - double(iy);
        n = inv(COLS * iy + ix) * ((1.0 - dx) * (1.0 - dy)) + inv(COLS * iy + ix + 1) * (dx * (1.0 - dy)) +
            inv(COLS * (iy + 1) + ix) * ((1.0 - dx) * dy) + inv(COLS * (iy + 1) + ix + 1) * (dx * dy);
    }
    return n;
}
PTG_HD inline int cp_harmonic(const double* P, const CpScalars& c, double a, double eta, double frac) {
    double n = ceil(cp_harmonic_cont(P, c, a, eta, frac));
    return (n >= 1.0 && n <= 1.0e7) ? int(n) : 0;
}
// cp_harmonic at the hand of a whole wave (all lanes call it with the same
// arguments and return the same value): the two or four row inversions are
// independent, so lane q of every quad inverts corner q (lanes 0-3 supply the
// values) and the wave waits for one table search and root iteration instead
// of four in a row. The corners are combined as cp_harmonic_cont combines them.
__device__ inline int cp_harmonic_wave(const double* P, const CpScalars& c, double a, double eta, double frac) {
    using namespace cp_pack;
    const int q = int(__lane_id()) & 3;
    double la = log(a), le = log(eta);
    if (la <= c.min[0]) return 1;
    double fx = (la - c.min[0]) / c.step[0];
    if (!(fx < double(COLS - 1))) return 0;
    int ix = int(fx);
    double dx = fx - double(ix);
    const double* T = P + CDF;
    const double* L = P + CDF_LN;
    auto inv = [&](int row) { return cp_invert_one(T + 64 * row, L + 62 * row, frac); };
    double n;
    if (le <= c.min[1]) {
        const double v = inv(ix + (q & 1));
        const double v0 = __shfl(v, 0, 64), v1 = __shfl(v, 1, 64);
        n = v0 * (1.0 - dx) + v1 * dx;
    } else {
        double fy = (le - c.min[1]) / c.step[1];
        if (!(fy < double(ROWS - 1))) return 0;
        int iy = int(fy);
        double dy = fy - double(iy);
        const double v = inv(COLS * (iy + (q >> 1)) + ix + (q & 1));
        const double v0 = __shfl(v, 0, 64), v1 = __shfl(v, 1, 64), v2 = __shfl(v, 2, 64), v3 = __shfl(v, 3, 64);
        n = v0 * ((1.0 - dx) * (1.0 - dy)) + v1 * (dx * (1.0 - dy)) + v2 * ((1.0 - dx) * dy) + v3 * (dx * dy);
    }
    n = ceil(n);
    return (n >= 1.0 && n <= 1.0e7) ? int(n) : 0;
}
// classical harmonic (cp/classical.rs:83-123)
PTG_HD inline int cp_classical_harmonic(const double* P, const CpScalars& c, double a, double frac) {
    using namespace cp_pack;
    double la = log(a);
    if (la <= c.cl_cdf_min) return 1;
    double fx = (la - c.cl_cdf_min) / c.cl_cdf_step;
    if (!(fx < double(CL_N - 1))) return 0;
    int ix = int(fx);
    double dx = fx - double(ix);
    const double* T = P + CL_CDF;
    const double* L = P + CL_CDF_LN;
    double n = cp_invert_one(T + 64 * ix, L + 62 * ix, frac) * (1.0 - dx) +
               cp_invert_one(T + 64 * (ix + 1), L + 62 * (ix + 1), frac) * dx;
    n = ceil(n);
    return (n >= 1.0 && n <= 1.0e7) ? int(n) : 0;
}

// ---------------------------------------------------------------- photon Stokes parameters
// In the basis (k x k', k'_perp), then rotated to (e_1, e_2) by pi/2 + theta
// (cp/mod.rs:203-232, cp/classical.rs:155-175). The reference forms 4 n^2 in
// 32-bit integer arithmetic, which wraps for n > 23170; reproduced.
PTG_HD __forceinline__ double cp_four_n_sqd(int n) { return double(int32_t(uint32_t(4) * uint32_t(n) * uint32_t(n))); }
PTG_HD inline void cp_stokes(TP tb, int n, double a, double s, double sn, double v, double theta, bool classical, double sv[4]) {
    constexpr double LASER_S3 = -1.0;  // cp/mod.rs:12
    double z = classical ? sqrt(cp_four_n_sqd(n) * (a * a / (1.0 + a * a)) * v * (1.0 - v))
                         : sqrt(cp_four_n_sqd(n) * (a * a / (1.0 + a * a)) * (s / (sn * (1.0 - s))) * (1.0 - s / (sn * (1.0 - s))));
    double j[3];
    bessel_j_pm(tb, n, z, j);
    double jm = j[0], jn = j[1], jp = j[2];
    double base = jm * jm + jp * jp - 2.0 * jn * jn;
    double ja = 2.0 * jn / a;
    double xi0, xi3;
    if (classical) {
        xi0 = 2.0 * base - ja * ja;
        xi3 = 2.0 * (1.0 - 2.0 * v) * (jm * jm - jp * jp);
    } else {
        xi0 = (1.0 - s + 1.0 / (1.0 - s)) * base - ja * ja;
        xi3 = LASER_S3 * (1.0 - s + 1.0 / (1.0 - s)) * (1.0 - 2.0 * s / (sn * (1.0 - s))) * (jm * jm - jp * jp);
    }
    double nz = double(n) / z;
    double xi1 = 2.0 * base + 8.0 * jn * jn * (1.0 - nz * nz - 0.5 / (a * a));
    double q = xi1 / xi0, u = 0.0 / xi0, w = xi3 / xi0;
    double s2, c2;
    sincos(2.0 * (0.5 * PI_D + theta), &s2, &c2);  // polarization.rs:44-53
    sv[0] = 1.0;
    sv[1] = c2 * q + s2 * u;
    sv[2] = -s2 * q + c2 * u;
    sv[3] = w;
}

// ---------------------------------------------------------------- CP nonlinear Breit-Wheeler (pair_creation/cp/mod.rs)
constexpr double CP_LASER_S3 = -1.0;  // pair_creation/cp/mod.rs:13
// s range at harmonic n: [s_min, 1/2], the other half by symmetry (:35-40)
PTG_HD inline void cpbw_s_bounds(int n, double a, double eta, double& s_min, double& s_max) {
    double sn = 2.0 * double(n) * eta / (1.0 + a * a);
    s_min = fmin(0.5 * (1.0 - sqrt(1.0 - 4.0 / sn)), 0.5);
    s_max = 0.5;
}
// unpolarised, S_1/S_2 and S_3 coefficients of the double-differential rate (:45-66)
PTG_HD inline void cpbw_elements(TP tb, int n, double a, double eta, double s, double e[3]) {
    double ell = double(n);
    double a2 = a * a;
    double sn = 2.0 * ell * eta / (1.0 + a2);
    double tmp = 1.0 / (sn * s * (1.0 - s));
    double z = sqrt((4.0 * ell * ell) * (a2 / (1.0 + a2)) * tmp * (1.0 - tmp));
    double j[3];
    bessel_j_pm(tb, n, z, j);
    double jm2 = j[0] * j[0], jn2 = j[1] * j[1], jp2 = j[2] * j[2];
    double sym = (1.0 - s) / s + s / (1.0 - s);
    double ela = ell * a / z;
    e[0] = jn2 + 0.25 * a2 * sym * (jm2 + jp2 - 2.0 * jn2);
    e[1] = (2.0 * (ela * ela) - (1.0 + a2)) * jn2 - 0.5 * a2 * (jm2 + jp2);
    e[2] = 0.25 * a2 * sym * (1.0 - 2.0 / (sn * s * (1.0 - s))) * (jm2 - jp2);
}
// rate differential in s and theta for Stokes (S_1, S_2, S_3) (:71-78); the reference
// binds (sin 2theta, cos 2theta) to names in the opposite order, kept as written
PTG_HD inline double cpbw_fully_resolved(TP tb, int n, double a, double eta, double s, double theta, double sv1, double sv2,
                                         double sv3) {
    double e[3];
    cpbw_elements(tb, n, a, eta, s, e);
    double first, second;
    sincos(2.0 * theta, &first, &second);
    return (e[0] - (sv1 * first + sv2 * second) * e[1] - CP_LASER_S3 * sv3 * e[2]) / (2.0 * PI_D);
}
// integrated over theta at S_3 = +1 (re) and -1 (im) (:80-93)
PTG_HD __forceinline__ cplx cpbw_over_theta(TP tb, int n, double a, double eta, double s) {
    double e[3];
    cpbw_elements(tb, n, a, eta, s, e);
    return cplx{e[0] - CP_LASER_S3 * e[2], e[0] + CP_LASER_S3 * e[2]};
}
// partial rate of harmonic n at S_3 = +1 (re) and -1 (im) by 32-point Gauss-Legendre
// on one to three s intervals split at the estimated peak (:98-184)
PTG_HD inline cplx cpbw_partial_rate(TP tb, const double* g32n, const double* g32w, int n, double a, double eta) {
    double ell = double(n);
    double sn = 2.0 * ell * eta / (1.0 + a * a);
    if (sn <= 4.0) return cplx{0.0, 0.0};
    double a2m2 = a * a - 2.0;
    double ie = 1.0 / ell + 1.0;
    double beta = (a * a) * (a * a) * (ie * ie) + 16.0 * (a2m2 * a2m2) / (sn * sn) - 8.0 * a * a * a2m2 / sn;
    beta = sqrt(beta) / a2m2;
    double alpha = (a * a + 2.0 * ell) / (ell * (2.0 - a * a)) - 4.0 / sn;
    double s_peak = 0.5 * (1.0 - sqrt(alpha + beta));
    double s_min, s_max;
    cpbw_s_bounds(n, a, eta, s_min, s_max);
    auto seg = [&](double lo, double hi) {
        cplx acc{0.0, 0.0};
        for (int i = 0; i < 32; ++i) {
            double s = 0.5 * (hi - lo) * g32n[i] + 0.5 * (lo + hi);
            acc = acc + (0.5 * (hi - lo) * g32w[i]) * cpbw_over_theta(tb, n, a, eta, s);
        }
        return acc;
    };
    if (__builtin_isfinite(s_peak)) {
        double s_mid = 2.0 * s_peak - s_min;
        if (s_mid > s_max) {
            cplx lower = seg(s_min, s_peak), upper = seg(s_peak, s_max);
            return 2.0 * (upper + lower);
        }
        cplx lower = seg(s_min, s_peak), middle = seg(s_peak, s_mid), upper = seg(s_mid, s_max);
        return 2.0 * ((lower + middle) + upper);
    }
    return 2.0 * seg(s_min, s_max);
}
// harmonics contributing to the total (:368-381)
PTG_HD inline void cpbw_sum_limits(double a, double eta, int& lo, int& hi) {
    double n_min = ceil(2.0 * (1.0 + a * a) / eta);
    double range = 2.0 * (0.25 + a * a) / eta + 13.0 * a * (2.0 + a);
    double test = 0.25 - (1.0 + a * a) / (2.0 * n_min * eta);
    if (test <= 2.220446049250313e-16) { lo = int(n_min) + 1; hi = int(n_min + 1.0 + range); }
    else { lo = int(n_min); hi = int(n_min + range); }
}
PTG_HD __forceinline__ bool cpbw_is_too_small(double a, double eta) {
    double l = log10(a) + 2.0;
    return log10(eta) < -1.0 - l * l / 4.5;
}
// table lookup: ln f bilinear in (ln a, ln eta) where a eta > 2, else 1 / ln f
// bilinear in (a, eta) (:406-449); false outside the table
PTG_HD inline bool cpbw_by_lookup(const double* P, const CpScalars& c, double a, double eta, double f[2]) {
    using namespace cp_pack;
    double x = log(a), y = log(eta);
    if (y < c.bw_ln_min_eta) { f[0] = f[1] = 0.0; return true; }
    double fx = (x - c.bw_ln_min_a) / c.bw_ln_a_step, fy = (y - c.bw_ln_min_eta) / c.bw_ln_eta_step;
    if (!(fx >= 0.0 && fx < double(BW_COLS - 1) && fy < double(BW_ROWS - 1))) return false;
    int ix = int(fx), iy = int(fy);
    const double* t00 = P + BW_TOTAL + (iy * BW_COLS + ix) * 2;
    const double* t10 = t00 + BW_COLS * 2;
    if (a * eta > 2.0) {
        double dx = fx - double(ix), dy = fy - double(iy);
        double w0 = (1.0 - dx) * (1.0 - dy), w1 = dx * (1.0 - dy), w2 = (1.0 - dx) * dy, w3 = dx * dy;
        for (int k = 0; k < 2; ++k) f[k] = exp(w0 * t00[k] + w1 * t00[2 + k] + w2 * t10[k] + w3 * t10[2 + k]);
    } else {
        double a_min = exp(c.bw_ln_min_a + double(ix) * c.bw_ln_a_step), a_max = exp(c.bw_ln_min_a + double(ix + 1) * c.bw_ln_a_step);
        double e_min = exp(c.bw_ln_min_eta + double(iy) * c.bw_ln_eta_step);
        double e_max = exp(c.bw_ln_min_eta + double(iy + 1) * c.bw_ln_eta_step);
        double dx = (a - a_min) / (a_max - a_min), dy = (eta - e_min) / (e_max - e_min);
        double w0 = (1.0 - dx) * (1.0 - dy), w1 = dx * (1.0 - dy), w2 = (1.0 - dx) * dy, w3 = dx * dy;
        for (int k = 0; k < 2; ++k) {
            double g = w0 / t00[k] + w1 / t00[2 + k] + w2 / t10[k] + w3 / t10[2 + k];
            f[k] = exp(1.0 / g);
        }
    }
    return true;
}
// (S_3 = +1, S_3 = -1) totals (:271-281): 0 below a = 0.02 or where negligible,
// the explicit sum over harmonics below the table's a range, else the table
PTG_HD inline bool cpbw_total(TP tb, const double* P, const CpScalars& c, double a, double eta, double f[2]) {
    f[0] = f[1] = 0.0;
    if (a < 0.02 || cpbw_is_too_small(a, eta)) return true;
    if (a < exp(c.bw_ln_min_a)) {
        int lo, hi;
        cpbw_sum_limits(a, eta, lo, hi);
        cplx t{0.0, 0.0};
        for (int n = lo; n <= hi; ++n) t = t + cpbw_partial_rate(tb, P + cp_pack::G32N, P + cp_pack::G32W, n, a, eta);
        f[0] = t.re;
        f[1] = t.im;
        return true;
    }
    return cpbw_by_lookup(P, c, a, eta, f);
}
// probability over dphi and the Stokes parameters after the step (:286-322)
PTG_HD inline double cpbw_probability(TP tb, const double* P, const CpScalars& c, double a, double eta, sv4& sv, double dphi,
                                      bool& in_table) {
    double f[2];
    in_table = cpbw_total(tb, P, c, a, eta, f);
    double prob = ALPHA * (0.5 * ((f[0] + f[1]) + sv.v * (f[0] - f[1]))) * dphi / eta;
    double prob_avg = ALPHA * (0.5 * (f[0] + f[1])) * dphi / eta;
    double delta = ALPHA * (0.5 * (f[0] - f[1])) * dphi / eta;
    sv = {sv.i, sv.q * (1.0 - prob_avg) / (1.0 - prob), sv.u * (1.0 - prob_avg) / (1.0 - prob),
          (sv.v * (1.0 - prob_avg) - delta) / (1.0 - prob)};
    return prob;
}
// ---------------------------------------------------------------- sampler
// (s, theta, Stokes) of one emission at harmonic n (cp/mod.rs:181-234,
// cp/classical.rs:128-178): v = s / s_max (quantum) or s / s_n (classical) by
// rejection under the tabulated ceiling, theta uniform. `cap` bounds the trials;
// hitting it sets `capped` and keeps the last trial.
__device__ inline void cp_sample_at(TP tb, const double* P, int n, double a, double eta, bool classical, Rng& rng, double& s,
                                    double& theta, sv4& sv, uint32_t& trials, bool& capped) {
    double mx = classical ? cp_classical_ceiling(tb, n, a) : cp_ceiling(tb, P + cp_pack::G16N, n, a, eta);
    double v = 0.0;
    const uint32_t cap = 1u << 20;
    for (uint32_t t = 0;; ++t) {
        double u;
        rng_uniform2(rng, v, u);
        ++trials;
        if (t >= cap) { capped = true; break; }
        double f = classical ? cp_spectrum_low_eta(tb, n, a, v) : cp_ddpr(tb, n, a, eta, v);
        if (u <= f / mx) break;
    }
    double sn = 2.0 * double(n) * eta / (1.0 + a * a);
    s = classical ? v * sn : v * (sn / (1.0 + sn));
    theta = 2.0 * PI_D * rng_uniform(rng);
    double st[4];
    cp_stokes(tb, n, a, s, sn, v, theta, classical, st);
    sv = {st[0], st[1], st[2], st[3]};
}

// (s, theta) of one pair at harmonic n (pair_creation/cp/mod.rs:188-245): s by
// rejection under 1.5 x the largest of 32 samples, mirrored into (1/2, 1) with
// probability 1/2; theta by rejection under 1.01 x the maximum of the
// theta-dependence, found at tan 2 theta = S_2 / S_1
__device__ inline void cpbw_sample_at(TP tb, const double* P, int n, double a, double eta, double sv1, double sv2, double sv3, Rng& rng,
                                      double& s_out, double& theta_out, uint32_t& trials, bool& capped) {
    const double* g32n = P + cp_pack::G32N;
    double s_min, s_max;
    cpbw_s_bounds(n, a, eta, s_min, s_max);
    double mx = NAN;
    for (int i = 0; i < 32; ++i) {
        cplx sp = cpbw_over_theta(tb, n, a, eta, 0.5 * (s_max - s_min) * g32n[i] + 0.5 * (s_min + s_max));
        mx = fmax(mx, 0.5 * (sp.re * (1.0 + sv3) + sp.im * (1.0 - sv3)));
    }
    mx = 1.5 * mx;
    const uint32_t cap = 1u << 20;
    double s = s_min;
    for (uint32_t t = 0;; ++t) {
        double u1, u2;
        rng_uniform2(rng, u1, u2);
        ++trials;
        s = s_min + (s_max - s_min) * u1;
        if (t >= cap) { capped = true; break; }
        cplx sp = cpbw_over_theta(tb, n, a, eta, s);
        if (mx * u2 < 0.5 * (sp.re * (1.0 + sv3) + sp.im * (1.0 - sv3))) break;
    }
    if (int(rng_uniform(rng) * 2.0) == 0) s = 1.0 - s;
    double theta_opt = 0.5 * atan2(sv2, sv1);
    double mt = 0.0;
    for (int i = 0; i < 4; ++i) {
        double f = cpbw_fully_resolved(tb, n, a, eta, s, theta_opt + double(i) * (PI_D / 2.0), sv1, sv2, sv3);
        mt = i == 0 ? f : fmax(mt, f);
    }
    mt = 1.01 * mt;
    double theta = 0.0;
    for (uint32_t t = 0;; ++t) {
        double u1, u2;
        rng_uniform2(rng, u1, u2);
        ++trials;
        theta = 2.0 * PI_D * u1;
        if (t >= cap) { capped = true; break; }
        if (mt * u2 < cpbw_fully_resolved(tb, n, a, eta, s, theta, sv1, sv2, sv3)) break;
    }
    s_out = s;
    theta_out = theta;
}

}  // namespace ptg
