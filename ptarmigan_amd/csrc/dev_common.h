// Device-side primitives for the particle-shower kernels (gfx950).
//
// Vector algebra, complex arithmetic, constants, the counter-based RNG and
// monotone-cubic (PWMCI) table interpolation. Everything here is __device__
// inline code meant to stay in registers inside one-thread-per-particle loops.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ptg {

// Pure-math helpers that the host-side table generator shares with the kernels.
#define PTG_HD __host__ __device__

// Physical constants, SI, 7 significant digits as in the reference (src/constants.rs:4-29).
constexpr double C_LIGHT = 2.997925e8;
constexpr double C_LIGHT_SQD = 89875517873681764.0;
constexpr double E_CHARGE = 1.602177e-19;  // elementary charge
constexpr double M_ELECTRON = 9.109383e-31;
constexpr double M_E_MEV = 0.510999;
constexpr double E_CRIT = 1.323285e18;
constexpr double ALPHA = 7.29735257e-3;
constexpr double T_COMPTON = 1.28808867e-21;
constexpr double RQM_ELECTRON = -1.602177e-19 / 9.109383e-31;
constexpr double PI_D = 3.14159265358979323846;
constexpr double LN2_D = 0.693147180559945309417;
constexpr double SQRT3_D = 1.73205080756887729353;

// ---------------------------------------------------------------- sin/cos of the carrier phase
// sin and cos of x for |x| < 2^20 pi/2 (carrier phases are at most ~1e4 rad):
// Cody-Waite reduction by pi/2 split into three parts (33-bit heads, so n*P1 and
// n*P2 are exact), then the Taylor series of sin and cos on [-pi/4, pi/4]
// through r^19 and r^20 (truncation < 1e-21). About twice as accurate as the
// 1-ulp target and much lighter on registers than a general-argument sincos.
__device__ __forceinline__ void sincos_phase(double x, double* s, double* c) {
    const double TWO_OVER_PI = 0.63661977236758138;
    const double P1 = 1.5707963267341256, P2 = 6.077100506303966e-11, P3 = 2.0222662487959506e-21;
    double n = rint(x * TWO_OVER_PI);
    double r = fma(-n, P1, x);
    r = fma(-n, P2, r);
    r = fma(-n, P3, r);
    double r2 = r * r;
    double ps = 8.2206352466243295e-18;
    ps = fma(ps, -r2, 2.8114572543455206e-15);
    ps = fma(ps, -r2, 7.6471637318198164e-13);
    ps = fma(ps, -r2, 1.6059043836821613e-10);
    ps = fma(ps, -r2, 2.505210838544172e-08);
    ps = fma(ps, -r2, 2.7557319223985893e-06);
    ps = fma(ps, -r2, 0.00019841269841269841);
    ps = fma(ps, -r2, 0.0083333333333333332);
    ps = fma(ps, -r2, 0.16666666666666666);
    double sr = fma(-r * r2, ps, r);  // r - r^3/3! + ...
    double pc = 4.1103176233121648e-19;
    pc = fma(pc, -r2, 1.5619206968586225e-16);
    pc = fma(pc, -r2, 4.7794773323873853e-14);
    pc = fma(pc, -r2, 1.1470745597729725e-11);
    pc = fma(pc, -r2, 2.08767569878681e-09);
    pc = fma(pc, -r2, 2.7557319223985888e-07);
    pc = fma(pc, -r2, 2.4801587301587302e-05);
    pc = fma(pc, -r2, 0.0013888888888888889);
    pc = fma(pc, -r2, 0.041666666666666664);
    pc = fma(pc, -r2, 0.5);
    double cr = fma(-r2, pc, 1.0);  // 1 - r^2/2! + ...
    int q = int(n) & 3;
    double ss = (q & 1) ? cr : sr, cc = (q & 1) ? sr : cr;
    *s = (q & 2) ? -ss : ss;
    *c = ((q + 1) & 2) ? -cc : cc;
}

// ---------------------------------------------------------------- 3- and 4-vectors
struct d3 { double x, y, z; };
__device__ __forceinline__ d3 operator+(d3 a, d3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ d3 operator-(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ d3 operator*(double s, d3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ d3 operator-(d3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ d3 cross(d3 a, d3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
__device__ __forceinline__ double nsq(d3 a) { return dot(a, a); }
__device__ __forceinline__ d3 unit(d3 a) { return (1.0 / sqrt(nsq(a))) * a; }
// Right-handed rotation of v about the unit axis `ax` by angle (sin s, cos c).
__device__ __forceinline__ d3 rotate(d3 v, d3 ax, double s, double c) {
    double k = 1.0 - c;
    return {(c + ax.x * ax.x * k) * v.x + (ax.x * ax.y * k - ax.z * s) * v.y + (ax.x * ax.z * k + ax.y * s) * v.z,
            (ax.y * ax.x * k + ax.z * s) * v.x + (c + ax.y * ax.y * k) * v.y + (ax.y * ax.z * k - ax.x * s) * v.z,
            (ax.z * ax.x * k - ax.y * s) * v.x + (ax.z * ax.y * k + ax.x * s) * v.y + (c + ax.z * ax.z * k) * v.z};
}
__device__ __forceinline__ d3 rot_z(d3 v, double s, double c) { return {c * v.x - s * v.y, c * v.y + s * v.x, v.z}; }

struct d4 { double t, x, y, z; };
__device__ __forceinline__ d3 sp(d4 a) { return {a.x, a.y, a.z}; }
__device__ __forceinline__ double mdot(d4 a, d4 b) { return a.t * b.t - a.x * b.x - a.y * b.y - a.z * b.z; }

// ---------------------------------------------------------------- complex
struct cplx { double re, im; };
PTG_HD __forceinline__ cplx operator+(cplx a, cplx b) { return {a.re + b.re, a.im + b.im}; }
PTG_HD __forceinline__ cplx operator-(cplx a, cplx b) { return {a.re - b.re, a.im - b.im}; }
PTG_HD __forceinline__ cplx operator*(cplx a, cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
PTG_HD __forceinline__ cplx operator*(double s, cplx a) { return {s * a.re, s * a.im}; }
PTG_HD __forceinline__ cplx operator+(double s, cplx a) { return {s + a.re, a.im}; }
PTG_HD __forceinline__ cplx mul_i(cplx a) { return {-a.im, a.re}; }      // i * a
PTG_HD __forceinline__ cplx mul_mi(cplx a) { return {a.im, -a.re}; }     // -i * a

// ---------------------------------------------------------------- Stokes vectors
struct sv4 { double i, q, u, v; };

// ---------------------------------------------------------------- RNG
// Philox4x32-10 (Salmon et al., SC'11): a counter-based generator. Each particle
// owns a 64-bit stream id; the 64-bit counter advances once per block of four
// 32-bit words. The key is derived from the run seed. Streams of secondaries are
// derived from the parent's stream and event index (splitmix64 finaliser), so
// results do not depend on the order in which threads run.
struct Rng {
    uint32_t k0, k1;       // key (seed)
    uint32_t s0, s1;       // stream id
    uint32_t c0, c1;       // counter
};
__device__ __forceinline__ void philox_round(uint32_t& x0, uint32_t& x1, uint32_t& x2, uint32_t& x3, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    uint32_t hi0 = __umulhi(M0, x0), lo0 = M0 * x0;
    uint32_t hi1 = __umulhi(M1, x2), lo1 = M1 * x2;
    uint32_t y0 = hi1 ^ x1 ^ k0, y2 = hi0 ^ x3 ^ k1;
    x0 = y0; x1 = lo1; x2 = y2; x3 = lo0;
}
// Two doubles uniform on [0, 1) with 53 random bits each.
__device__ __forceinline__ void rng_uniform2(Rng& g, double& a, double& b) {
    uint32_t x0 = g.c0, x1 = g.c1, x2 = g.s0, x3 = g.s1;
    uint32_t k0 = g.k0, k1 = g.k1;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        philox_round(x0, x1, x2, x3, k0, k1);
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    if (++g.c0 == 0) ++g.c1;
    uint64_t w0 = (uint64_t(x0) << 32 | x1) >> 11, w1 = (uint64_t(x2) << 32 | x3) >> 11;
    const double SCALE = 1.0 / 9007199254740992.0;
    a = double(w0) * SCALE;
    b = double(w1) * SCALE;
}
__device__ __forceinline__ double rng_uniform(Rng& g) { double a, b; rng_uniform2(g, a, b); return a; }
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
__device__ __forceinline__ Rng rng_make(uint64_t key, uint64_t stream) {
    return Rng{uint32_t(key), uint32_t(key >> 32), uint32_t(stream), uint32_t(stream >> 32), 0u, 0u};
}
__device__ __forceinline__ uint64_t rng_child(uint64_t stream, uint64_t event) {
    return mix64(stream + 0x9e3779b97f4a7c15ULL * (event + 1));
}

// ---------------------------------------------------------------- PWMCI
// Monotone piecewise-cubic Hermite interpolation of a table of (x, f) pairs,
// with the tangent limiter of Fritsch-Carlson (tangents clipped at 3x the
// secant), and its inverse by safeguarded inverse-quadratic interpolation.
// Mirrors src/pwmci.rs:47-229: same bracketing rule, same iteration cap (32)
// and relative tolerance (1e-6), so inverses agree with the reference to that
// tolerance. `t` points at [n][2] doubles (global or LDS).
struct Hermite { double x0, x1, f0, f1, m0, m1; };

template <typename P>
PTG_HD __forceinline__ Hermite hermite_segment(P t, int i, int n) {
    double xm1 = t[2 * (i - 1)], fm1 = t[2 * (i - 1) + 1];
    double x0 = t[2 * i], f0 = t[2 * i + 1];
    double s1 = (f0 - fm1) / (x0 - xm1);
    double s0 = s1, s2 = s1;
    if (i >= 2) s0 = (fm1 - t[2 * (i - 2) + 1]) / (xm1 - t[2 * (i - 2)]);
    if (i <= n - 2) s2 = (t[2 * (i + 1) + 1] - f0) / (t[2 * (i + 1)] - x0);
    double m0 = s0 * s1 > 0.0 ? 0.5 * (s0 + s1) : 0.0;
    double m1 = s1 * s2 > 0.0 ? 0.5 * (s1 + s2) : 0.0;
    if (s1 != 0.0 && m0 / s1 > 3.0) m0 = 3.0 * s1;
    if (s2 != 0.0 && m1 / s2 > 3.0) m1 = 3.0 * s2;
    return {xm1, x0, fm1, f0, m0, m1};
}
PTG_HD __forceinline__ double hermite_eval(const Hermite& h, double x) {
    double d = h.x1 - h.x0;
    double t = (x - h.x0) / d;
    double omt = 1.0 - t;
    double t2 = t * t, omt2 = omt * omt;
    return h.f0 * (1.0 + 2.0 * t) * omt2 + h.f1 * t2 * (3.0 - 2.0 * t) + d * (h.m0 * t * omt2 + h.m1 * t2 * (t - 1.0));
}
// Segment index i (1..n-1) such that t[i-1][col] < v <= t[i][col]; returns -1 if
// v <= t[0][col] and n if v > t[n-1][col]. Binary search on a monotone column.
template <typename P>
PTG_HD __forceinline__ int pwmci_locate(P t, int n, int col, double v) {
    if (!(v > t[col])) return -1;
    if (v > t[2 * (n - 1) + col]) return n;
    int lo = 0, hi = n - 1;  // t[lo] < v <= t[hi]
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (v > t[2 * mid + col]) lo = mid; else hi = mid;
    }
    return hi;
}
// Evaluate; out-of-range -> linear extrapolation if ext, else returns false.
template <typename P>
PTG_HD __forceinline__ bool pwmci_eval(P t, int n, double x, bool ext, double& out) {
    int i = pwmci_locate(t, n, 0, x);
    if (i > 0 && i < n) { out = hermite_eval(hermite_segment(t, i, n), x); return true; }
    if (!ext) return false;
    int j = i <= 0 ? 1 : n - 1;
    double x1 = t[2 * (j - 1)], f1 = t[2 * (j - 1) + 1], x2 = t[2 * j], f2 = t[2 * j + 1];
    out = f1 + (f2 - f1) * (x - x1) / (x2 - x1);
    return true;
}
template <typename P>
PTG_HD __forceinline__ bool pwmci_invert(P t, int n, double f, bool ext, double& out) {
    int i = pwmci_locate(t, n, 1, f);
    if (!(i > 0 && i < n)) {
        if (!ext) return false;
        int j = i <= 0 ? 1 : n - 1;
        double x1 = t[2 * (j - 1)], f1 = t[2 * (j - 1) + 1], x2 = t[2 * j], f2 = t[2 * j + 1];
        out = x1 + (f - f1) * (x2 - x1) / (f2 - f1);
        return true;
    }
    Hermite h = hermite_segment(t, i, n);
    // roots: a = best estimate, b = bracket partner, c = previous best, d = one before
    double ax = h.x0, af = h.f0 - f, bx = h.x1, bf = h.f1 - f;
    if (fabs(bf) < fabs(af)) { double tx = ax, tf = af; ax = bx; af = bf; bx = tx; bf = tf; }
    double cx = bx, cf = bf, dx = 0.0;
    bool prev_bisect = false;
    for (int it = 0; it < 32; ++it) {
        double s;
        if (af != cf && bf != cf) {
            double r = af / cf, q = af / bf, tt = bf / cf;
            double p = q * (tt * (r - tt) * (cx - ax) - (1.0 - r) * (ax - bx));
            double den = (tt - 1.0) * (r - 1.0) * (q - 1.0);
            s = ax + p / den;
        } else {
            s = ax - af * (ax - bx) / (af - bf);
        }
        double lo = 0.25 * (3.0 * bx + ax);
        bool inside = lo < ax ? (s > lo && s < ax) : (ax < lo ? (s > ax && s < lo) : true);
        if (!inside || (prev_bisect && fabs(s - ax) >= 0.5 * fabs(ax - cx)) ||
            (!prev_bisect && fabs(s - ax) >= 0.5 * fabs(cx - dx))) {
            prev_bisect = true;
            s = 0.5 * (ax + bx);
        } else {
            prev_bisect = false;
        }
        dx = cx;
        cx = ax; cf = af;
        ax = s;
        af = hermite_eval(h, s) - f;
        if (bf * af >= 0.0) { bx = ax; bf = af; ax = cx; af = cf; }
        if (fabs(bf) < fabs(af)) { double tx = ax, tf = af; ax = bx; af = bf; bx = tx; bf = tf; }
        if (fabs(bx - ax) / fabs(ax) < 1.0e-6 || af == 0.0) break;
    }
    out = ax;
    return true;
}

}  // namespace ptg
