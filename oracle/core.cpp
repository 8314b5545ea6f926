// Ptarmigan CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.hpp).
// Tables, RNG, Stokes-basis algebra, PWMCI and Airy functions.
#include "oracle.hpp"

#include <cstdio>
#include <cstdlib>
#include <map>

namespace orc {

// ---------------------------------------------------------------- tables
static std::vector<double> g_blob;
static std::map<std::string, Table> g_tables;

bool load_tables(const char* path, std::string* err) {
    FILE* f = std::fopen(path, "rb");
    if (!f) { *err = std::string("cannot open ") + path; return false; }
    std::fseek(f, 0, SEEK_END);
    long len = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    std::vector<char> raw(len);
    if (std::fread(raw.data(), 1, len, f) != size_t(len)) { std::fclose(f); *err = "short read"; return false; }
    std::fclose(f);
    if (len < 16 || std::memcmp(raw.data(), "PTGTBL01", 8) != 0) { *err = "bad magic"; return false; }
    uint32_t n;
    std::memcpy(&n, raw.data() + 8, 4);
    const size_t rec = 96 + 8 + 8 + 4 + 16 + 4;
    size_t base = 16 + n * rec;
    g_blob.assign((len - base) / 8, 0.0);
    std::memcpy(g_blob.data(), raw.data() + base, g_blob.size() * 8);
    g_tables.clear();
    for (uint32_t i = 0; i < n; ++i) {
        const char* r = raw.data() + 16 + i * rec;
        char name[97] = {0};
        std::memcpy(name, r, 96);
        uint64_t off, cnt;
        std::memcpy(&off, r + 96, 8);
        std::memcpy(&cnt, r + 104, 8);
        g_tables[name] = Table{g_blob.data() + off, size_t(cnt)};
    }
    return true;
}

bool has_table(const char* name) { return g_tables.count(name) != 0; }

Table table(const char* name) {
    auto it = g_tables.find(name);
    if (it == g_tables.end()) {
        std::fprintf(stderr, "oracle: table %s not loaded\n", name);
        std::abort();
    }
    return it->second;
}

// ---------------------------------------------------------------- RNG
static uint64_t splitmix_next(uint64_t& x) {
    x += 0x9e3779b97f4a7c15ULL;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

Xoshiro::Xoshiro(uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) s[i] = splitmix_next(x);
}

uint64_t Xoshiro::next_u64() {
    uint64_t result = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
}

void Xoshiro::jump() {
    static const uint64_t J[4] = {0x180ec6d33cfd0abaULL, 0xd5a61266f0c9392cULL, 0xa9582618e03fc9aaULL, 0x39abdc4529b1661cULL};
    uint64_t t[4] = {0, 0, 0, 0};
    for (int i = 0; i < 4; ++i)
        for (int b = 0; b < 64; ++b) {
            if (J[i] & (1ULL << b))
                for (int k = 0; k < 4; ++k) t[k] ^= s[k];
            next_u64();
        }
    for (int k = 0; k < 4; ++k) s[k] = t[k];
}

double Xoshiro::normal() {
    if (has_spare) { has_spare = false; return spare; }
    for (;;) {
        double u = 2.0 * gen_f64() - 1.0, v = 2.0 * gen_f64() - 1.0;
        double s2 = u * u + v * v;
        if (s2 > 0.0 && s2 < 1.0) {
            double m = std::sqrt(-2.0 * std::log(s2) / s2);
            spare = v * m;
            has_spare = true;
            return u * m;
        }
    }
}

// ---------------------------------------------------------------- Stokes algebra
SV rotate_by(SV s, double theta) {
    double sn = std::sin(2.0 * theta), cs = std::cos(2.0 * theta);
    return {s.i, cs * s.q + sn * s.u, -sn * s.q + cs * s.u, s.v};
}

BasisSV in_basis(SV s, V3 a, V3 k) {
    double k_mag = std::hypot(k[0], k[2]);
    double a_mag = std::sqrt(norm_sqr(a));
    double cos_theta = (k_mag == 0.0) ? a[0] / a_mag : (k[2] * a[0] - k[0] * a[2]) / (k_mag * a_mag);
    double det = (k_mag == 0.0) ? -k[1] * a[2] + k[2] * a[1]
                                : a[1] * (powi(k[0], 2) + powi(k[2], 2)) - k[1] * (a[0] * k[0] + a[2] * k[2]);
    double sign = -signum(det);
    cos_theta = std::fmin(std::fmax(cos_theta, -1.0), 1.0);
    double sin_theta = sign * std::sqrt(1.0 - cos_theta * cos_theta);
    double c2 = cos_theta * cos_theta - sin_theta * sin_theta;
    double s2 = 2.0 * cos_theta * sin_theta;
    return {{s.i, c2 * s.q - s2 * s.u, s2 * s.q + c2 * s.u, s.v}, c2, s2};
}

static void lma_basis_rotation(V4 k, double* sin_t, double* cos_t) {
    double k_minus = k[0] - k[3];
    double cos_theta = (k[3] * k_minus - powi(k[1], 2)) / (k_minus * std::hypot(k[1], k[3]));
    double det = k[1] * k[2];
    double sign = signum(det);
    cos_theta = std::fmin(std::fmax(cos_theta, -1.0), 1.0);
    *sin_t = sign * std::sqrt(1.0 - cos_theta * cos_theta);
    *cos_t = cos_theta;
}

SV in_lma_basis(SV s, V4 k) {
    double st, ct;
    lma_basis_rotation(k, &st, &ct);
    double c2 = ct * ct - st * st, s2 = 2.0 * ct * st;
    return {s.i, c2 * s.q - s2 * s.u, s2 * s.q + c2 * s.u, s.v};
}

SV from_lma_basis(SV s, V4 k) {
    double st, ct;
    lma_basis_rotation(k, &st, &ct);
    st = -st;
    double c2 = ct * ct - st * st, s2 = 2.0 * ct * st;
    return {s.i, c2 * s.q - s2 * s.u, s2 * s.q + c2 * s.u, s.v};
}

// ---------------------------------------------------------------- pwmci
namespace {
enum Loc { Above, Within, Below };
struct Fit {
    double x[2], f[2], m[2];
    double eval(double xx) const {
        double t = (xx - x[0]) / (x[1] - x[0]);
        double h00 = (1.0 + 2.0 * t) * powi(1.0 - t, 2);
        double h10 = t * powi(1.0 - t, 2);
        double h01 = powi(t, 2) * (3.0 - 2.0 * t);
        double h11 = powi(t, 2) * (t - 1.0);
        return f[0] * h00 + f[1] * h01 + (x[1] - x[0]) * (m[0] * h10 + m[1] * h11);
    }
};
// pwmci.rs:241-297
Fit construct(int i, const double* t, int len) {
    auto X = [&](int j) { return t[2 * j]; };
    auto F = [&](int j) { return t[2 * j + 1]; };
    double sec[3];
    if (i == 1) {
        sec[0] = (F(i) - F(i - 1)) / (X(i) - X(i - 1));
        sec[1] = sec[0];
        sec[2] = (F(i + 1) - F(i)) / (X(i + 1) - X(i));
    } else if (i == len - 1) {
        sec[0] = (F(i - 1) - F(i - 2)) / (X(i - 1) - X(i - 2));
        sec[1] = (F(i) - F(i - 1)) / (X(i) - X(i - 1));
        sec[2] = sec[1];
    } else {
        sec[0] = (F(i - 1) - F(i - 2)) / (X(i - 1) - X(i - 2));
        sec[1] = (F(i) - F(i - 1)) / (X(i) - X(i - 1));
        sec[2] = (F(i + 1) - F(i)) / (X(i + 1) - X(i));
    }
    double tan0 = sec[0] * sec[1] > 0.0 ? 0.5 * (sec[0] + sec[1]) : 0.0;
    double tan1 = sec[1] * sec[2] > 0.0 ? 0.5 * (sec[1] + sec[2]) : 0.0;
    if (sec[1] != 0.0 && tan0 / sec[1] > 3.0) tan0 = 3.0 * sec[1];
    if (sec[2] != 0.0 && tan1 / sec[2] > 3.0) tan1 = 3.0 * sec[2];
    return Fit{{X(i - 1), X(i)}, {F(i - 1), F(i)}, {tan0, tan1}};
}
bool is_between(double x, double mn, double mx) {
    if (mn < mx) return x > mn && x < mx;
    if (mx < mn) return x > mx && x < mn;
    return true;
}
void locate(const double* t, int len, int col, double v, int* i, Loc* loc) {
    // reverse scan for the last j with v > t[j][col]
    int j = len - 1;
    for (; j >= 0; --j)
        if (v > t[2 * j + col]) break;
    if (j < 0) { *i = 1; *loc = Below; }
    else if (j == len - 1) { *i = len - 1; *loc = Above; }
    else { *i = j + 1; *loc = Within; }
}
}  // namespace

bool Interp::evaluate(double x, double* out) const {
    int i;
    Loc loc;
    locate(tbl, n, 0, x, &i, &loc);
    if (loc == Within) { *out = construct(i, tbl, n).eval(x); return true; }
    if (!ext) return false;
    double x1 = tbl[2 * (i - 1)], f1 = tbl[2 * (i - 1) + 1], x2 = tbl[2 * i], f2 = tbl[2 * i + 1];
    *out = f1 + (f2 - f1) * (x - x1) / (x2 - x1);
    return true;
}

bool Interp::invert(double f, double* out) const {
    int i;
    Loc loc;
    locate(tbl, n, 1, f, &i, &loc);
    if (loc != Within) {
        if (!ext) return false;
        double x1 = tbl[2 * (i - 1)], f1 = tbl[2 * (i - 1) + 1], x2 = tbl[2 * i], f2 = tbl[2 * i + 1];
        *out = x1 + (f - f1) * (x2 - x1) / (f2 - f1);
        return true;
    }
    Fit fp = construct(i, tbl, n);
    struct R { double x, f; } root[4] = {};
    root[0] = {tbl[2 * (i - 1)], tbl[2 * (i - 1) + 1] - f};
    root[1] = {tbl[2 * i], tbl[2 * i + 1] - f};
    if (std::fabs(root[1].f) < std::fabs(root[0].f)) std::swap(root[0], root[1]);
    root[2] = root[1];
    bool prev_bisect = false;
    for (int it = 0; it < 32; ++it) {
        double s;
        if (root[0].f != root[2].f && root[1].f != root[2].f) {
            double r = root[0].f / root[2].f;
            double ss = root[0].f / root[1].f;
            double t = root[1].f / root[2].f;
            double p = ss * (t * (r - t) * (root[2].x - root[0].x) - (1.0 - r) * (root[0].x - root[1].x));
            double q = (t - 1.0) * (r - 1.0) * (ss - 1.0);
            s = root[0].x + p / q;
        } else {
            s = root[0].x - (root[0].f * (root[0].x - root[1].x) / (root[0].f - root[1].f));
        }
        if (!is_between(s, 0.25 * (3.0 * root[1].x + root[0].x), root[0].x) ||
            (prev_bisect && std::fabs(s - root[0].x) >= 0.5 * std::fabs(root[0].x - root[2].x)) ||
            (!prev_bisect && std::fabs(s - root[0].x) >= 0.5 * std::fabs(root[2].x - root[3].x))) {
            prev_bisect = true;
            s = 0.5 * (root[0].x + root[1].x);
        } else {
            prev_bisect = false;
        }
        root[3] = root[2];
        root[2] = root[0];
        root[0].x = s;
        root[0].f = fp.eval(s) - f;
        if (root[1].f * root[0].f >= 0.0) {
            root[1] = root[0];
            root[0] = root[2];
        }
        if (std::fabs(root[1].f) < std::fabs(root[0].f)) std::swap(root[0], root[1]);
        double error = std::fabs(root[1].x - root[0].x) / std::fabs(root[0].x);
        if (error < 1.0e-6 || root[0].f == 0.0) break;
    }
    *out = root[0].x;
    return true;
}

// ---------------------------------------------------------------- Airy (airy.rs:41-136)
static double series_up_to(Table s, double x, int max) {
    // s = [a[20], n[20]] (special_functions/mod.rs:25-35)
    double sum = 0.0;
    for (int k = 0; k < 20; ++k) {
        int p = int(s[20 + k]);
        if (!(p < max)) break;
        sum += s[k] * powi(x, p);
    }
    return sum;
}

static double gl_sum(const char* nodes, const char* weights, double s, double expo) {
    Table N = table(nodes), W = table(weights);
    double acc = 0.0;
    for (size_t k = 0; k < N.n; ++k) acc += W[k] * std::pow(2.0 + N[k] / s, expo);
    return acc;
}

bool airy_ai(double x, double* out) {
    if (x < 0.0) return false;
    if (x < 1.0) { *out = series_up_to(table("special_functions/airy:AIRY_SMALL_X"), x, 20); return true; }
    double s = 2.0 * std::pow(x, 1.5) / 3.0;
    double a = 0.262183997088323 * std::pow(s, -1.0 / 6.0) * std::exp(-s);
    double integral;
    if (x < 2.0) integral = gl_sum("special_functions/airy:GL_m1_6_NODES_40", "special_functions/airy:GL_m1_6_WEIGHTS_40", s, -1.0 / 6.0);
    else if (x < 10.0) integral = gl_sum("special_functions/airy:GL_m1_6_NODES_16", "special_functions/airy:GL_m1_6_WEIGHTS_16", s, -1.0 / 6.0);
    else integral = gl_sum("special_functions/airy:GL_m1_6_NODES_4", "special_functions/airy:GL_m1_6_WEIGHTS_4", s, -1.0 / 6.0);
    *out = a * integral;
    return true;
}

bool airy_ai_prime(double x, double* out) {
    if (x < 0.0) return false;
    if (x < 1.0) { *out = series_up_to(table("special_functions/airy:AIRY_PRIME_SMALL_X"), x, 20); return true; }
    double s = 2.0 * std::pow(x, 1.5) / 3.0;
    double b = -0.2898380090915846 * std::pow(s, 1.0 / 6.0) * std::exp(-s);
    double integral;
    if (x < 2.0) integral = gl_sum("special_functions/airy:GL_p1_6_NODES_32", "special_functions/airy:GL_p1_6_WEIGHTS_32", s, 1.0 / 6.0);
    else if (x < 4.0) integral = gl_sum("special_functions/airy:GL_p1_6_NODES_16", "special_functions/airy:GL_p1_6_WEIGHTS_16", s, 1.0 / 6.0);
    else if (x < 11.0) integral = gl_sum("special_functions/airy:GL_p1_6_NODES_8", "special_functions/airy:GL_p1_6_WEIGHTS_8", s, 1.0 / 6.0);
    else integral = gl_sum("special_functions/airy:GL_p1_6_NODES_4", "special_functions/airy:GL_p1_6_WEIGHTS_4", s, 1.0 / 6.0);
    *out = b * integral;
    return true;
}

bool bessel_k_1_3(double x, double* out) {
    double y;
    if (!airy_ai(std::pow(1.5 * x, 2.0 / 3.0), &y)) return false;
    *out = PI * std::cbrt(2.0 * std::sqrt(3.0) / x) * y;
    return true;
}

bool bessel_k_2_3(double x, double* out) {
    double y;
    if (!airy_ai_prime(std::pow(1.5 * x, 2.0 / 3.0), &y)) return false;
    *out = -PI * std::pow(3.0, -1.0 / 6.0) * std::pow(2.0 / x, 2.0 / 3.0) * y;
    return true;
}

}  // namespace orc
