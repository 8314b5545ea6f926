// Ptarmigan CPU oracle -- TEST INFRASTRUCTURE ONLY.
//
// A scalar C++17 restatement of the reference's per-particle Monte Carlo loop
// (tgblackburn/ptarmigan v1.5.0, Rust). It exists to check the HIP product path:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// it. Every function cites the reference file:line it follows. Arithmetic is
// written in the reference's operation order and compiled with
// -ffp-contract=off, because Rust does not contract a*b+c into FMA.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

namespace orc {

// ---------------------------------------------------------------- constants
// src/constants.rs:4-29 (7-significant-digit SI values, used verbatim)
constexpr double SPEED_OF_LIGHT = 2.997925e8;
constexpr double SPEED_OF_LIGHT_SQD = 89875517873681764.0;
constexpr double VACUUM_PERMITTIVITY = 8.854188e-12;
constexpr double ELECTRON_CHARGE = -1.602177e-19;
constexpr double ELEMENTARY_CHARGE = 1.602177e-19;
constexpr double ELECTRON_MASS = 9.109383e-31;
constexpr double ELECTRON_MASS_MEV = 0.510999;
constexpr double CRITICAL_FIELD = 1.323285e18;
constexpr double ALPHA_FINE = 7.29735257e-3;
constexpr double COMPTON_TIME = 1.28808867e-21;
constexpr double PI = 3.14159265358979323846;
constexpr double FRAC_PI_2 = 1.57079632679489661923;
constexpr double FRAC_PI_3 = 1.04719755119659774615;
constexpr double LN_2 = 0.693147180559945309417;
constexpr double SQRT_2 = 1.41421356237309504880;

// Rust's f64::powi with a constant exponent lowers to square-and-multiply
// (llvm.powi / compiler-rt __powidf2); restated so the rounding matches.
inline double powi(double a, int b) {
    bool recip = b < 0;
    unsigned n = recip ? unsigned(-b) : unsigned(b);
    double r = 1.0;
    for (;;) {
        if (n & 1) r *= a;
        n >>= 1;
        if (!n) break;
        a *= a;
    }
    return recip ? 1.0 / r : r;
}
inline double signum(double x) { return std::signbit(x) ? -1.0 : 1.0; }  // Rust f64::signum (x != NaN)
inline double fract(double x) { return x - std::trunc(x); }             // Rust f64::fract

// ---------------------------------------------------------------- vectors
// src/geometry/three_vector.rs:25-120, 153-205
struct V3 {
    double x, y, z;
    double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline double operator*(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator/(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline V3 operator-(V3 a) { return -1.0 * a; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double norm_sqr(V3 a) { return a * a; }
inline V3 normalize(V3 a) { double m = std::sqrt(norm_sqr(a)); return a / m; }
// three_vector.rs:79-95
inline V3 rotate_around(V3 v, V3 ax, double th) {
    double s = std::sin(th), c = std::cos(th);
    return {(c + ax.x * ax.x * (1.0 - c)) * v.x + (ax.x * ax.y * (1.0 - c) - ax.z * s) * v.y + (ax.x * ax.z * (1.0 - c) + ax.y * s) * v.z,
            (ax.y * ax.x * (1.0 - c) + ax.z * s) * v.x + (c + ax.y * ax.y * (1.0 - c)) * v.y + (ax.y * ax.z * (1.0 - c) - ax.x * s) * v.z,
            (ax.z * ax.x * (1.0 - c) - ax.y * s) * v.x + (ax.z * ax.y * (1.0 - c) + ax.x * s) * v.y + (c + ax.z * ax.z * (1.0 - c)) * v.z};
}
// three_vector.rs:109-117
inline V3 rotate_around_z(V3 v, double th) {
    double s = std::sin(th), c = std::cos(th);
    return {c * v.x - s * v.y, c * v.y + s * v.x, v.z};
}
// three_vector.rs:103-106 -> rotate_around(y axis)
inline V3 rotate_around_y(V3 v, double th) { return rotate_around(v, V3{0.0, 1.0, 0.0}, th); }

// src/geometry/four_vector.rs:22-110, 128-205
struct V4 {
    double t, x, y, z;
    double operator[](int i) const { return i == 0 ? t : (i == 1 ? x : (i == 2 ? y : z)); }
};
inline V4 operator+(V4 a, V4 b) { return {a.t + b.t, a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V4 operator-(V4 a, V4 b) { return {a.t - b.t, a.x - b.x, a.y - b.y, a.z - b.z}; }
inline double operator*(V4 a, V4 b) { return a.t * b.t - a.x * b.x - a.y * b.y - a.z * b.z; }
inline V4 operator*(double s, V4 a) { return {s * a.t, s * a.x, s * a.y, s * a.z}; }
inline V4 operator*(V4 a, double s) { return {a.t * s, a.x * s, a.y * s, a.z * s}; }
inline V4 operator/(V4 a, double s) { return {a.t / s, a.x / s, a.y / s, a.z / s}; }
inline V3 spatial(V4 a) { return {a.x, a.y, a.z}; }
inline V4 lightlike(double x, double y, double z) { return {std::sqrt(powi(x, 2) + powi(y, 2) + powi(z, 2)), x, y, z}; }
inline V4 unitize(V4 a) { return {std::sqrt(1.0 + powi(a.x, 2) + powi(a.y, 2) + powi(a.z, 2)), a.x, a.y, a.z}; }
inline V4 with_sqr(V4 a, double b) { return {std::sqrt(b + powi(a.x, 2) + powi(a.y, 2) + powi(a.z, 2)), a.x, a.y, a.z}; }
// four_vector.rs:68-80
inline V4 boost_by(V4 s, V4 u) {
    double gamma = u.t;
    double beta = std::sqrt(1.0 - 1.0 / (gamma * gamma));
    V3 n = normalize(spatial(u));
    double a = s.t;
    V3 z = spatial(s);
    return {gamma * (a - beta * (n * z)),
            z.x + (gamma - 1.0) * (n * z) * n.x - gamma * beta * a * n.x,
            z.y + (gamma - 1.0) * (n * z) * n.y - gamma * beta * a * n.y,
            z.z + (gamma - 1.0) * (n * z) * n.z - gamma * beta * a * n.z};
}
// four_vector.rs:98-107
inline V4 rotate_around_z(V4 v, double th) {
    double s = std::sin(th), c = std::cos(th);
    return {v.t, c * v.x - s * v.y, c * v.y + s * v.x, v.z};
}

// src/geometry/polarization.rs:11-214
struct SV {
    double i, q, u, v;
    double operator[](int k) const { return k == 0 ? i : (k == 1 ? q : (k == 2 ? u : v)); }
};
inline SV unpolarized() { return {1.0, 0.0, 0.0, 0.0}; }
struct BasisSV { SV sv; double cos_2theta, sin_2theta; };
BasisSV in_basis(SV s, V3 a, V3 k);             // polarization.rs:58-101
SV rotate_by(SV s, double theta);               // polarization.rs:44-53
SV in_lma_basis(SV s, V4 k);                    // polarization.rs:120-133
SV from_lma_basis(SV s, V4 k);                  // polarization.rs:141-157

// ---------------------------------------------------------------- complex (num-complex 0.4.0 semantics)
struct C {
    double re, im;
};
inline C operator+(C a, C b) { return {a.re + b.re, a.im + b.im}; }
inline C operator-(C a, C b) { return {a.re - b.re, a.im - b.im}; }
inline C operator*(C a, C b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline C operator/(C a, C b) {
    double ns = b.re * b.re + b.im * b.im;
    double re = a.re * b.re + a.im * b.im;
    double im = a.im * b.re - a.re * b.im;
    return {re / ns, im / ns};
}
inline C operator*(C a, double s) { return {a.re * s, a.im * s}; }
inline C operator*(double s, C a) { return {s * a.re, s * a.im}; }
inline C operator/(C a, double s) { return {a.re / s, a.im / s}; }
inline C operator+(C a, double s) { return {a.re + s, a.im}; }
inline C operator+(double s, C a) { return {s + a.re, a.im}; }
inline C operator-(C a, double s) { return {a.re - s, a.im}; }
inline C operator-(double s, C a) { return {s - a.re, 0.0 - a.im}; }  // as Complex - Complex: +0 imag for real a
inline C operator-(C a) { return {-a.re, -a.im}; }
inline C cexp(C a) { double e = std::exp(a.re); return {e * std::cos(a.im), e * std::sin(a.im)}; }
inline C cpowf(C a, double p) {
    double r = std::hypot(a.re, a.im), th = std::atan2(a.im, a.re);
    double rp = std::pow(r, p), tp = th * p;
    return {rp * std::cos(tp), rp * std::sin(tp)};
}

// ---------------------------------------------------------------- tables
struct Table {
    const double* p = nullptr;
    size_t n = 0;
    double operator[](size_t i) const { return p[i]; }
};
bool load_tables(const char* path, std::string* err);
Table table(const char* name);  // aborts if absent (oracle misuse is a bug)
bool has_table(const char* name);

// ---------------------------------------------------------------- RNG
// rand_xoshiro 0.4.0 Xoshiro256StarStar; seed_from_u64 via SplitMix64; jump();
// rand 0.7 Standard f64 = (next_u64 >> 11) * 2^-53.
struct Xoshiro {
    uint64_t s[4];
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    explicit Xoshiro(uint64_t seed);
    uint64_t next_u64();
    double gen_f64() { return double(next_u64() >> 11) * (1.0 / 9007199254740992.0); }
    double normal();  // standard normal (Marsaglia polar; statistically equivalent to rand_distr Ziggurat)
    void jump();
private:
    bool has_spare = false;
    double spare = 0.0;
};

// ---------------------------------------------------------------- pwmci (src/pwmci.rs)
struct Interp {
    const double* tbl;  // [n][2]
    int n;
    bool ext = false;
    bool evaluate(double x, double* out) const;  // pwmci.rs:47-85
    bool invert(double f, double* out) const;    // pwmci.rs:102-229
};

// ---------------------------------------------------------------- special functions
bool airy_ai(double x, double* out);        // airy.rs:41-75
bool airy_ai_prime(double x, double* out);  // airy.rs:77-116
bool bessel_k_1_3(double x, double* out);   // airy.rs:118-127
bool bessel_k_2_3(double x, double* out);   // airy.rs:129-136

// ---------------------------------------------------------------- LCFA
double lcfa_rate(double chi, double gamma);                                                 // lcfa/photon_emission/mod.rs:13-33
double lcfa_classical_rate(double chi, double gamma);                                       // lcfa/photon_emission/classical.rs:10-13
void lcfa_sample(double chi, double gamma, double r1, double r2, double r3, double* omega, bool* theta_ok, double* theta, double* cphi);  // mod.rs:80-231
void lcfa_classical_sample(double chi, double gamma, double r1, double r2, double r3, double* omega, bool* theta_ok, double* theta, double* cphi);  // classical.rs:24-60
SV lcfa_stokes(V4 k, double chi, double gamma, V3 v, V3 w);                                 // mod.rs:239-298
SV lcfa_classical_stokes(V4 k, double chi, double gamma, V3 v, V3 w);                       // classical.rs:67-110
double gaunt_factor(double chi);                                                            // mod.rs:330-346
void auxiliary_t(double chi, double* par, double* perp);                                    // lcfa/pair_creation/mod.rs:13-40
double pair_probability(V4 ell, SV sv, double chi, V3 a_perp, double dt, SV* sv_new);      // lcfa/pair_creation/mod.rs:48-73
void pair_sample(V4 ell, SV sv, double chi, V3 a_perp, Xoshiro& rng, double* gamma_p, double* cos_theta, double* phi, double* s, double* z);  // :180-217
double pair_spectrum(double s, double chi, double sv1);                                     // :80-89
double pair_spectrum_ceiling(double chi, double sv1);                                       // :93-110
double pair_angular_spectrum(double y, double s, double chi, double sv1);                   // :116-127
double pair_angular_spectrum_ceiling(double s, double chi, double sv1);                     // :131-143

// ---------------------------------------------------------------- Bessel J, CP nonlinear Compton (cp.cpp)
void bessel_j_pm(int n, double x, double out[3]);  // special_functions/bessel.rs:26-33
double bessel_j(int n, double x);                  // bessel.rs:18-24
bool cp_rate(double a, double eta, double* out);   // nonlinear_compton/cp/mod.rs:162-176
bool cp_classical_rate(double a, double eta, double* out);  // cp/classical.rs:22-44
void cp_sample(double a, double eta, Xoshiro& rng, int fixed_n, int* n, double* s, double* theta, SV* sv, uint64_t* trials);  // cp/mod.rs:181-234
void cp_classical_sample(double a, double eta, Xoshiro& rng, int fixed_n, int* n, double* s, double* theta, SV* sv,
                         uint64_t* trials);  // cp/classical.rs:128-178

double cpbw_probability(double a, double eta, SV sv, double dphi, SV* out);  // pair_creation/cp/mod.rs:286-322
void cpbw_sample(double a, double eta, double sv1, double sv2, double sv3, Xoshiro& rng, int* n, double* s, double* theta);  // :327-365

// ---------------------------------------------------------------- fields
enum class Kind : int { PlaneWave = 0, FocusedLaser = 1, FastPlaneWave = 2, FastFocusedLaser = 3 };
enum class Env : int { CosSquared = 0, Flattop = 1, Gaussian = 2 };
enum class Pol : int { Linear = 0, Circular = 1 };
enum class Eqn : int { Lorentz = 0, LandauLifshitz = 1, ModifiedLandauLifshitz = 2 };

struct Laser {
    Kind kind;
    double a0, wavelength, waist, n_cycles;
    Env env;
    Pol pol;
    double pol_angle, chirp_b;
    bool finite_bandwidth;
    // derived
    V4 wavevector;
    double duration;
    void init();
    double max_timestep() const;
    bool contains(V4 r) const;
    double ideal_initial_z() const;
    void fields(V4 r, V3* E, V3* B, double* a) const;
};

struct PushOut { V4 r, u; double dt, work; };
PushOut vay_push(V4 r, V4 ui, V3 E, V3 B, double rqm, double dt, Eqn eqn);  // field/lcf.rs:18-78
PushOut field_push(const Laser& L, V4 r, V4 u, double rqm, double dt, Eqn eqn);  // field/mod.rs:147-152 (and LMA overrides)

}  // namespace orc
