// Ptarmigan CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.hpp).
// LCFA photon emission (src/lcfa/photon_emission/{mod,classical}.rs) and
// LCFA pair creation (src/lcfa/pair_creation/mod.rs).
#include "oracle.hpp"

namespace orc {

static const char* PE = "lcfa/photon_emission/tables:";
static Table pe(const char* n) { return table((std::string(PE) + n).c_str()); }

// lcfa/photon_emission/mod.rs:13-33
double lcfa_rate(double chi, double gamma) {
    double h;
    if (chi < 0.01) {
        h = 5.0 * FRAC_PI_3 * (1.0 - 8.0 * chi / (5.0 * std::sqrt(3.0)));
    } else if (chi >= 100.0) {
        double c = std::cbrt(chi);
        h = -1019.4661473121777 + 1786.716527650374 * c * c;
        h = 1750.6263395722715 + c * c * h;
        h = -2260.1819695887225 + c * h;
        h = 0.00296527643253334 * h / (chi * chi);
    } else {
        Table T = pe("LN_H_CHI_TABLE");
        double delta = pe("DELTA_LN_CHI")[0];
        double index = (std::log(chi) - T[0]) / delta;
        double weight = fract(index);
        size_t i = size_t(std::floor(index));
        double ln_h = (1.0 - weight) * T[2 * i + 1] + weight * T[2 * (i + 1) + 1];
        h = std::exp(ln_h);
    }
    return std::sqrt(3.0) * ALPHA_FINE * chi * h / (2.0 * PI * gamma * COMPTON_TIME);
}

// lcfa/photon_emission/classical.rs:10-13
double lcfa_classical_rate(double chi, double gamma) {
    double h = 5.0 * FRAC_PI_3;
    return std::sqrt(3.0) * ALPHA_FINE * chi * h / (2.0 * PI * gamma * COMPTON_TIME);
}

// lcfa/photon_emission/mod.rs:35-69. cdf = [coeff, power, table[31][2]]
static double from_linear_cdf_table(double global_zero, double local_zero, double rand, const double* cdf) {
    double coeff = cdf[0], power = cdf[1];
    const double* tbl = cdf + 2;
    Interp ip{tbl, 31};
    double r_zero;
    if (local_zero < tbl[0]) {
        r_zero = coeff * std::pow(local_zero - global_zero, power);
    } else {
        double tmp;
        if (!ip.evaluate(local_zero, &tmp)) return local_zero;
        r_zero = tmp;
    }
    double r = r_zero + (1.0 - r_zero) * rand;
    double y;
    if (r <= tbl[1]) {
        double tmp = (std::log(r) - std::log(coeff)) / power;
        y = std::exp(tmp) + global_zero;
    } else if (!ip.invert(r, &y)) {
        y = local_zero;
    }
    return y;
}

// lcfa/photon_emission/mod.rs:80-231
void lcfa_sample(double chi, double gamma, double rand1, double rand2, double rand3,
                 double* omega, bool* theta_ok, double* theta_out, double* cphi) {
    const double LN_CHI_MIN = pe("LN_CHI_MIN")[0], LN_CHI_STEP = pe("LN_CHI_STEP")[0];
    const double LN_DELTA_MIN = pe("LN_DELTA_MIN")[0], LN_DELTA_STEP = pe("LN_DELTA_STEP")[0];
    Table Q = pe("QUANTUM_CDF"), YC = pe("Y_CDF"), YI = pe("Y_INFINITE_DELTA_CDF");
    Table HC = pe("HIGH_CHI_CDF"), HP = pe("HIGH_CHI_PEAK_CDF");

    double index = (std::log(chi) - LN_CHI_MIN) / LN_CHI_STEP;
    double weight = fract(index);
    size_t idx = size_t(std::floor(index));

    if (std::log(chi) <= LN_CHI_MIN) {
        double om;
        lcfa_classical_sample(chi, gamma, rand1, rand2, rand3, &om, theta_ok, theta_out, cphi);
        *omega = om * gamma / (gamma + om);
        return;
    }

    double u;
    if (idx >= 41 - 1) {
        double u_bdy = (chi - 8.0) / 8.0;
        Interp lower{HC.p, 32};
        double p_lower;
        if (std::log(u_bdy) < HC[0]) p_lower = 1.06327715546 * std::cbrt(u_bdy);
        else if (std::log(u_bdy) < HC[2 * 31]) {
            double y;
            lower.evaluate(std::log(u_bdy), &y);
            p_lower = 1.0 - std::exp(-1.0 * std::exp(y));
        } else p_lower = 1.0 - 0.265819288864 * std::pow(u_bdy, -2.0 / 3.0);

        double y_bdy = u_bdy / chi - 1.0 / 9.0;
        double scale = 0.84626000931 * std::pow(chi, -2.0 / 3.0);
        Interp higher{HP.p, 32, true};
        double p_higher;
        if (std::log(y_bdy) < HP[0]) p_higher = scale * (1.0 - 7.7567692587 * y_bdy);
        else if (std::log(y_bdy) < HP[2 * 31]) {
            double t;
            higher.evaluate(std::log(y_bdy), &t);
            p_higher = scale * std::exp(-1.0 * std::exp(t));
        } else p_higher = scale * 0.46534912384 * std::exp(-0.45 * u_bdy / chi) * std::pow(y_bdy, -5.0 / 3.0);

        double r = rand1 * (p_lower + p_higher);
        if (r <= p_lower) {
            double target = std::log(-std::log(1.0 - r));
            if (target < HC[1]) {
                double uu = powi(r / 1.06327715546, 3);
                u = uu + 1.5 * uu * uu;
            } else if (target < HC[2 * 31 + 1]) {
                double x;
                lower.invert(target, &x);
                u = std::exp(x);
            } else {
                u = std::pow(0.265819288864 / (1.0 - r), 1.5);
            }
        } else {
            double target = (p_lower + p_higher - r) / scale;
            double y;
            if (target < HP[1]) y = (1.0 - target) / 7.7567692587;
            else {
                double tg = std::log(-1.0 * std::log(target));
                double x;
                higher.invert(tg, &x);
                y = std::exp(x);
            }
            u = (y + 1.0 / 9.0) * chi;
        }
    } else {
        auto ln_u_of = [&](const double* cdf) {
            double coeff = cdf[0], power = cdf[1];
            const double* tbl = cdf + 2;
            double lr = std::log(rand1);
            if (lr <= tbl[1]) return (lr - std::log(coeff)) / power;
            double res;
            if (Interp{tbl, 31}.invert(lr, &res)) return res;
            return tbl[2 * 30];
        };
        double ln_u_lower = ln_u_of(Q.p + 64 * idx);
        double ln_u_upper = ln_u_of(Q.p + 64 * (idx + 1));
        u = std::exp((1.0 - weight) * ln_u_lower + weight * ln_u_upper);
    }

    // angle
    double beta = 2.0 * u / (3.0 * chi);
    double delta = (1.0 + powi(1.0 + u, 2)) * std::pow(beta, -2.0 / 3.0) / (1.0 + u);
    double dindex = (std::log(delta) - LN_DELTA_MIN) / LN_DELTA_STEP;
    double dweight = fract(dindex);
    size_t di = size_t(std::floor(dindex));
    double y;
    if (di >= 31 - 1) {
        y = from_linear_cdf_table(0.0, beta, rand2, YI.p);
    } else {
        double y_lower = from_linear_cdf_table(std::pow(delta, -1.5), beta, rand2, YC.p + 64 * di);
        double y_upper = from_linear_cdf_table(std::pow(delta, -1.5), beta, rand2, YC.p + 64 * (di + 1));
        y = (1.0 - dweight) * y_lower + dweight * y_upper;
    }
    double z = std::fmax(y / beta, 1.0);
    double cos_theta = (gamma - std::pow(z, 2.0 / 3.0) / (2.0 * gamma)) / std::sqrt(powi(gamma, 2) - 1.0);
    if (cos_theta >= 1.0) { *theta_ok = true; *theta_out = 0.0; }
    else if (cos_theta >= -1.0) { *theta_ok = true; *theta_out = std::acos(cos_theta); }
    else { *theta_ok = false; *theta_out = 0.0; }
    *omega = gamma * u / (1.0 + u);
    *cphi = 2.0 * PI * rand3;
}

// lcfa/photon_emission/classical.rs:24-60
void lcfa_classical_sample(double chi, double gamma, double rand1, double rand2, double rand3,
                           double* omega, bool* theta_ok, double* theta_out, double* cphi) {
    Table CS = pe("CLASSICAL_SPECTRUM_TABLE");
    double delta = std::acos((-9.0 + 50.0 * rand2 - 25.0 * powi(rand2, 2)) / 16.0);
    double z = powi((2.0 + 4.0 * std::cos(delta / 3.0)) / (5.0 * (1.0 - rand2)), 3);
    double ln_rand = std::log(rand1);
    double x;
    if (ln_rand < CS[1]) x = 1.020377255 * std::pow(rand1, 0.6);
    else {
        double lx;
        if (!Interp{CS.p, 41}.invert(ln_rand, &lx)) lx = CS[2 * 40];
        x = std::exp(lx);
    }
    double u = 3.0 * chi * x / (2.0 * z);
    *omega = u * gamma;
    double cos_theta = (gamma - std::pow(z, 2.0 / 3.0) / (2.0 * gamma)) / std::sqrt(powi(gamma, 2) - 1.0);
    if (cos_theta >= 1.0) { *theta_ok = true; *theta_out = 0.0; }
    else if (cos_theta >= -1.0) { *theta_ok = true; *theta_out = std::acos(cos_theta); }
    else { *theta_ok = false; *theta_out = 0.0; }
    *cphi = 2.0 * PI * rand3;
}

static void k13_k23(double eta, double* k1, double* k2) {
    if (!bessel_k_1_3(eta, k1)) *k1 = 0.0;
    if (!bessel_k_2_3(eta, k2)) *k2 = 0.0;
}

// lcfa/photon_emission/mod.rs:239-298
SV lcfa_stokes(V4 k, double chi, double gamma, V3 v, V3 w) {
    v = normalize(v);
    w = normalize(w);
    V3 n = normalize(spatial(k));
    double u = k[0] / (gamma - k[0]);
    double beta = std::asin(n * normalize(cross(v, w)));
    double mu = std::sqrt(beta * beta + 1.0 / (gamma * gamma));
    double eta = u * powi(gamma * mu, 3) / (3.0 * chi);
    double k1_3, k2_3;
    k13_k23(eta, &k1_3, &k2_3);
    double g = 0.5 * powi(u * mu, 2) * (k1_3 * k1_3 + k2_3 * k2_3) / (1.0 + u) + powi(mu * k2_3, 2) + powi(beta * k1_3, 2);
    V3 xi{(powi(mu * k2_3, 2) - powi(beta * k1_3, 2)) / g, 0.0,
          2.0 * (1.0 + 0.5 * u * u / (1.0 + u)) * beta * mu * k1_3 * k2_3 / g};
    V3 e1{k[3], 0.0, -k[1]};
    if (norm_sqr(e1) == 0.0) e1 = V3{1.0, 0.0, 0.0};
    else e1 = normalize(e1);
    double phi = std::acos(std::fmin(std::fmax(e1 * w, -1.0), 1.0));
    double sign = signum(cross(w, e1) * n);
    phi = sign * phi;
    V3 x2{std::cos(2.0 * phi) * xi.x + std::sin(2.0 * phi) * xi.y,
          -std::sin(2.0 * phi) * xi.x + std::cos(2.0 * phi) * xi.y, xi.z};
    return {1.0, x2.x, x2.y, x2.z};
}

// lcfa/photon_emission/classical.rs:67-110
SV lcfa_classical_stokes(V4 k, double chi, double gamma, V3 v, V3 w) {
    v = normalize(v);
    w = normalize(w);
    V3 n = normalize(spatial(k));
    double u = k[0] / gamma;
    double beta = std::asin(n * normalize(cross(v, w)));
    double mu = std::sqrt(beta * beta + 1.0 / (gamma * gamma));
    double eta = u * powi(gamma * mu, 3) / (3.0 * chi);
    double k1_3, k2_3;
    k13_k23(eta, &k1_3, &k2_3);
    double g = powi(beta * k1_3, 2) + powi(mu * k2_3, 2);
    V3 xi{(powi(mu * k2_3, 2) - powi(beta * k1_3, 2)) / g, 0.0, 2.0 * beta * mu * k1_3 * k2_3 / g};
    double phi = std::asin((w - (n * w) * n) * V3{0.0, 1.0, 0.0});
    V3 x2{std::cos(2.0 * phi) * xi.x + std::sin(2.0 * phi) * xi.y,
          -std::sin(2.0 * phi) * xi.x + std::cos(2.0 * phi) * xi.y, xi.z};
    return {1.0, x2.x, x2.y, x2.z};
}

// lcfa/photon_emission/mod.rs:330-346
double gaunt_factor(double chi) {
    if (chi <= 1.0e-3) return 1.0 - 55.0 * std::sqrt(3.0) * chi / 16.0 + 48.0 * chi * chi;
    if (chi > 500.0) {
        double coeff = 0.5563810015417746;
        double c = std::cbrt(chi);
        double g = -1.0000341464898563 + coeff * c * c;
        g = 2.6235881916750565 + c * c * g;
        g = -3.8106418922532588 + c * g;
        return g / powi(chi, 3);
    }
    double y = 0.0;
    Interp{table("lcfa/photon_emission/mod:GAUNT_FACTOR_TABLE").p, 25}.evaluate(std::log(chi), &y);
    return std::exp(y);
}

// ---------------------------------------------------------------- pair creation
// lcfa/pair_creation/mod.rs:13-40
void auxiliary_t(double chi, double* par, double* perp) {
    Table T = table("lcfa/pair_creation/tables:LN_T_CHI_TABLE");
    double D = table("lcfa/pair_creation/tables:DELTA_LN_CHI")[0];
    if (chi <= 0.01) { *par = 0.0; *perp = 0.0; }
    else if (chi < 1.0) {
        size_t i = size_t((std::log(chi) - T[0]) / D);
        double dx = (chi - std::exp(T[3 * i])) / (std::exp(T[3 * (i + 1)]) - std::exp(T[3 * i]));
        double pa = (1.0 - dx) / T[3 * i + 1] + dx / T[3 * (i + 1) + 1];
        double pe_ = (1.0 - dx) / T[3 * i + 2] + dx / T[3 * (i + 1) + 2];
        *par = std::exp(1.0 / pa);
        *perp = std::exp(1.0 / pe_);
    } else if (chi < 100.0) {
        size_t i = size_t((std::log(chi) - T[0]) / D);
        double dx = (std::log(chi) - T[3 * i]) / D;
        double pa = (1.0 - dx) * T[3 * i + 1] + dx * T[3 * (i + 1) + 1];
        double pe_ = (1.0 - dx) * T[3 * i + 2] + dx * T[3 * (i + 1) + 2];
        *par = std::exp(pa);
        *perp = std::exp(pe_);
    } else {
        *par = 0.3036898468348568 / std::cbrt(chi) - 2.0 / (3.0 * chi);
        *perp = 0.4555347702522852 / std::cbrt(chi) - 2.0 / (3.0 * chi);
    }
}

// lcfa/pair_creation/mod.rs:48-73
double pair_probability(V4 ell, SV sv0, double chi, V3 a_perp, double dt, SV* sv_new) {
    BasisSV b = in_basis(sv0, a_perp, spatial(ell));
    SV sv = b.sv;
    double parallel_proj = 0.5 * (1.0 + sv[1]);
    double perp_proj = 0.5 * (1.0 - sv[1]);
    double t_par, t_perp;
    auxiliary_t(chi, &t_par, &t_perp);
    double prob = ALPHA_FINE * chi * (parallel_proj * t_par + perp_proj * t_perp) * dt / (COMPTON_TIME * ell[0]);
    double prob_avg = ALPHA_FINE * chi * 0.5 * (t_par + t_perp) * dt / (COMPTON_TIME * ell[0]);
    double delta = ALPHA_FINE * chi * 0.5 * (t_par - t_perp) * dt / (COMPTON_TIME * ell[0]);
    double sv1 = (sv[1] * (1.0 - prob_avg) - delta) / (1.0 - prob);
    double sv2 = sv[2] * (1.0 - prob_avg) / (1.0 - prob);
    *sv_new = {sv[0], b.cos_2theta * sv1 + b.sin_2theta * sv2, -b.sin_2theta * sv1 + b.cos_2theta * sv2,
               sv[3] * (1.0 - prob_avg) / (1.0 - prob)};
    return prob;
}

// lcfa/pair_creation/mod.rs:80-89
double pair_spectrum(double s, double chi, double sv1) {
    Table N = table("quadrature/tables:GL_NODES"), W = table("quadrature/tables:GL_WEIGHTS");
    double acc = 0.0;
    for (int k = 0; k < 32; ++k) {
        double t = N[k], w = W[k];
        double xi = 2.0 / (3.0 * chi * s * (1.0 - s));
        double prefactor = std::exp(-xi * std::cosh(t) + t);
        acc += w * prefactor * ((1.0 / (s * (1.0 - s)) - sv1 - 2.0) * std::cosh(2.0 * t / 3.0) + std::cosh(t / 3.0) / std::cosh(t));
    }
    return acc;
}

// lcfa/pair_creation/mod.rs:93-110
double pair_spectrum_ceiling(double chi, double sv1) {
    double chi_switch = std::exp((1.5 - sv1) / std::sqrt(3.0));
    double mx;
    if (chi < chi_switch) mx = pair_spectrum(0.5, chi, sv1);
    else if (chi > 100.0) mx = pair_spectrum(4.0 / (3.0 * chi), chi, sv1);
    else {
        double m = -0.94866 + 0.170159 * sv1;
        double s = 0.5 * std::pow(chi_switch, -m) * std::pow(chi, m);
        mx = pair_spectrum(s, chi, sv1);
    }
    return 1.05 * mx;
}

// lcfa/pair_creation/mod.rs:116-127
double pair_angular_spectrum(double y, double s, double chi, double sv1) {
    double xi = 2.0 / (3.0 * chi * s * (1.0 - s));
    double prefactor = s / (1.0 - s) + (1.0 - s) / s - sv1;
    double z = 1.0 + 4.0 * chi * y * y;
    double k;
    if (!bessel_k_1_3(xi * z, &k)) k = 0.0;
    return y * (1.0 + prefactor * std::pow(z, 2.0 / 3.0)) * k;
}

// lcfa/pair_creation/mod.rs:131-143
double pair_angular_spectrum_ceiling(double s, double chi, double sv1) {
    double y_min = 0.216, y_max = 0.259;
    double y_peak = y_min + (y_max - y_min) * (1.0 - std::tanh(std::pow(8.3 / chi, 2.0 / 3.0)));
    double y = y_peak * std::sqrt(1.0 - powi(1.0 - 2.0 * s, 2));
    return 1.05 * pair_angular_spectrum(y, s, chi, sv1);
}

// lcfa/pair_creation/mod.rs:145-173
static double sample_azimuthal_angle(double s, double z, double chi, SV sv, Xoshiro& rng) {
    double arg = 2.0 * z / (3.0 * chi * s * (1.0 - s));
    double k_ratio = arg < 1.0e-4 ? 0.6368498843179743 / std::cbrt(arg)
                                  : 1.0 + 1.0 / (1.4624087952220928 * std::cbrt(arg) + 1.023821552056939 * std::sqrt(arg) + 6.0 * arg);
    double a = 1.0 + std::pow(z, 2.0 / 3.0) * (powi(s, 2) + powi(1.0 - s, 2)) / (s * (1.0 - s));
    double b = 1.0;
    double c = std::pow(z, 2.0 / 3.0);
    double d = std::pow(z, 2.0 / 3.0) - 1.0;
    double e = std::pow(z, 1.0 / 3.0) * (std::pow(z, 2.0 / 3.0) - 1.0) * (powi(s, 2) + powi(1.0 - s, 2)) * k_ratio / (s * (1.0 - s));
    auto spec = [&](double phi) {
        return a + b * (std::cos(2.0 * phi) * (1.0 - c) - c) * sv[1] - d * std::sin(2.0 * phi) * sv[2] + e * std::sin(phi) * sv[3];
    };
    double mx = spec(0.0);
    for (int i = 1; i < 32; ++i) mx = std::fmax(mx, spec(2.0 * PI * double(i) / 32.0));
    mx = 1.1 * mx;
    for (;;) {
        double phi = 2.0 * PI * rng.gen_f64();
        double u = rng.gen_f64();
        if (u <= spec(phi) / mx) return phi;
    }
}

// lcfa/pair_creation/mod.rs:180-217
void pair_sample(V4 ell, SV sv0, double chi, V3 a_perp, Xoshiro& rng,
                 double* gamma_p, double* cos_theta, double* phi, double* s_out, double* z_out) {
    double gamma = ell[0];
    SV sv = in_basis(sv0, a_perp, spatial(ell)).sv;
    double mx = pair_spectrum_ceiling(chi, sv[1]);
    double s;
    for (;;) {
        s = rng.gen_f64();
        double u = rng.gen_f64();
        if (u <= pair_spectrum(s, chi, sv[1]) / mx) break;
    }
    double mx2 = pair_angular_spectrum_ceiling(s, chi, sv[1]);
    double z;
    for (;;) {
        double y = rng.gen_f64();
        double u = rng.gen_f64();
        if (u <= pair_angular_spectrum(y, s, chi, sv[1]) / mx2) { z = 1.0 + 4.0 * chi * y * y; break; }
    }
    *phi = sample_azimuthal_angle(s, z, chi, sv, rng);
    double gp = s * gamma;
    double ct = (2.0 * gp * gp - std::pow(z, 2.0 / 3.0)) / (2.0 * gp * gp - 1.0);
    *cos_theta = std::fmax(ct, -1.0);
    *gamma_p = gp;
    *s_out = s;
    *z_out = z;
}

}  // namespace orc
