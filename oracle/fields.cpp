// Ptarmigan CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.hpp).
// Laser fields with the fast carrier (src/field/fast_focused_laser.rs,
// src/field/fast_plane_wave.rs) and the Vay pusher (src/field/lcf.rs:18-78).
#include "oracle.hpp"

namespace orc {

void Laser::init() {
    // fast_focused_laser.rs:24-38, fast_plane_wave.rs:23-38, plane_wave.rs, focused_laser.rs
    wavevector = (2.0 * PI / wavelength) * V4{1.0, 0.0, 0.0, 1.0};
    duration = n_cycles * wavelength / SPEED_OF_LIGHT;
    if (pol == Pol::Circular) pol_angle = 0.0;
}

double lma_max_timestep(const Laser& L);
bool lma_contains(const Laser& L, V4 r);
double lma_ideal_initial_z(const Laser& L);

static double ffl_omega(const Laser& L) { return SPEED_OF_LIGHT * L.wavevector[0]; }
static double ffl_rayleigh(const Laser& L) { return 0.5 * L.wavevector[0] * powi(L.waist, 2); }
// fast_focused_laser.rs:129-131
static double ffl_n_cycles(const Laser& L) { return SPEED_OF_LIGHT * L.duration * L.wavevector[0] / (2.0 * PI); }

double Laser::max_timestep() const {
    switch (kind) {
    case Kind::FastFocusedLaser: {  // fast_focused_laser.rs:155-159
        double dt = 1.0 / ffl_omega(*this);
        double multiplier = std::fmin(std::sqrt(3.0) / (5.0 * ALPHA_FINE * a0), 0.1);
        return dt * multiplier;
    }
    case Kind::FastPlaneWave: {  // fast_plane_wave.rs:89-98
        double chirp = 1.0 + 2.0 * chirp_b * PI * n_cycles;
        double dt = 1.0 / (SPEED_OF_LIGHT * wavevector[0] * chirp);
        double multiplier = std::fmin(std::sqrt(3.0) / (5.0 * ALPHA_FINE * a0), 0.1);
        return dt * multiplier;
    }
    default: return lma_max_timestep(*this);
    }
}

bool Laser::contains(V4 r) const {
    if (kind == Kind::FastFocusedLaser || kind == Kind::FastPlaneWave) {
        // fast_focused_laser.rs:161-169 (n_cycles recomputed from the duration) / fast_plane_wave.rs:100-108
        double nc = kind == Kind::FastFocusedLaser ? ffl_n_cycles(*this) : n_cycles;
        double phase = wavevector * r;
        double max_phase = env == Env::CosSquared ? PI * nc : (env == Env::Flattop ? PI * (nc + 1.0) : 6.0 * PI * nc);
        return phase < max_phase;
    }
    return lma_contains(*this, r);
}

double Laser::ideal_initial_z() const {
    if (kind == Kind::FastFocusedLaser || kind == Kind::FastPlaneWave) {
        double nc = kind == Kind::FastFocusedLaser ? ffl_n_cycles(*this) : n_cycles;
        double wl = 2.0 * PI / wavevector[0];
        return env == Env::CosSquared ? 0.5 * wl * nc : (env == Env::Flattop ? 0.5 * wl * (nc + 1.0) : 2.0 * wl * nc);
    }
    return lma_ideal_initial_z(*this);
}

// fast_focused_laser.rs:61-111
struct Beam { V3 reE, imE, reB, imB; };
static Beam ffl_beam(const Laser& L, V4 r) {
    r = rotate_around_z(r, -L.pol_angle);
    double x = r[1] / L.waist;
    double y = r[2] / L.waist;
    double z = r[3] / ffl_rayleigh(L);
    double rho = std::hypot(x, y);
    double e = L.waist / ffl_rayleigh(L);
    const C i{0.0, 1.0};
    C f = i / (z + i);
    C prefactor = f * cexp(-f * powi(rho, 2) + i * (L.wavevector * r + 0.0));
    C Ex = -i * prefactor * (1.0 + powi(e, 2) * f * f * (powi(x, 2) - f * powi(rho, 4) / 4.0) +
                             powi(e, 4) * f * f * (1.0 / 8.0 - f * powi(rho, 2) / 4.0 + f * f * (powi(x * rho, 2) - powi(rho, 4) / 16.0) +
                                                   f * f * f * powi(rho, 4) * (-x * x - 0.5 * powi(rho, 2)) / 4.0 + f * f * f * f * powi(rho, 8) / 32.0));
    C Ey = -i * prefactor * (powi(e, 2) * f * f * x * y + powi(e, 4) * x * y * cpowf(f, 4.0) * powi(rho, 2) * (1.0 - f * powi(rho, 2) / 4.0));
    C Ez = prefactor * (e * f * x + powi(e, 3) * f * f * x * (-0.5 + f * powi(rho, 2) - f * f * powi(rho, 4) / 4.0));
    C Bx{0.0, 0.0};
    C By = -i * prefactor * (1.0 + powi(e, 2) * f * f * powi(rho, 2) * (1.0 - f * powi(rho, 2) / 2.0) / 2.0 +
                             powi(e, 4) * f * f * (-1.0 + 2.0 * f * powi(rho, 2) + 2.5 * f * f * powi(rho, 4) - 2.0 * f * f * f * powi(rho, 6) + 0.25 * f * f * f * f * powi(rho, 8)) / 8.0);
    C Bz = prefactor * (e * f * y + powi(e, 3) * f * f * y * (1.0 + f * powi(rho, 2) - f * f * powi(rho, 4) / 2.0) / 2.0);
    double E0 = ELECTRON_MASS * SPEED_OF_LIGHT * ffl_omega(L) * L.a0 / ELEMENTARY_CHARGE;
    double B0 = E0 / SPEED_OF_LIGHT;
    return {E0 * rotate_around_z(V3{Ex.re, Ey.re, Ez.re}, L.pol_angle), E0 * rotate_around_z(V3{Ex.im, Ey.im, Ez.im}, L.pol_angle),
            B0 * rotate_around_z(V3{Bx.re, By.re, Bz.re}, L.pol_angle), B0 * rotate_around_z(V3{Bx.im, By.im, Bz.im}, L.pol_angle)};
}

// fast_focused_laser.rs:135-153
static void ffl_envelope(const Laser& L, double phase, double* f, double* df) {
    double nc = ffl_n_cycles(L);
    switch (L.env) {
    case Env::CosSquared:
        if (std::fabs(phase) < PI * nc) {
            double env = powi(std::cos(phase / (2.0 * nc)), 2);
            *f = env;
            *df = -1.0 * std::tan(phase / (2.0 * nc)) * env / nc;
        } else { *f = 0.0; *df = 0.0; }
        return;
    case Env::Flattop:
        if (std::fabs(phase) > PI * (nc + 1.0)) { *f = 0.0; *df = 0.0; }
        else if (std::fabs(phase) > PI * (nc - 1.0)) {
            double arg = 0.25 * (std::fabs(phase) - (nc - 1.0) * PI);
            *f = powi(std::cos(arg), 2);
            *df = -0.25 * signum(phase) * std::sin(2.0 * arg);
        } else { *f = 1.0; *df = 0.0; }
        return;
    case Env::Gaussian: {
        double tau = ffl_omega(L) * L.duration;
        double env = std::exp(-2.0 * LN_2 * powi(phase, 2) / powi(tau, 2));
        *f = env;
        *df = -4.0 * LN_2 * phase * env / powi(tau, 2);
        return;
    }
    }
}

// fast_focused_laser.rs:190-218
static void ffl_fields(const Laser& L, V4 r, V3* E_out, V3* B_out, double* a_out) {
    double phase = L.wavevector * r;
    double f = 0.0, df = 0.0;
    ffl_envelope(L, phase, &f, &df);
    Beam b = ffl_beam(L, r);
    V3 E_x = f * b.reE + df * b.imE, B_x = f * b.reB + df * b.imB;
    V3 E_y{0.0, 0.0, 0.0}, B_y{0.0, 0.0, 0.0};
    if (L.pol == Pol::Circular) {
        V3 axis = normalize(spatial(L.wavevector));
        V3 rp = rotate_around(spatial(r), axis, -FRAC_PI_2);
        V4 r_prime{r[0], rp.x, rp.y, rp.z};
        Beam c = ffl_beam(L, r_prime);
        V3 ey = f * c.imE - df * c.reE, by = f * c.imB - df * c.reB;
        E_y = rotate_around(ey, axis, FRAC_PI_2);
        B_y = rotate_around(by, axis, FRAC_PI_2);
    }
    V3 E = E_x + E_y, B = B_x + B_y;
    *E_out = E;
    *B_out = B;
    *a_out = ELEMENTARY_CHARGE * std::sqrt(norm_sqr(E)) / (ELECTRON_MASS * SPEED_OF_LIGHT * ffl_omega(L));
}

// fast_plane_wave.rs:110-185 (default build: no compensating chirp)
static void fpw_fields(const Laser& L, V4 r, V3* E_out, V3* B_out, double* a_out) {
    double delta = L.pol == Pol::Linear ? 0.0 : 1.0;
    double phi = L.wavevector * r;
    double psi = phi * (1.0 + L.chirp_b * phi);
    double dpsi = 1.0 + 2.0 * L.chirp_b * phi;
    double f = 0.0, df = 0.0;
    double nc = L.n_cycles;
    switch (L.env) {
    case Env::CosSquared:
        if (std::fabs(phi) < nc * PI) { f = powi(std::cos(phi / (2.0 * nc)), 2); df = -std::sin(phi / nc) / (2.0 * nc); }
        else { f = 0.0; df = 0.0; }
        break;
    case Env::Flattop:
        if (std::fabs(phi) > PI * (nc + 1.0)) { f = 0.0; df = 0.0; }
        else if (std::fabs(phi) > PI * (nc - 1.0)) {
            double arg = 0.25 * (std::fabs(phi) - (nc - 1.0) * PI);
            f = powi(std::cos(arg), 2);
            df = -signum(phi) * 0.25 * std::sin(2.0 * arg);
        } else { f = 1.0; df = 0.0; }
        break;
    default: {
        double arg = -0.5 * powi(phi / (PI * nc), 2);
        f = std::exp2(arg);
        df = -LN_2 * phi * std::exp2(arg) / powi(PI * nc, 2);
    }
    }
    double dax = std::sin(psi) * df + std::cos(psi) * dpsi * f;
    double day = delta * (std::cos(psi) * df - std::sin(psi) * dpsi * f);
    double amplitude = (ELECTRON_MASS * SPEED_OF_LIGHT_SQD * L.wavevector[0] * L.a0) / ELEMENTARY_CHARGE;
    V3 E = -amplitude * V3{dax, day, 0.0};
    V3 B = (amplitude / SPEED_OF_LIGHT) * V3{day, -dax, 0.0};
    E = rotate_around_z(E, L.pol_angle);
    B = rotate_around_z(B, L.pol_angle);
    *E_out = E;
    *B_out = B;
    *a_out = ELEMENTARY_CHARGE * std::sqrt(norm_sqr(E)) / (ELECTRON_MASS * SPEED_OF_LIGHT * (SPEED_OF_LIGHT * L.wavevector[0]));
}

void Laser::fields(V4 r, V3* E, V3* B, double* a) const {
    if (kind == Kind::FastFocusedLaser) ffl_fields(*this, r, E, B, a);
    else if (kind == Kind::FastPlaneWave) fpw_fields(*this, r, E, B, a);
    else { *E = {0, 0, 0}; *B = {0, 0, 0}; *a = 0.0; }  // field/mod.rs:183-186 default
}

// field/lcf.rs:18-78
PushOut vay_push(V4 r, V4 ui, V3 E, V3 B, double rqm, double dt, Eqn eqn) {
    V3 u = spatial(ui);
    double gamma = std::sqrt(1.0 + u * u);
    V3 v = SPEED_OF_LIGHT * u / gamma;
    double alpha = rqm * dt / (2.0 * SPEED_OF_LIGHT);
    V3 u_half = u + alpha * (E + cross(v, B));
    double gamma_half = std::sqrt(1.0 + u_half * u_half);
    double dwork = rqm * (E * u_half) * dt / (gamma_half * SPEED_OF_LIGHT);
    V3 u_rad{0.0, 0.0, 0.0};
    if (eqn != Eqn::Lorentz) {
        double mag = std::sqrt(norm_sqr(u_half));
        V3 beta = u_half / mag;
        double E_rf_sqd = norm_sqr(E + SPEED_OF_LIGHT * cross(beta, B)) - powi(E * beta, 2);
        double chi = E_rf_sqd > 0.0 ? gamma_half * std::sqrt(E_rf_sqd) / CRITICAL_FIELD : 0.0;
        double power = 2.0 * ALPHA_FINE * chi * chi / (3.0 * COMPTON_TIME);
        double g_chi = eqn == Eqn::ModifiedLandauLifshitz ? gaunt_factor(chi) : 1.0;
        u_rad = g_chi * power * dt * u_half / mag;
    }
    V3 u_prime = u_half + alpha * E;
    double gamma_prime_sqd = 1.0 + u_prime * u_prime;
    V3 tau = alpha * SPEED_OF_LIGHT * B;
    double u_star = u_prime * tau;
    double sigma = gamma_prime_sqd - tau * tau;
    gamma = std::sqrt(0.5 * sigma + std::sqrt(0.25 * powi(sigma, 2) + tau * tau + powi(u_star, 2)));
    V3 t = tau / gamma;
    double s = 1.0 / (1.0 + t * t);
    V3 u_new = s * (u_prime + (u_prime * t) * t + cross(u_prime, t));
    u_new = u_new - u_rad;
    gamma = std::sqrt(1.0 + u_new * u_new);
    V4 u4{gamma, u_new.x, u_new.y, u_new.z};
    V4 r_new = r + 0.5 * SPEED_OF_LIGHT * u4 * dt / gamma;
    return {r_new, u4, dt, dwork};
}

PushOut lma_push(const Laser& L, V4 r, V4 u, double rqm, double dt, Eqn eqn);

// field/mod.rs:147-152
PushOut field_push(const Laser& L, V4 r, V4 u, double rqm, double dt, Eqn eqn) {
    if (L.kind == Kind::PlaneWave || L.kind == Kind::FocusedLaser) return lma_push(L, r, u, rqm, dt, eqn);
    V4 rh = r + 0.5 * SPEED_OF_LIGHT * u * dt / u[0];
    V3 E, B;
    double a;
    L.fields(rh, &E, &B, &a);
    return vay_push(rh, u, E, B, rqm, dt, eqn);
}

}  // namespace orc
