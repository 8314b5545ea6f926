// Ptarmigan CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.hpp).
// collide() (src/main.rs:134-289), the LCFA event handlers (src/field/lcf.rs:85-193),
// beam generation (src/particle/builder.rs:173-219) and the extern "C" entry
// points used by the tests and the CPU baseline.
#include <algorithm>
#include <chrono>
#include <thread>

#include "oracle.hpp"
#include "../include/ptg.h"

namespace orc {

struct Particle {
    int species;  // 0 e-, 1 e+, 2 photon
    V4 r[2], u[2];
    SV pol;
    double payload, parent_chi, interaction_count, work, weight;
    uint64_t id, parent_id;
};

// particle/mod.rs:108-126
static Particle create(int species, V4 r) {
    V4 u = species == 2 ? V4{0, 0, 0, 0} : V4{1, 0, 0, 0};
    return Particle{species, {r, r}, {u, u}, unpolarized(), 0, 0, 0, 0, 1.0, 0, 0};
}
static double rqm(const Particle& p) {  // particle/mod.rs:141-147
    if (p.species == 0) return ELECTRON_CHARGE / ELECTRON_MASS;
    if (p.species == 1) return -ELECTRON_CHARGE / ELECTRON_MASS;
    return 0.0;
}

struct Options {
    double dt_multiplier, rate_increase, t_stop, min_energy, max_angle;
    int64_t multiplicity;
    bool discard_bg_e, discard_bg_ph, rr, tracking_photons, tracking_secondaries, keep_decayed_photons,
        pol_resolved, rotate_stokes_pars, classical, gaunt_factor;
};

struct RadEvent { V4 k, u_prime; SV pol; double a_eff, chi, absorption; };
struct PairEvent { V4 u_e, u_p; double frac, a_eff, chi, absorption; };

struct Counters { uint64_t lepton_steps = 0, photon_steps = 0, emissions = 0, pairs = 0; };

// LMA event handlers (plane_wave.rs:251-316, focused_laser.rs), defined in lma.cpp
bool lma_radiate(const Laser& L, V4 r, V4 u, double dt, Xoshiro& rng, bool classical, RadEvent* ev);
double lma_pair_create(const Laser& L, V4 r, V4 ell, SV pol, double dt, Xoshiro& rng, double rate_increase, SV* pol_new,
                       bool* happened, PairEvent* ev);

// field/lcf.rs:85-134
static bool lcf_radiate(V4 u, V3 E, V3 B, double a, double dt, Xoshiro& rng, bool classical, RadEvent* ev) {
    V3 beta = spatial(u) / u[0];
    double E_rf_sqd = norm_sqr(E + SPEED_OF_LIGHT * cross(beta, B)) - powi(E * beta, 2);
    double chi = E_rf_sqd > 0.0 ? u[0] * std::sqrt(E_rf_sqd) / CRITICAL_FIELD : 0.0;
    double prob = classical ? dt * lcfa_classical_rate(chi, u[0]) : dt * lcfa_rate(chi, u[0]);
    if (!(rng.gen_f64() < prob)) return false;
    double r1 = rng.gen_f64(), r2 = rng.gen_f64(), r3 = rng.gen_f64();
    double omega, theta, cphi;
    bool ok;
    if (classical) lcfa_classical_sample(chi, u[0], r1, r2, r3, &omega, &ok, &theta, &cphi);
    else lcfa_sample(chi, u[0], r1, r2, r3, &omega, &ok, &theta, &cphi);
    if (!ok) return false;
    V3 lng = normalize(beta);
    V3 w = -normalize(E - (lng * E) * lng / std::sqrt(norm_sqr(E)) + SPEED_OF_LIGHT * cross(beta, B));
    V3 perp = rotate_around(w, lng, cphi);
    V3 k3 = omega * (std::cos(theta) * lng + std::sin(theta) * perp);
    V4 k = lightlike(k3.x, k3.y, k3.z);
    SV pol = classical ? lcfa_classical_stokes(k, chi, u[0], beta, w) : lcfa_stokes(k, chi, u[0], beta, w);
    *ev = RadEvent{k, u - k, pol, a, chi, 0.0};
    return true;
}

// field/lcf.rs:144-193
static double lcf_pair_create(V4 u, SV sv, V3 E, V3 B, double a, double dt, Xoshiro& rng, double rate_increase,
                              SV* sv_out, bool* happened, PairEvent* ev) {
    V3 n = normalize(spatial(u));
    V3 a_perp = E - (E * n) * n + SPEED_OF_LIGHT * cross(n, B);
    double E_rf_sqd = norm_sqr(a_perp);
    double chi = 0.0, prob = 0.0;
    SV sv_new = sv;
    if (E_rf_sqd > 0.0) {
        chi = u[0] * std::sqrt(E_rf_sqd) / CRITICAL_FIELD;
        prob = pair_probability(u, sv, chi, a_perp, dt, &sv_new);
    }
    if (prob * rate_increase > 0.1) rate_increase = 0.1 / prob;
    *sv_out = sv_new;
    *happened = false;
    if (rng.gen_f64() < prob * rate_increase) {
        double gamma_p, cos_theta, cphi, s_, z_;
        pair_sample(u, sv, chi, a_perp, rng, &gamma_p, &cos_theta, &cphi, &s_, &z_);
        double sin_theta = std::sqrt(1.0 - cos_theta * cos_theta);
        double up = gamma_p * std::sqrt(1.0 - 1.0 / (gamma_p * gamma_p));
        V3 e_1 = normalize(a_perp);
        V3 e_2 = cross(n, e_1);
        V3 u_p = up * (cos_theta * n + sin_theta * std::cos(cphi) * e_1 + sin_theta * std::sin(cphi) * e_2);
        V3 u_e = spatial(u) - u_p;
        *ev = PairEvent{unitize(V4{0.0, u_e.x, u_e.y, u_e.z}), unitize(V4{0.0, u_p.x, u_p.y, u_p.z}), 1.0 / rate_increase, a, chi, 0.0};
        *happened = true;
    }
    return prob;
}

static bool radiate(const Laser& L, V4 r, V4 u, double dt, Xoshiro& rng, bool classical, RadEvent* ev) {
    if (L.kind == Kind::PlaneWave || L.kind == Kind::FocusedLaser) return lma_radiate(L, r, u, dt, rng, classical, ev);
    V3 E, B;
    double a;
    L.fields(r, &E, &B, &a);  // field/mod.rs:158-161
    return lcf_radiate(u, E, B, a, dt, rng, classical, ev);
}

static double pair_create(const Laser& L, V4 r, V4 ell, SV pol, double dt, Xoshiro& rng, double rate_increase, SV* pol_new,
                          bool* happened, PairEvent* ev) {
    if (L.kind == Kind::PlaneWave || L.kind == Kind::FocusedLaser)
        return lma_pair_create(L, r, ell, pol, dt, rng, rate_increase, pol_new, happened, ev);
    V3 E, B;
    double a;
    L.fields(r, &E, &B, &a);  // field/mod.rs:174-177
    return lcf_pair_create(ell, pol, E, B, a, dt, rng, rate_increase, pol_new, happened, ev);
}

// src/main.rs:134-289
static void collide(const Laser& L, const Particle& incident, Xoshiro& rng, uint64_t& current_id, const Options& o,
                    std::vector<Particle>& secondaries_out, std::vector<Particle>& intermediates_out, Counters& cnt) {
    std::vector<Particle> primaries{incident};
    std::vector<Particle> secondaries, intermediates;
    double dt = L.max_timestep() * o.dt_multiplier;
    uint64_t primary_id = incident.id;
    Eqn eqn = (o.classical && o.rr) ? (o.gaunt_factor ? Eqn::ModifiedLandauLifshitz : Eqn::LandauLifshitz) : Eqn::Lorentz;
    bool classical_mode = o.classical && !o.gaunt_factor;
    bool electron_recoils = !o.classical && o.rr;

    while (!primaries.empty()) {
        Particle pt = primaries.back();
        primaries.pop_back();
        if (!o.tracking_secondaries && pt.id != primary_id) { secondaries.push_back(pt); continue; }
        if (pt.species != 2) {
            while (L.contains(pt.r[1]) && pt.r[1][0] / SPEED_OF_LIGHT < o.t_stop) {
                cnt.lepton_steps++;
                PushOut po = field_push(L, pt.r[1], pt.u[1], rqm(pt), dt, eqn);
                V4 u = po.u;
                RadEvent ev;
                if (radiate(L, po.r, u, po.dt, rng, classical_mode, &ev)) {
                    cnt.emissions++;
                    uint64_t id = current_id++;
                    Particle ph = create(2, po.r);
                    ph.payload = ev.a_eff;
                    ph.parent_chi = ev.chi;
                    ph.weight = pt.weight;
                    ph.id = id;
                    ph.parent_id = pt.id;
                    ph.pol = ev.pol;
                    ph.u[1] = ev.k;
                    primaries.push_back(ph);
                    if (electron_recoils) u = ev.u_prime;
                    pt.interaction_count += 1.0;
                    if (!o.classical) pt.work += ev.absorption;
                }
                pt.r[1] = po.r;
                pt.u[1] = u;
                pt.work += po.work;
            }
            if (pt.id != primary_id || !o.discard_bg_e || pt.interaction_count > 0.0) secondaries.push_back(pt);
        } else {
            bool has_decayed = false;
            while (L.contains(pt.r[1]) && pt.r[1][0] / SPEED_OF_LIGHT < o.t_stop && !has_decayed && o.tracking_photons) {
                cnt.photon_steps++;
                V4 ell = pt.u[1];
                V4 r = pt.r[1] + SPEED_OF_LIGHT * ell * dt / ell[0];
                SV pol = o.pol_resolved ? pt.pol : unpolarized();
                SV pol_new;
                bool happened;
                PairEvent ev;
                double prob = pair_create(L, r, ell, pol, dt, rng, o.rate_increase, &pol_new, &happened, &ev);
                if (happened) {
                    cnt.pairs++;
                    uint64_t id = current_id;
                    current_id += 2;
                    Particle e = create(0, r), p = create(1, r);
                    for (Particle* q : {&e, &p}) {
                        q->weight = ev.frac * pt.weight;
                        q->payload = ev.a_eff;
                        q->parent_chi = ev.chi;
                        q->parent_id = pt.id;
                        q->work += 0.5 * ev.absorption;
                    }
                    e.id = id;
                    p.id = id + 1;
                    e.u[1] = ev.u_e;
                    p.u[1] = ev.u_p;
                    primaries.push_back(e);
                    primaries.push_back(p);
                    if (ev.frac == 1.0) has_decayed = true;
                    else pt.weight = pt.weight * (1.0 - ev.frac);
                }
                if (o.rotate_stokes_pars && o.pol_resolved) pt.pol = pol_new;
                pt.interaction_count += prob;
                pt.r[1] = r;
            }
            if (!has_decayed && (pt.id != primary_id || !o.discard_bg_ph)) secondaries.push_back(pt);
            if (has_decayed && o.keep_decayed_photons) intermediates.push_back(pt);
        }
    }

    V3 n0 = normalize(spatial(ELECTRON_MASS_MEV * incident.u[1]));
    std::vector<Particle> kept;
    for (const Particle& pt : secondaries) {
        V4 p = ELECTRON_MASS_MEV * pt.u[1];
        V3 n = normalize(spatial(p));
        if (p[0] > o.min_energy && n0 * n > std::cos(o.max_angle)) kept.push_back(pt);
    }
    if (o.multiplicity >= 0 && kept.size() != size_t(o.multiplicity) + 1) kept.clear();
    secondaries_out.insert(secondaries_out.end(), kept.begin(), kept.end());
    intermediates_out.insert(intermediates_out.end(), intermediates.begin(), intermediates.end());
}

}  // namespace orc

// ======================================================================== C API
using namespace orc;

static Laser to_laser(const ptg_laser_desc* d) {
    Laser L{Kind(d->kind), d->a0, d->wavelength, d->waist, d->n_cycles, Env(d->envelope), Pol(d->polarization),
            d->pol_angle, d->chirp_b, d->finite_bandwidth != 0, {}, 0.0};
    L.init();
    return L;
}

static Options to_options(const ptg_collide_options* o) {
    return Options{o->dt_multiplier, o->rate_increase, o->t_stop, o->min_energy, o->max_angle, o->multiplicity,
                   o->discard_bg_e != 0, o->discard_bg_ph != 0, o->rr != 0, o->tracking_photons != 0,
                   o->tracking_secondaries != 0, o->keep_decayed_photons != 0, o->pol_resolved != 0,
                   o->rotate_stokes_pars != 0, o->classical != 0, o->gaunt_factor != 0};
}

static V4 get4(const double* a, size_t i) { return {a[4 * i], a[4 * i + 1], a[4 * i + 2], a[4 * i + 3]}; }
static void put4(double* a, size_t i, V4 v) { a[4 * i] = v.t; a[4 * i + 1] = v.x; a[4 * i + 2] = v.y; a[4 * i + 3] = v.z; }

static Particle read_particle(const ptg_particles* in, size_t i) {
    Particle p = create(in->species[i], get4(in->r, i));
    p.u[1] = get4(in->u, i);  // u[0] keeps create()'s default, as with_normalized_momentum does
    if (in->r0) p.r[0] = get4(in->r0, i);
    if (in->u0) p.u[0] = get4(in->u0, i);
    if (in->pol) { V4 s = get4(in->pol, i); p.pol = {s.t, s.x, s.y, s.z}; }
    if (in->payload) p.payload = in->payload[i];
    if (in->parent_chi) p.parent_chi = in->parent_chi[i];
    if (in->interaction_count) p.interaction_count = in->interaction_count[i];
    if (in->work) p.work = in->work[i];
    if (in->weight) p.weight = in->weight[i];
    p.id = in->id ? in->id[i] : i;
    p.parent_id = in->parent_id ? in->parent_id[i] : p.id;
    return p;
}

static void write_particle(const ptg_particles* out, size_t i, const Particle& p) {
    if (out->species) out->species[i] = uint8_t(p.species);
    if (out->r) put4(out->r, i, p.r[1]);
    if (out->u) put4(out->u, i, p.u[1]);
    if (out->r0) put4(out->r0, i, p.r[0]);
    if (out->u0) put4(out->u0, i, p.u[0]);
    if (out->pol) put4(out->pol, i, V4{p.pol.i, p.pol.q, p.pol.u, p.pol.v});
    if (out->payload) out->payload[i] = p.payload;
    if (out->parent_chi) out->parent_chi[i] = p.parent_chi;
    if (out->interaction_count) out->interaction_count[i] = p.interaction_count;
    if (out->work) out->work[i] = p.work;
    if (out->weight) out->weight[i] = p.weight;
    if (out->id) out->id[i] = p.id;
    if (out->parent_id) out->parent_id[i] = p.parent_id;
}

struct OrcRun {
    std::vector<Particle> secondaries, intermediates;
    Counters cnt;
};

extern "C" {

int orc_init(const char* table_path) {
    std::string err;
    return load_tables(table_path, &err) ? 0 : -2;
}

/* Runs collide() over n primaries with `nthreads` emulated MPI ranks: rank k takes
 * the k-th contiguous slice (src/main.rs:783-789) and a xoshiro256** stream seeded
 * as src/main.rs:1036-1040 (jumped k times). Returns an opaque handle. */
void* orc_collide(const ptg_laser_desc* ld, const ptg_collide_options* od, const ptg_particles* in, size_t n,
                  uint64_t rng_seed, int nthreads, double* seconds) {
    Laser L = to_laser(ld);
    Options o = to_options(od);
    if (nthreads < 1) nthreads = 1;
    std::vector<OrcRun> parts(nthreads);
    std::vector<Particle> prim(n);
    for (size_t i = 0; i < n; ++i) prim[i] = read_particle(in, i);
    auto t0 = std::chrono::steady_clock::now();
    auto work = [&](int rank) {
        size_t lo = n * size_t(rank) / size_t(nthreads), hi = n * size_t(rank + 1) / size_t(nthreads);
        uint64_t seed = 0x8658b90036b165ebULL + rng_seed * 0x32f55cddaebae910ULL;
        Xoshiro rng(seed);
        for (int j = 0; j < rank; ++j) rng.jump();
        uint64_t current_id = n;  // ids unique per rank, as in the reference before the post-loop offset
        OrcRun& R = parts[rank];
        for (size_t i = lo; i < hi; ++i) collide(L, prim[i], rng, current_id, o, R.secondaries, R.intermediates, R.cnt);
    };
    if (nthreads == 1) work(0);
    else {
        std::vector<std::thread> th;
        for (int k = 0; k < nthreads; ++k) th.emplace_back(work, k);
        for (auto& t : th) t.join();
    }
    auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    OrcRun* out = new OrcRun;
    for (auto& p : parts) {
        out->secondaries.insert(out->secondaries.end(), p.secondaries.begin(), p.secondaries.end());
        out->intermediates.insert(out->intermediates.end(), p.intermediates.begin(), p.intermediates.end());
        out->cnt.lepton_steps += p.cnt.lepton_steps;
        out->cnt.photon_steps += p.cnt.photon_steps;
        out->cnt.emissions += p.cnt.emissions;
        out->cnt.pairs += p.cnt.pairs;
    }
    return out;
}

void orc_run_counts(void* h, uint64_t* c /* [6]: lepton_steps, photon_steps, emissions, pairs, n_sec, n_int */) {
    OrcRun* R = static_cast<OrcRun*>(h);
    c[0] = R->cnt.lepton_steps; c[1] = R->cnt.photon_steps; c[2] = R->cnt.emissions; c[3] = R->cnt.pairs;
    c[4] = R->secondaries.size(); c[5] = R->intermediates.size();
}

void orc_run_fetch(void* h, ptg_particles* sec, ptg_particles* inter) {
    OrcRun* R = static_cast<OrcRun*>(h);
    if (sec) for (size_t i = 0; i < R->secondaries.size(); ++i) write_particle(sec, i, R->secondaries[i]);
    if (inter) for (size_t i = 0; i < R->intermediates.size(); ++i) write_particle(inter, i, R->intermediates[i]);
}

void orc_run_free(void* h) { delete static_cast<OrcRun*>(h); }

/* Field-only push, stopping when a particle leaves the field (Field::push). */
void orc_push(const ptg_laser_desc* ld, const uint8_t* species, double* r, double* u, size_t n, uint32_t n_steps,
              double dt_multiplier, int eqn, uint32_t* steps_taken, double* work) {
    Laser L = to_laser(ld);
    double dt = L.max_timestep() * dt_multiplier;
    for (size_t i = 0; i < n; ++i) {
        V4 ri = get4(r, i), ui = get4(u, i);
        double q = species[i] == 0 ? ELECTRON_CHARGE / ELECTRON_MASS : (species[i] == 1 ? -ELECTRON_CHARGE / ELECTRON_MASS : 0.0);
        uint32_t k = 0;
        double w = 0.0;
        for (; k < n_steps && L.contains(ri); ++k) {
            PushOut po = field_push(L, ri, ui, q, dt, Eqn(eqn));
            ri = po.r;
            ui = po.u;
            w += po.work;
        }
        if (work) work[i] = w;
        put4(r, i, ri);
        put4(u, i, ui);
        if (steps_taken) steps_taken[i] = k;
    }
}

/* BeamBuilder::build for the normal-gamma beam (particle/builder.rs:173-219). */
void orc_build_beam(const ptg_beam_desc* b, double initial_z, uint64_t rng_seed, ptg_particles* out) {
    uint64_t seed = 0x8658b90036b165ebULL + rng_seed * 0x32f55cddaebae910ULL;
    Xoshiro rng(seed);
    double rho = -b->energy_chirp;
    for (uint64_t i = 0; i < b->n; ++i) {
        double gamma, dz;
        for (;;) {
            double n0 = rng.normal(), n1 = rng.normal();
            double n2 = rho * n0 + std::sqrt(1.0 - rho * rho) * n1;
            dz = b->length * n0;
            gamma = b->gamma + b->sigma * n2;
            if (gamma > 1.0) break;
        }
        double uu = b->species == 2 ? -gamma : -std::sqrt(gamma * gamma - 1.0);
        double theta_x = b->angle + b->rms_div * rng.normal();
        double theta_y = b->rms_div * rng.normal();
        V3 u3{uu * std::sin(theta_x) * std::cos(theta_y), uu * std::sin(theta_y), uu * std::cos(theta_x) * std::cos(theta_y)};
        u3 = rotate_around_z(u3, b->collision_plane_angle);
        V4 u4 = b->species == 2 ? lightlike(u3.x, u3.y, u3.z) : unitize(V4{0.0, u3.x, u3.y, u3.z});
        double t, z;
        if (b->offset[2] >= 0.0) { t = -initial_z; z = initial_z + b->offset[2] + dz; }
        else { t = -initial_z - std::fabs(b->offset[2]); z = initial_z + dz; }
        double x, y;
        if (b->radial == 0) { x = b->radius * rng.normal(); y = b->radius * rng.normal(); }
        else if (b->radial == 1) {
            for (;;) {
                x = b->radius * rng.normal();
                y = b->radius * rng.normal();
                if (x * x / (b->max_radius * b->max_radius) + y * y / (b->max_radius * b->max_radius) <= 1.0) break;
            }
        } else {
            double rr = b->radius * std::sqrt(rng.gen_f64());
            double th = 2.0 * PI * rng.gen_f64();
            x = rr * std::cos(th);
            y = rr * std::sin(th);
        }
        V3 r3{x + b->offset[0], y + b->offset[1], z};
        r3 = rotate_around_y(r3, b->angle);
        r3 = rotate_around_z(r3, b->collision_plane_angle);
        Particle p = create(b->species, V4{t, r3.x, r3.y, r3.z});
        p.u[1] = u4;
        p.pol = {b->pol[0], b->pol[1], b->pol[2], b->pol[3]};
        p.weight = b->weight;
        p.id = p.parent_id = i;
        write_particle(out, i, p);
    }
}

double orc_laser_max_timestep(const ptg_laser_desc* ld) { return to_laser(ld).max_timestep(); }
double orc_laser_ideal_initial_z(const ptg_laser_desc* ld) { return to_laser(ld).ideal_initial_z(); }
int orc_laser_contains(const ptg_laser_desc* ld, const double* r) { return to_laser(ld).contains(get4(r, 0)); }
void orc_laser_fields(const ptg_laser_desc* ld, const double* r, size_t n, double* out7) {
    Laser L = to_laser(ld);
    for (size_t i = 0; i < n; ++i) {
        V3 E, B;
        double a;
        L.fields(get4(r, i), &E, &B, &a);
        double* o = out7 + 7 * i;
        o[0] = E.x; o[1] = E.y; o[2] = E.z; o[3] = B.x; o[4] = B.y; o[5] = B.z; o[6] = a;
    }
}

// ---- scalar function probes (parity and known-answer tests)
double orc_lcfa_rate(double chi, double gamma) { return lcfa_rate(chi, gamma); }
void orc_lcfa_sample(double chi, double gamma, double r1, double r2, double r3, double* out4) {
    double om, th, ph;
    bool ok;
    lcfa_sample(chi, gamma, r1, r2, r3, &om, &ok, &th, &ph);
    out4[0] = om; out4[1] = ok ? 1.0 : 0.0; out4[2] = th; out4[3] = ph;
}
void orc_lcfa_stokes(const double* k4, double chi, double gamma, const double* v3, const double* w3, double* out4) {
    SV s = lcfa_stokes(get4(k4, 0), chi, gamma, V3{v3[0], v3[1], v3[2]}, V3{w3[0], w3[1], w3[2]});
    out4[0] = s.i; out4[1] = s.q; out4[2] = s.u; out4[3] = s.v;
}
double orc_gaunt(double chi) { return gaunt_factor(chi); }
void orc_aux_t(double chi, double* out2) { auxiliary_t(chi, &out2[0], &out2[1]); }
double orc_pair_spectrum(double s, double chi, double sv1) { return pair_spectrum(s, chi, sv1); }
double orc_pair_spectrum_ceiling(double chi, double sv1) { return pair_spectrum_ceiling(chi, sv1); }
double orc_pair_angular_spectrum(double y, double s, double chi, double sv1) { return pair_angular_spectrum(y, s, chi, sv1); }
double orc_pair_angular_spectrum_ceiling(double s, double chi, double sv1) { return pair_angular_spectrum_ceiling(s, chi, sv1); }
double orc_pair_probability(const double* ell4, const double* sv4, double chi, const double* a3, double dt, double* sv_out4) {
    SV n;
    double p = pair_probability(get4(ell4, 0), SV{sv4[0], sv4[1], sv4[2], sv4[3]}, chi, V3{a3[0], a3[1], a3[2]}, dt, &n);
    sv_out4[0] = n.i; sv_out4[1] = n.q; sv_out4[2] = n.u; sv_out4[3] = n.v;
    return p;
}
int orc_airy(double x, double* out4) {
    int ok = 0;
    ok |= airy_ai(x, &out4[0]) ? 1 : 0;
    ok |= airy_ai_prime(x, &out4[1]) ? 2 : 0;
    if (x > 0.0) {
        ok |= bessel_k_1_3(x, &out4[2]) ? 4 : 0;
        ok |= bessel_k_2_3(x, &out4[3]) ? 8 : 0;
    }
    return ok;
}
int orc_pwmci(const double* tbl, int n, double v, int invert, int ext, double* out) {
    Interp ip{tbl, n, ext != 0};
    return invert ? ip.invert(v, out) : ip.evaluate(v, out);
}
void orc_xoshiro_stream(uint64_t seed, int jumps, size_t n, uint64_t* out) {
    Xoshiro r(seed);
    for (int j = 0; j < jumps; ++j) r.jump();
    for (size_t i = 0; i < n; ++i) out[i] = r.next_u64();
}

}  // extern "C"
