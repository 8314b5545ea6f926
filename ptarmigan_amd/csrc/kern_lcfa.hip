// Transport kernels for the locally-constant-field path (FastFocusedLaser,
// FastPlaneWave): the electron/positron loop and the photon loop of collide()
// (src/main.rs:169-264), one thread per particle.
//
// Design
//  * Persistent lanes: each lane pulls the next particle index from an atomic
//    work counter as soon as its current particle leaves the field, so lanes of a
//    wave stay busy although particles need very different numbers of steps.
//  * The tables read every step (3 KB) are staged into LDS once per workgroup;
//    the event samplers read theirs from global memory (L2-resident).
//  * Emission/decay decisions use the survival method: a particle draws one
//    threshold U and multiplies S by (1 - p_k) every step; the event fires at
//    the first step with S < U. The event step has exactly the distribution of
//    the reference's per-step Bernoulli draws (rng < p_k), with one RNG call per
//    event instead of one per step.
//  * Secondaries, final states and forwarded records take their slots with one
//    atomic per wave (wave_reserve: ballot, mbcnt rank, readlane base).
#include "kern_common.h"

// Occupancy target of the transport kernels (waves per SIMD); bounds the
// registers of the step loop and of the out-of-line event samplers it calls.
#ifndef PTG_WAVES_PER_EU
#define PTG_WAVES_PER_EU 3
#endif

namespace ptg {

// Field variants are compile-time: KIND (focused / plane) and CP (polarisation).
template <int KIND, bool CP>
__device__ __forceinline__ void lcfa_fields(const LaserDev& L, d4 r, d3& E, d3& B) {
    if (KIND == K_FAST_FOCUSED_LASER) ffl_fields_t<CP>(L, r, E, B);
    else fpw_fields(L, r, E, B);
}

// Emission event (field/lcf.rs:101-133): samples the photon, writes it to the
// secondary queue and returns its momentum. Kept out of line so the registers
// it needs do not add to those of the step loop.
struct EmitOut { double kx, ky, kz, kt; uint32_t c0, c1; int ok; };

__device__ __noinline__ EmitOut lcfa_emission(const TransportArgs* __restrict__ A, const double* tb, double chi, double g,
                                              double bx, double by, double bz, double ex, double ey, double ez,
                                              double vx, double vy, double vz, double rt, double rx, double ry, double rz,
                                              uint32_t idx, uint32_t nev, Rng rng) {
    EmitOut o{0.0, 0.0, 0.0, 0.0, 0u, 0u, 0};
    double r1, r2, r3, un;
    rng_uniform2(rng, r1, r2);
    rng_uniform2(rng, r3, un);
    double omega, theta, cphi;
    bool ok;
    if (A->classical_mode) lcfa_classical_sample(tb, chi, g, r1, r2, r3, omega, ok, theta, cphi);
    else lcfa_sample(tb, A->sc, chi, g, r1, r2, r3, omega, ok, theta, cphi);
    o.c0 = rng.c0;
    o.c1 = rng.c1;
    if (!ok) return o;
    d3 beta{bx, by, bz}, E{ex, ey, ez}, vxb{vx, vy, vz};
    d3 lng = unit(beta);
    double Emag = sqrt(nsq(E));
    d3 w = -unit(E - (dot(lng, E) / Emag) * lng + vxb);  // as the reference: projection divided by |E|
    double sp_, cp_, st, ct;
    sincos(cphi, &sp_, &cp_);
    sincos(theta, &st, &ct);
    d3 perp = rotate(w, lng, sp_, cp_);
    d3 k3 = omega * (ct * lng + st * perp);
    d4 k{sqrt(nsq(k3)), k3.x, k3.y, k3.z};
    sv4 pol = lcfa_stokes(tb, k, chi, g, beta, w, A->classical_mode);
    uint32_t flags = 0;
    emit_photon(*A, d4{rt, rx, ry, rz}, k, pol, A->laser.a_norm * Emag, chi, A->in.d[P_WEIGHT][idx], A->in.id[idx],
                rng_child(A->in.stream[idx], nev), A->in.shower[idx], flags);
    o.kx = k.x; o.ky = k.y; o.kz = k.z; o.kt = k.t;
    o.ok = 1;
    return o;
}

// Electron/positron transport (src/main.rs:169-210 with field/mod.rs:147-161,
// field/lcf.rs:18-134).
template <int KIND, bool CP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PTG_WAVES_PER_EU))) void k_leptons_lcfa(const TransportArgs* __restrict__ A) {
    __shared__ double tb[lcfa_pack::HOT_SIZE];  // per-step tables; event tables are read from global memory
    for (int i = threadIdx.x; i < lcfa_pack::HOT_SIZE; i += blockDim.x) tb[i] = A->lcfa_tables[i];
    __syncthreads();
    const double* gt = A->lcfa_tables;
    const LaserDev& L = A->laser;
    const double dt = A->dt;
    const double half_cdt = 0.5 * C_LIGHT * dt;

    uint64_t n_steps = 0, n_emit = 0;
    uint32_t flags = 0, psteps = 0;

    // per-particle state
    uint32_t idx = 0;
    d4 r;
    d3 u;
    double ut = 1.0, work = 0.0, icount = 0.0, S = 1.0, U = 0.0;
    uint32_t nev = 0, meta = 0;
    double rqm = RQM_ELECTRON;
    Rng rng{};
    bool skip = false;

    auto fetch = [&]() -> bool {
        for (;;) {
            uint32_t i = wave_reserve32(A->work_ctr, 1u);
            if (i >= A->n_in) return false;
            const PBuf& in = A->in;
            meta = in.meta[i];
            if ((meta & 3u) == 2u) { forward_record(*A, i); continue; }
            idx = i;
            r = {in.d[P_RT][i], in.d[P_RX][i], in.d[P_RY][i], in.d[P_RZ][i]};
            u = {in.d[P_UX][i], in.d[P_UY][i], in.d[P_UZ][i]};
            ut = in.d[P_UT][i];
            work = in.d[P_WORK][i];
            icount = in.d[P_ICOUNT][i];
            nev = in.nev[i];
            rqm = (meta & 3u) == 1u ? -RQM_ELECTRON : RQM_ELECTRON;
            rng = rng_make(A->key, in.stream[i]);
            U = rng_uniform(rng);
            S = 1.0;
            skip = !A->opt.tracking_secondaries && !(meta & META_PRIMARY);
            psteps = 0;
            return true;
        }
    };
    bool live = fetch();
    while (live) {
        bool inside = !skip && field_contains(L, r) && r.t / C_LIGHT < A->opt.t_stop;
        if (inside && psteps >= A->step_cap) {
            flags |= PTG_FLAG_STEP_CAP;
            inside = false;
        }
        if (!inside) {
            finalize_lepton(*A, idx, r, ut, u, work, icount, meta);
            live = fetch();
            continue;
        }
        ++n_steps;
        ++psteps;
        // Field::push: half drift, fields at mid-step, Vay push, half drift
        double hg0 = half_cdt / ut;
        d4 rh{r.t + half_cdt, r.x + hg0 * u.x, r.y + hg0 * u.y, r.z + hg0 * u.z};
        d3 E, B;
        lcfa_fields<KIND, CP>(L, rh, E, B);
        // the work is added here, before the event branch (an LCFA emission adds no
        // absorption, so the sum is the one the reference forms after radiate): the
        // push's terms are then not kept live across that branch (measured: 11 % less
        // HBM traffic from register spills)
        work += vay_push(u, E, B, rqm, dt, A->rr_mode, [&](double c) { return gaunt_factor(tb, c); });
        double g = sqrt(1.0 + nsq(u));
        ut = g;
        double hg = half_cdt / g;
        r = {rh.t + half_cdt, rh.x + hg * u.x, rh.y + hg * u.y, rh.z + hg * u.z};
        // Field::radiate at the new position
        lcfa_fields<KIND, CP>(L, r, E, B);
        d3 beta = (1.0 / g) * u;
        d3 vxb = C_LIGHT * cross(beta, B);
        double eb = dot(E, beta);
        double E_rf_sqd = nsq(E + vxb) - eb * eb;
        double chi = E_rf_sqd > 0.0 ? g * sqrt(E_rf_sqd) / E_CRIT : 0.0;
        double p = A->classical_mode ? dt * lcfa_classical_rate(chi, g) : dt * lcfa_rate(tb, A->sc, chi, g);
        S *= (1.0 - p);
        if (S < U) {
            EmitOut o = lcfa_emission(A, gt, chi, g, beta.x, beta.y, beta.z, E.x, E.y, E.z, vxb.x, vxb.y, vxb.z, r.t, r.x,
                                      r.y, r.z, idx, nev, rng);
            rng.c0 = o.c0;
            rng.c1 = o.c1;
            if (o.ok) {
                ++n_emit;
                if (A->recoil) {
                    u = u - d3{o.kx, o.ky, o.kz};
                    ut = ut - o.kt;
                }
                icount += 1.0;
            }
            ++nev;
            U = rng_uniform(rng);
            S = 1.0;
        }
    }
    wave_add(A->stats + STAT_LEPTON_STEPS, n_steps);
    wave_add(A->stats + STAT_EMISSIONS, n_emit);
    if (flags) atomicOr(A->flags, flags);
}

// Pair-creation event (field/lcf.rs:164-190): samples the positron, writes the
// pair to the secondary queue. Returns the decayed fraction.
struct PairOut { uint32_t c0, c1, trials; int capped; };

__device__ __noinline__ PairOut lcfa_pair_event(const TransportArgs* __restrict__ A, const double* tb, double chi,
                                                double kt, double kx, double ky, double kz, double ax, double ay, double az,
                                                double ex, double ey, double ez, sv4 pin, double rt, double rx, double ry,
                                                double rz, double weight_frac, uint32_t idx, uint32_t nev, Rng rng) {
    double gamma_p, cos_theta, cphi;
    uint32_t tr = 0;
    bool capped = false;
    d3 k3{kx, ky, kz}, a_perp{ax, ay, az};
    pair_sample(tb, tb + lcfa_pack::PAIR_QN, kt, k3, pin, chi, a_perp, rng, gamma_p, cos_theta, cphi, tr, capped);
    d3 n = unit(k3);
    double sin_theta = sqrt(1.0 - cos_theta * cos_theta);
    double upm = gamma_p * sqrt(1.0 - 1.0 / (gamma_p * gamma_p));
    d3 e1 = unit(a_perp);
    d3 e2 = cross(n, e1);
    double sph, cph;
    sincos(cphi, &sph, &cph);
    d3 up = upm * (cos_theta * n + (sin_theta * cph) * e1 + (sin_theta * sph) * e2);
    d3 ue = k3 - up;
    double Emag = sqrt(ex * ex + ey * ey + ez * ez);
    uint32_t flags = 0;
    emit_pair(*A, d4{rt, rx, ry, rz}, ue, up, weight_frac, A->laser.a_norm * Emag, chi, A->in.id[idx],
              rng_child(A->in.stream[idx], nev), A->in.shower[idx], flags);
    return PairOut{rng.c0, rng.c1, tr, capped ? 1 : 0};
}

// Photon transport (src/main.rs:212-264 with field/lcf.rs:144-193 and
// lcfa/pair_creation/mod.rs).
template <int KIND, bool CP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PTG_WAVES_PER_EU))) void k_photons_lcfa(const TransportArgs* __restrict__ A) {
    __shared__ double tb[lcfa_pack::HOT_SIZE];  // per-step tables; event tables are read from global memory
    for (int i = threadIdx.x; i < lcfa_pack::HOT_SIZE; i += blockDim.x) tb[i] = A->lcfa_tables[i];
    __syncthreads();
    const double* gt = A->lcfa_tables;
    const LaserDev& L = A->laser;
    const double dt = A->dt;
    const double cdt = C_LIGHT * dt;

    uint64_t n_steps = 0, n_pairs = 0, trials = 0, ff_steps = 0;
    uint32_t flags = 0, psteps = 0;

    uint32_t idx = 0;
    d4 r, k;
    sv4 pol;
    double icount = 0.0, weight = 1.0, S = 1.0, U = 0.0;
    uint32_t nev = 0, meta = 0;
    Rng rng{};
    bool decayed = false, skip = false;

    auto fetch = [&]() -> bool {
        for (;;) {
            uint32_t i = wave_reserve32(A->work_ctr, 1u);
            if (i >= A->n_in) return false;
            const PBuf& in = A->in;
            meta = in.meta[i];
            if ((meta & 3u) != 2u) { forward_record(*A, i); continue; }
            idx = i;
            r = {in.d[P_RT][i], in.d[P_RX][i], in.d[P_RY][i], in.d[P_RZ][i]};
            k = {in.d[P_UT][i], in.d[P_UX][i], in.d[P_UY][i], in.d[P_UZ][i]};
            pol = {in.d[P_SI][i], in.d[P_SQ][i], in.d[P_SU][i], in.d[P_SV][i]};
            icount = in.d[P_ICOUNT][i];
            weight = in.d[P_WEIGHT][i];
            nev = in.nev[i];
            rng = rng_make(A->key, in.stream[i]);
            U = rng_uniform(rng);
            S = 1.0;
            decayed = false;
            skip = !A->opt.tracking_photons || (!A->opt.tracking_secondaries && !(meta & META_PRIMARY));
            psteps = 0;
            return true;
        }
    };
    bool live = fetch();
    while (live) {
        bool inside = !skip && !decayed && field_contains(L, r) && r.t / C_LIGHT < A->opt.t_stop;
        if (inside && psteps >= A->step_cap) {
            flags |= PTG_FLAG_STEP_CAP;
            inside = false;
        }
        if (!inside) {
            finalize_photon(*A, idx, r, k, pol, weight, icount, meta, decayed);
            live = fetch();
            continue;
        }
        double ik = cdt / k.t;
        if (psteps == A->ff_after && k.t < A->ff_kt_max) {
            // zero pair probability for the rest of the path: advance to two steps
            // before the exit (phase or t_stop) and let the loop take the last ones
            double dphi = L.k0 * (cdt - ik * k.z);
            double n_ph = dphi > 0.0 ? ceil((L.max_phase - L.k0 * (r.t - r.z)) / dphi) : 1.0e300;
            double n_t = isfinite(A->opt.t_stop) ? ceil((A->opt.t_stop * C_LIGHT - r.t) / cdt) : 1.0e300;
            double n_go = fmin(fmin(n_ph, n_t), double(A->step_cap - psteps)) - 2.0;
            if (n_go > 1.0) {
                r = {r.t + n_go * cdt, r.x + n_go * (ik * k.x), r.y + n_go * (ik * k.y), r.z + n_go * (ik * k.z)};
                psteps += uint32_t(n_go);
                ff_steps += uint64_t(n_go);
                continue;
            }
        }
        ++n_steps;
        ++psteps;
        r = {r.t + cdt, r.x + ik * k.x, r.y + ik * k.y, r.z + ik * k.z};
        sv4 pin = A->opt.pol_resolved ? pol : sv4{1.0, 0.0, 0.0, 0.0};
        d3 E, B;
        lcfa_fields<KIND, CP>(L, r, E, B);
        d3 k3 = sp(k);
        d3 n = unit(k3);
        d3 a_perp = E - dot(E, n) * n + C_LIGHT * cross(n, B);
        double E_rf_sqd = nsq(a_perp);
        double chi = 0.0, prob = 0.0;
        sv4 pnew = pin;
        if (E_rf_sqd > 0.0) {
            chi = k.t * sqrt(E_rf_sqd) / E_CRIT;
            prob = pair_probability(tb, A->sc, k.t, k3, pin, chi, a_perp, dt, pnew);
        }
        double ri = A->opt.rate_increase;
        if (prob * ri > 0.1) ri = 0.1 / prob;
        S *= (1.0 - prob * ri);
        if (S < U) {
            double frac = 1.0 / ri;
            PairOut o = lcfa_pair_event(A, gt, chi, k.t, k.x, k.y, k.z, a_perp.x, a_perp.y, a_perp.z, E.x, E.y, E.z, pin, r.t,
                                        r.x, r.y, r.z, frac * weight, idx, nev, rng);
            rng.c0 = o.c0;
            rng.c1 = o.c1;
            trials += o.trials;
            if (o.capped) flags |= PTG_FLAG_REJECT_CAP;
            ++n_pairs;
            if (frac == 1.0) decayed = true;
            else weight = weight * (1.0 - frac);
            ++nev;
            U = rng_uniform(rng);
            S = 1.0;
        }
        if (A->opt.rotate_stokes_pars && A->opt.pol_resolved) pol = pnew;
        icount += prob;
    }
    wave_add(A->stats + STAT_PHOTON_STEPS, n_steps);
    wave_add(A->stats + STAT_PHOTON_STEPS_CF, ff_steps);
    wave_add(A->stats + STAT_PAIRS, n_pairs);
    wave_add(A->stats + STAT_TRIALS, trials);
    if (flags) atomicOr(A->flags, flags);
}

void launch_leptons_lcfa(const TransportArgs* a, int kind, bool cp, int grid, hipStream_t s) {
    if (kind == K_FAST_FOCUSED_LASER) {
        if (cp) k_leptons_lcfa<K_FAST_FOCUSED_LASER, true><<<grid, 256, 0, s>>>(a);
        else k_leptons_lcfa<K_FAST_FOCUSED_LASER, false><<<grid, 256, 0, s>>>(a);
    } else {
        k_leptons_lcfa<K_FAST_PLANE_WAVE, false><<<grid, 256, 0, s>>>(a);
    }
}
// Resident 256-lane workgroups per CU for the transport kernel that
// launch_leptons_lcfa / launch_photons_lcfa would run (the persistent grid size).
int lcfa_blocks_per_cu(int kind, bool cp, bool photons) {
    int nb = 0;
    hipError_t e;
    if (kind == K_FAST_FOCUSED_LASER && cp)
        e = photons ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_photons_lcfa<K_FAST_FOCUSED_LASER, true>, 256, 0)
                    : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_leptons_lcfa<K_FAST_FOCUSED_LASER, true>, 256, 0);
    else if (kind == K_FAST_FOCUSED_LASER)
        e = photons ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_photons_lcfa<K_FAST_FOCUSED_LASER, false>, 256, 0)
                    : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_leptons_lcfa<K_FAST_FOCUSED_LASER, false>, 256, 0);
    else
        e = photons ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_photons_lcfa<K_FAST_PLANE_WAVE, false>, 256, 0)
                    : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_leptons_lcfa<K_FAST_PLANE_WAVE, false>, 256, 0);
    return (e == hipSuccess && nb > 0) ? nb : 1;
}

void launch_photons_lcfa(const TransportArgs* a, int kind, bool cp, int grid, hipStream_t s) {
    if (kind == K_FAST_FOCUSED_LASER) {
        if (cp) k_photons_lcfa<K_FAST_FOCUSED_LASER, true><<<grid, 256, 0, s>>>(a);
        else k_photons_lcfa<K_FAST_FOCUSED_LASER, false><<<grid, 256, 0, s>>>(a);
    } else {
        k_photons_lcfa<K_FAST_PLANE_WAVE, false><<<grid, 256, 0, s>>>(a);
    }
}

// Field-only push (test support for Field::push): n particles, at most n_steps each.
template <int KIND, bool CP>
__global__ __launch_bounds__(256) void k_push_lcfa(LaserDev L, double dt, int rr_mode, const uint8_t* species, double* r4, double* u4, uint32_t n,
                            uint32_t n_steps, uint32_t* taken, double* work) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    d4 r{r4[4 * i], r4[4 * i + 1], r4[4 * i + 2], r4[4 * i + 3]};
    d3 u{u4[4 * i + 1], u4[4 * i + 2], u4[4 * i + 3]};
    double ut = u4[4 * i];
    double rqm = species[i] == 1 ? -RQM_ELECTRON : (species[i] == 0 ? RQM_ELECTRON : 0.0);
    const double half_cdt = 0.5 * C_LIGHT * dt;
    uint32_t k = 0;
    double w = 0.0;
    for (; k < n_steps && field_contains(L, r); ++k) {
        d4 rh{r.t + half_cdt, r.x + half_cdt * u.x / ut, r.y + half_cdt * u.y / ut, r.z + half_cdt * u.z / ut};
        d3 E, B;
        lcfa_fields<KIND, CP>(L, rh, E, B);
        w += vay_push(u, E, B, rqm, dt, rr_mode, [](double) { return 1.0; });
        ut = sqrt(1.0 + nsq(u));
        double hg = half_cdt / ut;
        r = {rh.t + half_cdt, rh.x + hg * u.x, rh.y + hg * u.y, rh.z + hg * u.z};
    }
    r4[4 * i] = r.t; r4[4 * i + 1] = r.x; r4[4 * i + 2] = r.y; r4[4 * i + 3] = r.z;
    u4[4 * i] = ut; u4[4 * i + 1] = u.x; u4[4 * i + 2] = u.y; u4[4 * i + 3] = u.z;
    taken[i] = k;
    if (work) work[i] = w;
}

void launch_push_lcfa(const LaserDev& L, double dt, int rr_mode, const uint8_t* sp_, double* r4, double* u4, uint32_t n,
                      uint32_t n_steps, uint32_t* taken, double* work, hipStream_t s) {
    int grid = int((n + 255) / 256);
    if (L.kind == K_FAST_FOCUSED_LASER) {
        if (L.circular) k_push_lcfa<K_FAST_FOCUSED_LASER, true><<<grid, 256, 0, s>>>(L, dt, rr_mode, sp_, r4, u4, n, n_steps, taken, work);
        else k_push_lcfa<K_FAST_FOCUSED_LASER, false><<<grid, 256, 0, s>>>(L, dt, rr_mode, sp_, r4, u4, n, n_steps, taken, work);
    } else {
        k_push_lcfa<K_FAST_PLANE_WAVE, false><<<grid, 256, 0, s>>>(L, dt, rr_mode, sp_, r4, u4, n, n_steps, taken, work);
    }
}

}  // namespace ptg
