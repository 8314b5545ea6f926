// Transport kernels for the LMA fields (PlaneWave, FocusedLaser): collide()'s
// electron/positron loop (src/main.rs:169-210) with the step and the emission
// event split into two kernels.
//
//  * k_step_lma: persistent lanes advance electrons with the ponderomotive
//    push and accumulate the survival S *= 1 - P_k (P from the LP rate table).
//    When S < U the lane writes the particle, the local wavevector kappa and
//    <a^2> at the event into the event queue and pulls the next particle:
//    heavy, divergent event sampling never stalls a wave of stepping lanes.
//  * k_event_lma: one lane per queued event samples the harmonic, (s, theta)
//    by rejection on J(n, x, y), builds the photon in the
//    zero-momentum frame, writes it out and re-queues the recoiled electron.
// The host alternates the two until no electron is left in the field.
#include "kern_lma.h"

namespace ptg {

// RNG counter blocks: stepping of segment k (after k events) uses counter
// (c0, c1 = k); sampling event k uses (c0, c1 = k | EVENT_BLOCK).
constexpr uint32_t EVENT_BLOCK = 0x80000000u;

// Lanes of the persistent step kernels: stepping, holding a particle that left the
// field or reached its event (record not yet written), without a particle, or
// with no input left.
enum : uint32_t { LANE_ACTIVE = 0, LANE_EXIT = 1, LANE_EVENT = 2, LANE_EMPTY = 3, LANE_DONE = 4 };
// waiting lanes that trigger the wave's write-out and refill (8, 16, 32 measured on the
// a0 = 2 plane wave, the learned-sampler domain and the scan: 32 steps fastest)
#ifndef PTG_LMA_REFILL
#define PTG_LMA_REFILL 32
#endif

// rate table staged in LDS per emission model: LP [68][61], CP [70][61]; the
// classical tables (61 values) are read through the cache
template <int MODE>
struct StepTable {
    static constexpr int N = MODE == LMA_LP ? lma_pack::RATE_COLS * lma_pack::RATE_ROWS
                                            : (MODE == LMA_CP ? cp_pack::COLS * cp_pack::ROWS : 1);
    __device__ static const double* src(const LmaArgs& M) { return MODE == LMA_LP ? M.tables + lma_pack::RATE : M.cp + cp_pack::RATE; }
};

// MIX: mixed precision (LP and CP quantum modes): the rate table is staged in
// LDS as FP32 and looked up in FP32, and the plane-wave pusher algebra runs in
// FP32; the phase, the position and momentum updates, the mass shell and the
// survival product stay FP64.
template <int MODE, bool MIX>
__global__ __launch_bounds__(256) void k_step_lma(const TransportArgs* __restrict__ A, LmaArgs M) {
    using TT = typename std::conditional<MIX, float, double>::type;
    // The launch's laser, read once through the uniform argument pointer (its fields
    // stay in scalar registers); the lasers of a batch differ only in a0 (ptg_set_batch),
    // which each particle takes from its own collision when it is fetched. Reading the
    // whole laser through a per-particle pointer instead cost ~25 vector loads per
    // step, and the waves spent most of their time waiting on them.
    LaserDev L = A->laser;
    const double dt = A->dt;
    const double kc = C_LIGHT * T_COMPTON;
    __shared__ TT rt[StepTable<MODE>::N];
    if (A->n_in == 0u) return;  // a captured generation without input
    if (MODE == LMA_LP || MODE == LMA_CP) {
        const double* src = StepTable<MODE>::src(M);
        for (int i = threadIdx.x; i < StepTable<MODE>::N; i += blockDim.x) rt[i] = TT(src[i]);
        __syncthreads();
    }

    uint64_t n_steps = 0;
    uint32_t flags = 0, psteps = 0;
    uint32_t idx = 0, meta = 0, nev = 0;
    d4 r, u;
    double work = 0.0, S = 1.0, U = 0.0, rqm = RQM_ELECTRON;
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
            if (M.coll) L.a0 = M.lasers[M.coll[in.shower[i]]].a0;
            r = {in.d[P_RT][i], in.d[P_RX][i], in.d[P_RY][i], in.d[P_RZ][i]};
            u = {in.d[P_UT][i], in.d[P_UX][i], in.d[P_UY][i], in.d[P_UZ][i]};
            work = in.d[P_WORK][i];
            nev = in.nev[i];
            rqm = (meta & 3u) == 1u ? -RQM_ELECTRON : RQM_ELECTRON;
            rng = rng_make(A->key, in.stream[i]);
            rng.c1 = nev;
            U = rng_uniform(rng);
            S = 1.0;
            skip = !A->opt.tracking_secondaries && !(meta & META_PRIMARY);
            psteps = 0;
            return true;
        }
    };
    // A lane whose particle leaves the field or reaches its event holds it (its final
    // state or event record unwritten) while the other lanes step on; once a quarter
    // of the wave waits, or no lane steps, the waiting lanes write their records and
    // take new particles together, so the atomics and record loads of a refill are
    // paid once per refill rather than in nearly every step (a particle's stretch
    // between events is ~10^2 steps, so some lane of 64 ends one in most steps).
    uint32_t st = LANE_EMPTY;
    double k0_ev = 0.0;
    for (;;) {
        if (st == LANE_ACTIVE) {
            bool inside = !skip && field_contains(L, r) && r.t / C_LIGHT < A->opt.t_stop;
            if (inside && psteps >= A->step_cap) { flags |= PTG_FLAG_STEP_CAP; inside = false; }
            if (!inside) {
                st = LANE_EXIT;
            } else {
                ++n_steps;
                ++psteps;
                if (MIX && L.kind == K_PLANE_WAVE) work += pw_push_mixed(L, r, u, rqm, dt, A->rr_mode == 1);
                else work += lma_push(L, r, u, rqm, dt, A->rr_mode == 1);
                // PlaneWave::radiate (plane_wave.rs:251-282)
                double phase = L.k0 * (r.t - r.z);
                double chirp = 1.0 + 2.0 * L.chirp_b * phase;
                double width = 1.0;
                if (M.bandwidth != 0.0) {
                    double u1, u2;
                    rng_uniform2(rng, u1, u2);
                    double sn, cs;
                    sincos(2.0 * PI_D * u2, &sn, &cs);
                    width = 1.0 + M.bandwidth * sqrt(-2.0 * log(1.0 - u1)) * cs;
                }
                double k0 = kc * L.k0 * chirp * width;  // kappa = k0 (1, 0, 0, 1)
                double a_sqd_q = u.t * u.t - u.x * u.x - u.y * u.y - u.z * u.z - 1.0;
                double a_rms = a_sqd_q >= 0.0 ? sqrt(a_sqd_q) : 0.0;
                double eta = k0 * (u.t - u.z);
                double dphi = dt * eta / (T_COMPTON * u.t);
                double f;
                double prob = 0.0;
                bool ok;
                if constexpr (MIX && (MODE == LMA_LP || MODE == LMA_CP)) {
                    float ff = 0.0f;
                    if (MODE == LMA_LP) ok = lp_rate_f(rt, M.sc, a_rms * 1.4142135623730951, eta, ff);
                    else ok = cp_rate_f(rt, M.cp, M.csc, a_rms, eta, ff);
                    if (ok) prob = double(float(ALPHA) * ff * (float(dphi) / float(eta)));
                } else {
                    if (MODE == LMA_LP) ok = lp_rate(rt, M.sc, a_rms * 1.4142135623730951, eta, f);
                    else if (MODE == LMA_CP) ok = cp_rate(rt, M.cp, M.csc, a_rms, eta, f);
                    else if (MODE == LMA_CP_CLASSICAL) ok = cp_classical_rate(M.cp, M.csc, a_rms, eta, f);
                    else ok = lp_cl_rate(M.tables, M.sc, a_rms * 1.4142135623730951, eta, f);
                    if (ok) prob = ALPHA * f * dphi / eta;
                }
                if (!ok) flags |= PTG_FLAG_TABLE_RANGE;
                S *= (1.0 - prob);
                if (S < U) {
                    st = LANE_EVENT;
                    k0_ev = k0;
                }
            }
        }
        const uint64_t waiting = __ballot(st == LANE_EXIT || st == LANE_EVENT || st == LANE_EMPTY);
        const uint64_t active = __ballot(st == LANE_ACTIVE);
        if (!waiting) {
            if (!active) break;
            continue;
        }
        if (active && __popcll(waiting) < PTG_LMA_REFILL) continue;
        if (st == LANE_EXIT) {
            double icount = A->in.d[P_ICOUNT][idx];
            finalize_lepton(*A, idx, r, u.t, sp(u), work, icount, meta);
            st = LANE_EMPTY;
        } else if (st == LANE_EVENT) {
            // suspend: queue the particle with the event parameters
            unsigned long long j = wave_reserve(M.ev_ctr, 1u);
            if (j < M.ev_cap) {
                const PBuf& q = M.ev;
                copy_record(q, j, A->in, idx);
                q.d[P_RT][j] = r.t; q.d[P_RX][j] = r.x; q.d[P_RY][j] = r.y; q.d[P_RZ][j] = r.z;
                q.d[P_UT][j] = u.t; q.d[P_UX][j] = u.x; q.d[P_UY][j] = u.y; q.d[P_UZ][j] = u.z;
                q.d[P_WORK][j] = work;
                q.d[P_ICOUNT][j] = A->in.d[P_ICOUNT][idx];
                q.nev[j] = nev;
                M.ev_k0[j] = k0_ev;
                M.ev_a[j] = sqrt(lma_a_sqd(L, r));
            }
            st = LANE_EMPTY;
        }
        if (st == LANE_EMPTY) st = fetch() ? LANE_ACTIVE : LANE_DONE;
    }
    wave_add(A->stats + STAT_LEPTON_STEPS, n_steps);
    if (flags) atomicOr(A->flags, flags);
}

// The rest of an emission once (n, s, theta, Stokes) are sampled (nonlinear_compton/mod.rs:73-136,
// src/main.rs:172-200): the photon built in the zero-momentum frame and boosted back,
// written out (final state or photon queue), and the electron, recoiled or not,
// re-queued with its next RNG segment. `ok` false: no photon, the electron continues.
template <bool classical>
__device__ inline void lma_emit_finish(const TransportArgs* __restrict__ A, const LmaArgs& M, uint32_t i, double k0, double a, double eta,
                                       bool ok, int n, double s, double cphi, sv4 sv, uint64_t& n_emit, uint32_t& flags) {
    const PBuf& q = M.ev_in;
    d4 r{q.d[P_RT][i], q.d[P_RX][i], q.d[P_RY][i], q.d[P_RZ][i]};
    d4 u{q.d[P_UT][i], q.d[P_UX][i], q.d[P_UY][i], q.d[P_UZ][i]};
    double a_field = M.ev_a[i];
    uint32_t nev = q.nev[i];
    d4 u_new = u;
    double work = q.d[P_WORK][i], icount = q.d[P_ICOUNT][i];
    if (ok) {
        // photon momentum in the zero-momentum frame, boosted back (mod.rs:73-116)
        double ell = double(n);
        double p_zmf = classical ? ell * eta / sqrt(1.0 + a * a) : ell * eta / sqrt(1.0 + a * a + 2.0 * ell * eta);
        double cos_theta_zmf = classical ? 1.0 - s * (1.0 + a * a) / (ell * eta)
                                         : 1.0 - s * (1.0 + a * a + 2.0 * ell * eta) / (ell * eta);
        // the ZMF moves with q (classical) or q + n kappa (quantum); the photon's axis
        // is along n kappa or against q there
        d4 qk = classical ? u : d4{u.t + ell * k0, u.x, u.y, u.z + ell * k0};
        double inv = 1.0 / sqrt(mdot(qk, qk));
        d4 u_zmf{qk.t * inv, qk.x * inv, qk.y * inv, qk.z * inv};
        d3 along = classical ? unit(sp(boost4(d4{ell * k0, 0.0, 0.0, ell * k0}, u_zmf))) : -unit(sp(boost4(u, u_zmf)));
        d3 eps = unit(sp(boost4(d4{0.0, M.cos_pa, M.sin_pa, 0.0}, u_zmf)));
        eps = unit(cross(along, cross(eps, along)));
        double sp_, cp_;
        sincos(cphi, &sp_, &cp_);
        d3 perp = rotate(eps, along, sp_, cp_);
        d3 kp = p_zmf * (cos_theta_zmf * along + sqrt(1.0 - cos_theta_zmf * cos_theta_zmf) * perp);
        d4 kz{sqrt(nsq(kp)), kp.x, kp.y, kp.z};
        d4 k = boost4(kz, d4{u_zmf.t, -u_zmf.x, -u_zmf.y, -u_zmf.z});
        sv = stokes_from_lma(stokes_rotate(sv, -M.pol_angle), k);
        // sanity check of the construction (mod.rs:118-135)
        double s_new = (k0 * (k.t - k.z)) / eta;
        if (!(fabs(s - s_new) / s < 1.0e-2)) flags |= PTG_FLAG_SANITY;
        double chi = a_field * eta;
        // photon (src/main.rs:172-189)
        const uint32_t shower = q.shower[i];
        uint64_t id = A->id_base + wave_reserve(A->id_ctr, 1u);
        uint64_t pstream = rng_child(q.stream[i], nev);
        if (M.photons_to_out) {
            if (passes_cut(*A, shower, M_E_MEV * k.t, sp(k))) {
                uint64_t j = wave_reserve(A->out_ctr, 1u);
                write_new(A->out, j, 2u, r, k, d4{0.0, 0.0, 0.0, 0.0}, sv, a_field, chi, q.d[P_WEIGHT][i], id, q.id[i],
                          pstream, shower);
            }
        } else {
            uint64_t j = wave_reserve(M.ph_ctr, 1u);
            if (j < M.ph_cap)
                write_new(M.ph, j, 2u, r, k, d4{0.0, 0.0, 0.0, 0.0}, sv, a_field, chi, q.d[P_WEIGHT][i], id, q.id[i],
                          pstream, shower);
        }
        ++n_emit;
        if (A->recoil) u_new = {u.t + ell * k0 - k.t, u.x - k.x, u.y - k.y, u.z + ell * k0 - k.z};
        icount += 1.0;
        if (!classical) work += ell * k0;  // absorption (src/main.rs:198-200)
    }
    // the electron continues with its next RNG segment
    unsigned long long j = wave_reserve(M.next_ctr, 1u);
    const PBuf& o = M.next;
    copy_record(o, j, q, i);
    o.d[P_RT][j] = r.t; o.d[P_RX][j] = r.x; o.d[P_RY][j] = r.y; o.d[P_RZ][j] = r.z;
    o.d[P_UT][j] = u_new.t; o.d[P_UX][j] = u_new.x; o.d[P_UY][j] = u_new.y; o.d[P_UZ][j] = u_new.z;
    o.d[P_WORK][j] = work;
    o.d[P_ICOUNT][j] = icount;
    o.nev[j] = nev + 1;
}

__device__ __forceinline__ double wave_fmax(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// (s, theta) and the Stokes parameters of a CP emission at harmonic n
// (cp/mod.rs:181-234, cp/classical.rs:128-178), one emission per wave: all 64
// lanes call this with the same arguments and return the same values. The
// ceiling's points (32, or the classical 1 or 8) take lanes [0, m) while lanes
// [m, 64) take the first candidates, then rounds of 64; the first accepted
// candidate in candidate order is the sample. Candidate t draws (v, u) from
// counter 16 + t of `base`, counter 1 gives theta. Every lane of a wave works on
// the same harmonic, where one lane per emission would wait for the longest
// Bessel recurrence among 64 different harmonics.
template <bool classical>
__device__ __noinline__ void cp_sample_wave(TP tb, const double* P, int n, double a, double eta, const Rng& base, double& s,
                                            double& theta, sv4& sv, uint64_t& trials, uint32_t& flags) {
    const int lane = threadIdx.x;
    const uint32_t cap = 1u << 20;
    const int m = classical ? cp_classical_ceiling_count(n, a) : 32;
    auto rate = [&](double v) { return classical ? cp_spectrum_low_eta(tb, n, a, v) : cp_ddpr(tb, n, a, eta, v); };
    auto candidate = [&](uint32_t t, double& v, double& u) {
        Rng g = base;
        g.c0 = 16u + t;
        rng_uniform2(g, v, u);
    };
    double f = -INFINITY, v = 0.0, u = 0.0, fc = 0.0;
    if (lane < m) {
        f = rate(classical ? cp_classical_ceiling_arg(n, a, lane) : cp_ceiling_arg(P + cp_pack::G16N, n, a, eta, lane));
    } else {
        candidate(uint32_t(lane - m), v, u);
        fc = rate(v);
    }
    const double mx = (classical ? 1.1 : 1.2) * wave_fmax(f);
    int first = -1;
    double v_acc = 0.0;
    uint32_t t0 = uint32_t(64 - m);  // candidates consumed
    {
        const uint64_t mask = __ballot(lane >= m && u <= fc / mx);
        if (mask) {
            first = __ffsll((long long)mask) - 1;
            v_acc = __shfl(v, first, 64);
            trials += uint32_t(first - m) + 1;
        }
    }
    while (first < 0) {
        const uint32_t t = t0 + uint32_t(lane);
        bool acc = false;
        if (t < cap) {
            candidate(t, v, u);
            acc = u <= rate(v) / mx;
        }
        const uint64_t mask = __ballot(acc);
        if (mask) {
            first = __ffsll((long long)mask) - 1;
            v_acc = __shfl(v, first, 64);
            trials += t0 + uint32_t(first) + 1;
        } else {
            t0 += 64;
            if (t0 >= cap) {  // as the sequential sampler: give up and keep candidate `cap`
                flags |= PTG_FLAG_REJECT_CAP;
                candidate(cap, v_acc, u);
                trials += cap;
                first = 0;
            }
        }
    }
    const double sn = 2.0 * double(n) * eta / (1.0 + a * a);
    s = classical ? v_acc * sn : v_acc * (sn / (1.0 + sn));
    Rng g = base;
    g.c0 = 1u;
    theta = 2.0 * PI_D * rng_uniform(g);
    double st[4];
    cp_stokes(tb, n, a, s, sn, v_acc, theta, classical, st);
    sv = {st[0], st[1], st[2], st[3]};
}

// (ln n, ln(-ln(1 - F))) of the first 31 points of every CP CDF row: one thread per point.
__global__ void k_cp_cdf_logs(double* P) {
    using namespace cp_pack;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int row = t / 31, k = t % 31;
    if (row < CDF_N) cp_cdf_log_point(P + CDF + 64 * row, k, P + CDF_LN + 62 * row + 2 * k);
    else if (row < CDF_N + CL_N)
        cp_cdf_log_point(P + CL_CDF + 64 * (row - CDF_N), k, P + CL_CDF_LN + 62 * (row - CDF_N) + 2 * k);
}
void launch_cp_cdf_logs(double* cp, hipStream_t s) {
    const int n = (cp_pack::CDF_N + cp_pack::CL_N) * 31;
    k_cp_cdf_logs<<<(n + 255) / 256, 256, 0, s>>>(cp);
}

// (ln n, ln(-ln(1 - F))) of the first 15 points of every LP CDF row: one thread per point.
__global__ void k_lp_cdf_logs(double* P) {
    using namespace lma_pack;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int row = t / 15, k = t % 15;
    if (row < CDF_N) lp_cdf_log_point(P + CDF + 32 * row, k, P + CDF_LN + 30 * row + 2 * k);
}
void launch_lp_cdf_logs(double* lma, hipStream_t s) {
    const int n = lma_pack::CDF_N * 15;
    k_lp_cdf_logs<<<(n + 255) / 256, 256, 0, s>>>(lma);
}

// One CP emission per wave (nonlinear_compton::generate, mod.rs:58-138, with
// cp_sample_wave); lane 0 builds the photon.
template <int MODE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void k_event_lma(const TransportArgs* __restrict__ A,
                                                                                           LmaArgs M) {
    constexpr bool classical = MODE == LMA_CP_CLASSICAL;
    const int lane = threadIdx.x;
    uint64_t n_emit = 0, trials = 0;
    uint32_t flags = 0;
    const uint64_t n_ev = M.n_ev_dev ? *M.n_ev_dev : M.n_ev;
    for (uint64_t e = blockIdx.x; e < n_ev; e += gridDim.x) {
        const PBuf& q = M.ev_in;
        const uint32_t i = M.ev_list ? M.ev_list[e] : uint32_t(e);
        d4 u{q.d[P_UT][i], q.d[P_UX][i], q.d[P_UY][i], q.d[P_UZ][i]};
        const double k0 = M.ev_k0[i];
        Rng base = rng_make(A->key, q.stream[i]);
        base.c1 = q.nev[i] | EVENT_BLOCK;
        const double a = sqrt(u.t * u.t - u.x * u.x - u.y * u.y - u.z * u.z - 1.0);  // rms a
        const double eta = k0 * (u.t - u.z);
        Rng g = base;  // counter 0: the harmonic
        const double fr = rng_uniform(g);
        const int nh = classical ? cp_classical_harmonic(M.cp, M.csc, a, fr) : cp_harmonic_wave(M.cp, M.csc, a, eta, fr);
        const bool ok = nh >= 1;
        if (!ok) flags |= PTG_FLAG_TABLE_RANGE;
        double s = 0.0, cphi = 0.0;
        sv4 sv{1.0, 0.0, 0.0, 0.0};
        if (ok) cp_sample_wave<classical>(M.lcfa, M.cp, nh, a, eta, base, s, cphi, sv, trials, flags);
        if (lane == 0) lma_emit_finish<classical>(A, M, i, k0, a, eta, ok, ok ? nh : 0, s, cphi, sv, n_emit, flags);
    }
    if (lane == 0) {
        wave_add(A->stats + STAT_EMISSIONS, n_emit);
        wave_add(A->stats + STAT_TRIALS, trials);
        if (flags) atomicOr(A->flags, flags);
    }
}

// One emission per wave, for LP emission (quantum or classical), where every
// trial costs a J(n, x, y) evaluation: the 64 lanes evaluate the ceiling's
// sample points together (lp/mod.rs:53-140, lp/classical.rs:43-110), then
// draw (s, theta) candidates in rounds of 64 (one per lane) and take the first
// accepted one in candidate order. Candidates are independent draws (candidate
// t uses counters 16 + 2t, 17 + 2t of the event's RNG block), so the first
// accepted one has the distribution of sequential rejection sampling; an
// event's draws do not depend on how many waves run or on their order.
// From harmonic PTG_LMA_SHARED_N up, where the quadrature of J(n, x, y) has
// hundreds to tens of thousands of nodes, the wave instead shares every J
// evaluation (dbessel_quadrature_wave) and takes the ceiling's points and the
// candidates one after another: the same candidates and the same accepted one
// (up to the rounding of the partial sums), without the work of the rounds'
// candidates beyond it and at a fraction of the latency of one lane's J.
#ifndef PTG_LMA_WAVE_WPE
#define PTG_LMA_WAVE_WPE 2  // two waves per SIMD: 256 VGPRs, a little scratch
#endif
#ifndef PTG_LMA_SHARED_N
#define PTG_LMA_SHARED_N 200
#endif

// (s, theta) and the Stokes parameters of an LP emission at harmonic nh, for
// one emission per wave: all 64 lanes call this with the same arguments (a
// block is one wave) and return the same values. Counter 1 of `base` gives the
// quadrant of theta, candidate t draws (v, theta) from counter 16 + 2t and its
// acceptance from 17 + 2t.
template <bool classical>
__device__ __noinline__ void lp_sample_wave(int nh, double alp, double eta, const Rng& base, double* tb_s, double* tb_f,
                                            const double* ceil_tab, double& s, double& theta, sv4& sv, uint64_t& trials,
                                            uint32_t& flags, uint64_t& restarts) {
    const int lane = threadIdx.x;
    const uint32_t cap = 1u << 20;
    __syncthreads();  // the previous event's bound is no longer read
    const DBessel D = dbessel_of(nh);
    const double sn = 2.0 * double(nh) * eta / (1.0 + 0.5 * alp * alp);
    const double smax = classical ? sn : sn / (1.0 + sn);  // s = v * smax
    if (lane < 16) {  // the theta bound's 16 points, one per lane
        if (classical) theta_bound_low_eta_point(nh, alp, lane, tb_s[lane], tb_f[lane]);
        else theta_bound_point(nh, alp, eta, lane, tb_s[lane], tb_f[lane]);
    }
    __syncthreads();
    if (lane == 0) theta_bound_running_min(tb_f);
    __syncthreads();
    // A candidate t: (v, theta) from counter 16 + 2t, its acceptance draw from 17 + 2t.
    // f is the rate at the candidate (not needed if theta is out of bounds).
    auto candidate = [&](uint32_t t, double& v, double& th, double& uz, double& fc) -> bool {
        Rng gc = base;
        gc.c0 = 16u + 2u * t;
        rng_uniform2(gc, v, th);
        th *= PI_D / 2.0;
        if (th > theta_bound_at(tb_s, tb_f, classical ? v : smax * v)) return false;
        uz = rng_uniform(gc);
        fc = classical ? lp_cl_ddpr(D, alp, v, th) : lp_ddpr(D, alp, eta, smax * v, th);
        return true;
    };
    const int m = classical ? lp_cl_ceiling_count(nh, alp) : lp_ceiling_count(nh, alp, eta);
    int first = -1;
    double v_acc = 0.0, th_acc = 0.0;
    if (nh >= PTG_LMA_SHARED_N) {  // every lane holds the same values from here
        const DBesselWave W{nh};
        // the ceiling from the table where it covers (n, a, eta); a candidate whose
        // rate passes it (the tabulated envelope short there) restarts the event's
        // rejection from candidate 0 under the reference's own ceiling
        double tab = 0.0;
        bool from_table = !classical && ceil_tab && lp_ceiling_lookup(ceil_tab, nh, alp, eta, tab);
        auto direct = [&]() {
            double f = classical ? NAN : -INFINITY;
            for (int k = 0; k < m; ++k) {
                double sk, tk;
                if (classical) {
                    lp_cl_ceiling_arg(nh, alp, k, sk, tk);
                    f = fmax(f, lp_cl_ddpr(W, alp, sk, tk));
                } else {
                    lp_ceiling_arg(nh, alp, eta, k, sk, tk);
                    f = fmax(f, lp_ddpr(W, alp, eta, sk, tk));
                }
            }
            return (classical ? 1.1 : 1.15) * f;
        };
        double mx = from_table ? 1.15 * tab : direct();
        for (uint32_t t = 0; first < 0; ++t) {
            Rng gc = base;
            gc.c0 = 16u + 2u * t;
            double v, th;
            rng_uniform2(gc, v, th);
            th *= PI_D / 2.0;
            if (t >= cap) {  // as the sequential sampler: give up and keep candidate `cap`
                flags |= PTG_FLAG_REJECT_CAP;
                v_acc = v;
                th_acc = th;
                trials += cap;
                first = 0;
                break;
            }
            if (th > theta_bound_at(tb_s, tb_f, classical ? v : smax * v)) continue;
            const double uz = rng_uniform(gc);
            const double fc = classical ? lp_cl_ddpr(W, alp, v, th) : lp_ddpr(W, alp, eta, smax * v, th);
            if (from_table && fc > mx) {
                ++restarts;
                trials += t + 1;
                mx = direct();
                from_table = false;
                t = ~0u;  // the loop's increment starts again at candidate 0
                continue;
            }
            if (mx * uz < fc) {
                v_acc = v;
                th_acc = th;
                trials += t + 1;
                first = 0;
            }
        }
    }
    // round 0: the ceiling's sample points on lanes [0, m), the first 64 - m
    // candidates on the other lanes (their rates do not depend on the ceiling)
    double f = classical ? NAN : -INFINITY;
    double v = 0.0, th = 0.0, uz = 0.0, fc = 0.0;
    bool in_bound = false;
    for (int k = lane; first < 0 && k < m; k += 64) {
        double sk, tk;
        if (classical) {
            lp_cl_ceiling_arg(nh, alp, k, sk, tk);
            f = fmax(f, lp_cl_ddpr(D, alp, sk, tk));
        } else {
            lp_ceiling_arg(nh, alp, eta, k, sk, tk);
            f = fmax(f, lp_ddpr(D, alp, eta, sk, tk));
        }
    }
    if (first < 0 && lane >= m) in_bound = candidate(uint32_t(lane - m), v, th, uz, fc);
    const double mx = (classical ? 1.1 : 1.15) * wave_fmax(f);
    uint32_t t0 = uint32_t(m < 64 ? 64 - m : 0);  // candidates consumed
    if (first < 0) {
        const uint64_t mask = __ballot(lane >= m && in_bound && mx * uz < fc);
        if (mask) {
            first = __ffsll((long long)mask) - 1;
            v_acc = __shfl(v, first, 64);
            th_acc = __shfl(th, first, 64);
            trials += uint32_t(first - m) + 1;
        }
    }
    while (first < 0) {
        bool acc = false;
        const uint32_t t = t0 + uint32_t(lane);
        if (t < cap && candidate(t, v, th, uz, fc)) acc = mx * uz < fc;
        const uint64_t mask = __ballot(acc);
        if (mask) {
            first = __ffsll((long long)mask) - 1;
            v_acc = __shfl(v, first, 64);
            th_acc = __shfl(th, first, 64);
            trials += t0 + uint32_t(first) + 1;
        } else {
            t0 += 64;
            if (t0 >= cap) {  // as the sequential sampler: give up and keep candidate `cap`
                flags |= PTG_FLAG_REJECT_CAP;
                Rng gc = base;
                gc.c0 = 16u + 2u * cap;
                rng_uniform2(gc, v_acc, th_acc);
                th_acc *= PI_D / 2.0;
                trials += cap;
                first = 0;
            }
        }
    }
    Rng gq = base;  // counter 1: the quadrant
    gq.c0 = 1u;
    const int quadrant = int(rng_uniform(gq) * 4.0);
    theta = th_acc;
    if (quadrant == 1) theta = PI_D - theta;
    else if (quadrant == 2) theta = PI_D + theta;
    else if (quadrant == 3) theta = 2.0 * PI_D - theta;
    s = smax * v_acc;
    if (nh >= PTG_LMA_SHARED_N) {
        const DBesselWave W{nh};
        sv = classical ? lp_cl_stokes(W, alp, v_acc, theta) : lp_stokes(W, alp, eta, s, theta);
    } else {
        if (lane == 0) sv = classical ? lp_cl_stokes(D, alp, v_acc, theta) : lp_stokes(D, alp, eta, s, theta);
        sv = sv4{__shfl(sv.i, 0, 64), __shfl(sv.q, 0, 64), __shfl(sv.u, 0, 64), __shfl(sv.v, 0, 64)};
    }
}

template <int MODE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(PTG_LMA_WAVE_WPE))) void k_event_lma_wave(const TransportArgs* __restrict__ A, LmaArgs M) {
    constexpr bool classical = MODE == LMA_LP_CLASSICAL;
    const int lane = threadIdx.x;
    const double* cdf = M.tables + lma_pack::CDF;
    __shared__ double tb_s[16], tb_f[16];  // the event's theta bound, one copy per wave
    uint64_t n_emit = 0, trials = 0, restarts = 0;
    uint32_t flags = 0;
    const uint64_t n_ev = M.n_ev_dev ? *M.n_ev_dev : M.n_ev;
    for (uint64_t e = blockIdx.x; e < n_ev; e += gridDim.x) {
        const PBuf& q = M.ev_in;
        const uint32_t i = M.ev_list ? M.ev_list[e] : uint32_t(e);
        d4 u{q.d[P_UT][i], q.d[P_UX][i], q.d[P_UY][i], q.d[P_UZ][i]};
        const double k0 = M.ev_k0[i];
        Rng base = rng_make(A->key, q.stream[i]);
        base.c1 = q.nev[i] | EVENT_BLOCK;
        const double a = sqrt(u.t * u.t - u.x * u.x - u.y * u.y - u.z * u.z - 1.0);  // rms a
        const double eta = k0 * (u.t - u.z);
        const double alp = a * 1.4142135623730951;
        Rng g = base;  // counter 0: the harmonic
        const double fr = rng_uniform(g);
        // the harmonic is inverted from the FP64 CDF in mixed precision too: near F = 1
        // neighbouring FP32 entries round to the same value (and 1 - F to 0), where the
        // monotone inversion divides by zero; it is one lookup per event, not per step
        const int nh = classical ? lp_cl_harmonic(M.tables, M.sc, alp, fr) : lp_harmonic_wave(cdf, M.sc, alp, eta, fr);
        const bool ok = nh >= 1;
        if (!ok) flags |= PTG_FLAG_TABLE_RANGE;
        double s = 0.0, theta = 0.0;
        sv4 sv{1.0, 0.0, 0.0, 0.0};
        if (ok) {
            lp_sample_wave<classical>(nh, alp, eta, base, tb_s, tb_f, M.lp_ceil, s, theta, sv, trials, flags, restarts);
        }
        if (lane == 0) lma_emit_finish<classical>(A, M, i, k0, a, eta, ok, nh, s, theta, sv, n_emit, flags);
    }
    if (lane == 0) {
        wave_add(A->stats + STAT_EMISSIONS, n_emit);
        wave_add(A->stats + STAT_TRIALS, trials);
        wave_add(A->stats + STAT_CEIL_RESTART, restarts);
        if (flags) atomicOr(A->flags, flags);
    }
}

// ================================================================ photons: Breit-Wheeler
// Photon loop of collide() in the plane wave (src/main.rs:212-264 with
// plane_wave.rs:284-316 and pair_creation/mod.rs:21-41): straight-line steps,
// the decay probability and Stokes evolution of the photon, and the survival
// test. A photon that decays is suspended to the event queue with its state at
// the start of the step's update (position after the move, Stokes parameters
// before it); the pair-event kernel repeats the step's probability from that
// record and completes the step.
template <bool CIRC>
__device__ __forceinline__ double lma_photon_prob(const LmaArgs& M, const LaserDev& L, double dt, d4 r, d4 k, sv4 pin,
                                                  double& a, double& k0, double& eta, sv4& pnew, bool& in_table) {
    const double kc = C_LIGHT * T_COMPTON;
    double phase = L.k0 * (r.t - r.z);
    a = sqrt(lma_a_sqd(L, r));
    k0 = kc * L.k0 * (1.0 + 2.0 * L.chirp_b * phase);   // kappa = k0 (1, 0, 0, 1)
    eta = k0 * (k.t - k.z);
    double dphi = eta * dt / (T_COMPTON * k.t);
    sv4 s = stokes_rotate(stokes_to_lma(pin, k), M.pol_angle);
    // the LP rates take the peak amplitude a sqrt 2, the CP ones the rms a (pair_creation/mod.rs:28-36)
    double prob = CIRC ? cpbw_probability(M.lcfa, M.cp, M.csc, a, eta, s, dphi, in_table)
                       : bw_probability(M.bw, M.bsc, a * 1.4142135623730951, eta, s, dphi, in_table);
    pnew = stokes_from_lma(stokes_rotate(s, -M.pol_angle), k);
    return prob;
}

template <bool CIRC>
__global__ __launch_bounds__(256) void k_photon_step_lma(const TransportArgs* __restrict__ A, LmaArgs M) {
    LaserDev L = A->laser;  // uniform; a batch's per-collision a0 is taken at fetch (as k_step_lma)
    const double dt = A->dt, cdt = C_LIGHT * dt;
    uint64_t n_steps = 0;
    uint32_t flags = 0, psteps = 0;
    uint32_t idx = 0, meta = 0, nev = 0;
    d4 r, k;
    sv4 pol;
    double icount = 0.0, weight = 1.0, S = 1.0, U = 0.0;
    Rng rng{};
    bool skip = false;

    auto fetch = [&]() -> bool {
        for (;;) {
            uint32_t i = wave_reserve32(A->work_ctr, 1u);
            if (i >= A->n_in) return false;
            const PBuf& in = A->in;
            meta = in.meta[i];
            if ((meta & 3u) != 2u) { forward_record(*A, i); continue; }
            idx = i;
            if (M.coll) L.a0 = M.lasers[M.coll[in.shower[i]]].a0;
            r = {in.d[P_RT][i], in.d[P_RX][i], in.d[P_RY][i], in.d[P_RZ][i]};
            k = {in.d[P_UT][i], in.d[P_UX][i], in.d[P_UY][i], in.d[P_UZ][i]};
            pol = {in.d[P_SI][i], in.d[P_SQ][i], in.d[P_SU][i], in.d[P_SV][i]};
            icount = in.d[P_ICOUNT][i];
            weight = in.d[P_WEIGHT][i];
            nev = in.nev[i];
            rng = rng_make(A->key, in.stream[i]);
            rng.c1 = nev;
            U = rng_uniform(rng);
            S = 1.0;
            skip = !A->opt.tracking_photons || (!A->opt.tracking_secondaries && !(meta & META_PRIMARY));
            psteps = 0;
            return true;
        }
    };
    // finished lanes wait and are written out and refilled together (as k_step_lma)
    uint32_t st = LANE_EMPTY;
    double k0_ev = 0.0, a_ev = 0.0;
    for (;;) {
        if (st == LANE_ACTIVE) {
            bool inside = !skip && field_contains(L, r) && r.t / C_LIGHT < A->opt.t_stop;
            if (inside && psteps >= A->step_cap) { flags |= PTG_FLAG_STEP_CAP; inside = false; }
            if (!inside) {
                st = LANE_EXIT;
            } else {
                ++n_steps;
                ++psteps;
                double ik = cdt / k.t;
                r = {r.t + cdt, r.x + ik * k.x, r.y + ik * k.y, r.z + ik * k.z};
                sv4 pin = A->opt.pol_resolved ? pol : sv4{1.0, 0.0, 0.0, 0.0};
                double a, k0, eta;
                sv4 pnew;
                bool in_table;
                // outside both rate tables the probability is 0 without a diagnostic (lp/mod.rs:388-397)
                double prob = lma_photon_prob<CIRC>(M, L, dt, r, k, pin, a, k0, eta, pnew, in_table);
                double ri = A->opt.rate_increase;
                if (prob * ri > 0.1) ri = 0.1 / prob;
                S *= (1.0 - prob * ri);
                if (S < U) {
                    st = LANE_EVENT;
                    k0_ev = k0;
                    a_ev = a;
                } else {
                    if (A->opt.rotate_stokes_pars && A->opt.pol_resolved) pol = pnew;
                    icount += prob;
                }
            }
        }
        const uint64_t waiting = __ballot(st == LANE_EXIT || st == LANE_EVENT || st == LANE_EMPTY);
        const uint64_t active = __ballot(st == LANE_ACTIVE);
        if (!waiting) {
            if (!active) break;
            continue;
        }
        if (active && __popcll(waiting) < PTG_LMA_REFILL) continue;
        if (st == LANE_EXIT) {
            finalize_photon(*A, idx, r, k, pol, weight, icount, meta, false);
            st = LANE_EMPTY;
        } else if (st == LANE_EVENT) {
            unsigned long long j = wave_reserve(M.ev_ctr, 1u);
            if (j < M.ev_cap) {
                const PBuf& q = M.ev;
                copy_record(q, j, A->in, idx);
                q.d[P_RT][j] = r.t; q.d[P_RX][j] = r.x; q.d[P_RY][j] = r.y; q.d[P_RZ][j] = r.z;
                q.d[P_UT][j] = k.t; q.d[P_UX][j] = k.x; q.d[P_UY][j] = k.y; q.d[P_UZ][j] = k.z;
                q.d[P_SI][j] = pol.i; q.d[P_SQ][j] = pol.q; q.d[P_SU][j] = pol.u; q.d[P_SV][j] = pol.v;
                q.d[P_WORK][j] = A->in.d[P_WORK][idx];
                q.d[P_ICOUNT][j] = icount;
                q.d[P_WEIGHT][j] = weight;
                q.nev[j] = nev;
                M.ev_k0[j] = k0_ev;
                M.ev_a[j] = a_ev;
            }
            st = LANE_EMPTY;
        }
        if (st == LANE_EMPTY) st = fetch() ? LANE_ACTIVE : LANE_DONE;
    }
    wave_add(A->stats + STAT_PHOTON_STEPS, n_steps);
    if (flags) atomicOr(A->flags, flags);
}

// (s, theta) of an LP pair at harmonic D.n by rejection on the double-differential
// rate (lp/mod.rs:261-301), one decay per wave: all 64 lanes call this with the
// same arguments (a block is one wave). Every trial costs a J(n, x, y) evaluation,
// whose quadrature grows with n (thousands of nodes at the harmonics of a0 ~ 10-20),
// so the trials run in parallel: the theta bound's 16 points on lanes 0-15, the
// ceiling's 50 points along theta = 0 on lanes 0-49 while lanes 50-63 take the
// first 14 candidates, then rounds of 64 candidates; the sample is the
// first accepted candidate in candidate order. Candidate t draws (s, theta) from
// counter 16 + 2t of `base` and its acceptance from 17 + 2t, so the sample does
// not depend on the round sizes; counter 1 gives the side of s = 1/2 and the
// quadrant of theta. From harmonic PTG_LMA_SHARED_N up the wave shares each J
// evaluation instead and takes the ceiling's points and the candidates one after
// another (the same candidates and accepted one, up to the rounding of the
// partial sums). After 2^20 candidates the sampler gives up, keeps candidate
// 2^20 and sets `capped`. `trials` counts the candidates consumed.
__device__ __noinline__ void bw_sample_wave(const DBessel& D, double a, double eta, double sv1, double sv2, const Rng& base,
                                      double* tb_s, double* tb_f, double& s_out, double& theta_out, uint32_t& trials,
                                      bool& capped) {
    const int lane = threadIdx.x;
    const uint32_t cap = 1u << 20;
    constexpr int m = BW_CEILING_POINTS;
    double s_min, s_max;
    bw_s_bounds(D.n, a, eta, s_min, s_max);
    __syncthreads();  // the previous decay's bound is no longer read
    if (lane < 16) bw_theta_bound_point(D.n, a, eta, lane, tb_s[lane], tb_f[lane]);
    __syncthreads();
    if (lane == 0) bw_theta_bound_running_min(tb_f);
    __syncthreads();
    auto candidate = [&](uint32_t t, double& sc, double& thc, double& z, double& fc) -> bool {
        Rng g = base;
        g.c0 = 16u + 2u * t;
        double u1, u2;
        rng_uniform2(g, u1, u2);
        sc = s_min + (s_max - s_min) * u1;
        thc = 0.5 * PI_D * u2;
        if (thc > bw_theta_at(tb_s, tb_f, sc)) return false;
        z = rng_uniform(g);
        fc = bw_fully_resolved(D, a, eta, sc, thc, sv1, sv2);
        return true;
    };
    double lower, upper;
    bw_ceiling_range(D.n, a, eta, lower, upper);
    int first = -1;
    double s = 0.0, theta = 0.0;
    if (D.n >= PTG_LMA_SHARED_N) {  // the wave shares each J evaluation; every lane holds the same values
        const DBesselWave W{D.n};
        double f = -1.0e300;
        for (int k = 0; k < m; ++k) f = fmax(f, bw_fully_resolved(W, a, eta, lower + (upper - lower) * double(k) / double(m), 0.0, sv1, sv2));
        const double mx = 1.2 * f;
        for (uint32_t t = 0; first < 0; ++t) {
            Rng g = base;
            g.c0 = 16u + 2u * t;
            double u1, u2;
            rng_uniform2(g, u1, u2);
            const double st = s_min + (s_max - s_min) * u1, tht = 0.5 * PI_D * u2;
            if (t >= cap) {  // give up and keep candidate `cap`
                capped = true;
                s = st;
                theta = tht;
                trials += cap;
                first = 0;
                break;
            }
            if (tht > bw_theta_at(tb_s, tb_f, st)) continue;
            const double zt = rng_uniform(g);
            if (mx * zt < bw_fully_resolved(W, a, eta, st, tht, sv1, sv2)) {
                s = st;
                theta = tht;
                trials += t + 1;
                first = 0;
            }
        }
    }
    double f = -1.0e300, sc = 0.0, thc = 0.0, z = 0.0, fc = 0.0;
    bool in_bound = false;
    if (first < 0) {
        if (lane < m) f = bw_fully_resolved(D, a, eta, lower + (upper - lower) * double(lane) / double(m), 0.0, sv1, sv2);
        else in_bound = candidate(uint32_t(lane - m), sc, thc, z, fc);
    }
    const double mx = 1.2 * wave_fmax(f);
    uint32_t t0 = 64u - uint32_t(m);  // candidates consumed
    if (first < 0) {
        const uint64_t mask = __ballot(lane >= m && in_bound && mx * z < fc);
        if (mask) {
            first = __ffsll((long long)mask) - 1;
            s = __shfl(sc, first, 64);
            theta = __shfl(thc, first, 64);
            trials += uint32_t(first - m) + 1;
        }
    }
    while (first < 0) {
        bool acc = false;
        const uint32_t t = t0 + uint32_t(lane);
        if (t < cap && candidate(t, sc, thc, z, fc)) acc = mx * z < fc;
        const uint64_t mask = __ballot(acc);
        if (mask) {
            first = __ffsll((long long)mask) - 1;
            s = __shfl(sc, first, 64);
            theta = __shfl(thc, first, 64);
            trials += t0 + uint32_t(first) + 1;
        } else {
            t0 += 64;
            if (t0 >= cap) {  // give up and keep candidate `cap`
                capped = true;
                Rng g = base;
                g.c0 = 16u + 2u * cap;
                double u1, u2;
                rng_uniform2(g, u1, u2);
                s = s_min + (s_max - s_min) * u1;
                theta = 0.5 * PI_D * u2;
                trials += cap;
                first = 0;
            }
        }
    }
    Rng g = base;
    g.c0 = 1u;
    double u1, u2;
    rng_uniform2(g, u1, u2);
    if (int(u1 * 2.0) == 0) s = 1.0 - s;
    const int quadrant = int(u2 * 4.0);
    if (quadrant == 1) theta = PI_D - theta;
    else if (quadrant == 2) theta = PI_D + theta;
    else if (quadrant == 3) theta = 2.0 * PI_D - theta;
    s_out = s;
    theta_out = theta;
}

// Inclusive prefix sum over the 64 lanes of a wave (a block is one wave).
__device__ __forceinline__ double wave_inclusive_sum(double v) {
    const int lane = threadIdx.x;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const double t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}

// Harmonic of a CP pair by explicit summation of the partial rates until the
// cumulative rate passes frac x the total (pair_creation/cp/mod.rs:327-365),
// one decay per wave: all 64 lanes call this with the same arguments and return
// the same value. The reference sums one harmonic after another; at the
// amplitudes of a0 ~ 10-20 the sum runs over thousands of harmonics, each a
// 64-96 point quadrature of Bessel functions, so the lanes take 64 consecutive
// harmonics at a time and a prefix sum over the wave finds the first one whose
// cumulative rate passes the target (the same harmonic, up to the rounding of
// the partial sums). n_max - 1 with `fell_back` set if the sum never passes it,
// -1 if the total is not available.
__device__ __noinline__ int cpbw_harmonic_wave(TP tb, const double* P, const CpScalars& c, double a, double eta, double sv3,
                                               double frac, bool& fell_back) {
    const int lane = threadIdx.x;
    int lo, hi;
    cpbw_sum_limits(a, eta, lo, hi);
    double f[2];
    fell_back = false;
    if (!cpbw_total(tb, P, c, a, eta, f)) return -1;
    const double target = 0.5 * (f[0] * (1.0 + sv3) + f[1] * (1.0 - sv3)) * frac;
    double base = 0.0;
    for (int k0 = lo; k0 <= hi; k0 += 64) {
        const int k = k0 + lane;
        double v = 0.0;
        if (k <= hi) {
            const cplx pr = cpbw_partial_rate(tb, P + cp_pack::G32N, P + cp_pack::G32W, k, a, eta);
            v = 0.5 * ((1.0 + sv3) * pr.re + (1.0 - sv3) * pr.im);
        }
        const double cum = base + wave_inclusive_sum(v);
        const uint64_t mask = __ballot(k <= hi && cum > target);
        if (mask) return k0 + (__ffsll((long long)mask) - 1);
        base = __shfl(cum, 63, 64);
    }
    fell_back = true;
    return hi - 1;
}

// Pair events (pair_creation/mod.rs:47-107 with lp/mod.rs:424-442 and
// plane_wave.rs:305-316): harmonic from the CDF tables, (s, theta) by rejection
// on the double-differential rate, the positron built in the zero-momentum
// frame; the electron and positron go to the pair queue, the photon is kept
// (decayed) or re-queued with its reduced weight. One event per wave: every lane
// computes the event's scalars, the lanes share the harmonic sum (CP) or the
// rejection sampling (LP), and lane 0 writes.
template <bool CIRC>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(PTG_LMA_WAVE_WPE))) void k_pair_event_lma(
    const TransportArgs* __restrict__ A, LmaArgs M) {
    const uint32_t lane = threadIdx.x;
    const uint64_t e0 = blockIdx.x, stride = gridDim.x;
    const bool writer = lane == 0;
    __shared__ double tb_s[16], tb_f[16];  // LP: the decay's theta bound, one copy per wave
    uint64_t n_pairs = 0, trials = 0;
    uint32_t flags = 0;
    const uint64_t n_ev = M.n_ev_dev ? *M.n_ev_dev : M.n_ev;
    for (uint64_t e = e0; e < n_ev; e += stride) {
        const PBuf& q = M.ev_in;
        uint32_t i = M.ev_list ? M.ev_list[e] : uint32_t(e);
        // a batch of collisions: the photon's own collision's laser
        const LaserDev& L = M.coll ? M.lasers[M.coll[q.shower[i]]] : A->laser;
        d4 r{q.d[P_RT][i], q.d[P_RX][i], q.d[P_RY][i], q.d[P_RZ][i]};
        d4 k{q.d[P_UT][i], q.d[P_UX][i], q.d[P_UY][i], q.d[P_UZ][i]};
        sv4 pol{q.d[P_SI][i], q.d[P_SQ][i], q.d[P_SU][i], q.d[P_SV][i]};
        double icount = q.d[P_ICOUNT][i], weight = q.d[P_WEIGHT][i];
        uint32_t nev = q.nev[i], meta = q.meta[i];
        sv4 pin = A->opt.pol_resolved ? pol : sv4{1.0, 0.0, 0.0, 0.0};
        double a, k0, eta;
        sv4 pnew;
        bool in_table;
        double prob = lma_photon_prob<CIRC>(M, L, A->dt, r, k, pin, a, k0, eta, pnew, in_table);
        double ri = A->opt.rate_increase;
        if (prob * ri > 0.1) ri = 0.1 / prob;
        double frac = 1.0 / ri;
        Rng rng = rng_make(A->key, q.stream[i]);
        rng.c1 = nev | EVENT_BLOCK;
        double alp = a * 1.4142135623730951;
        sv4 sl = stokes_rotate(stokes_to_lma(pin, k), M.pol_angle);
        bool fell_back = false;
        int n = CIRC ? cpbw_harmonic_wave(M.lcfa, M.cp, M.csc, a, eta, sl.v, rng_uniform(rng), fell_back)
                     : bw_harmonic(M.bw, M.bsc, alp, eta, sl.q, rng_uniform(rng));
        bool made = false;
        if (n < 1) {
            flags |= PTG_FLAG_TABLE_RANGE;
        } else {
            double s, cphi;
            uint32_t tr = 0;
            bool capped = false;
            if (CIRC) cpbw_sample_at(M.lcfa, M.cp, n, a, eta, sl.q, sl.u, sl.v, rng, s, cphi, tr, capped);
            else bw_sample_wave(dbessel_of(n), alp, eta, sl.q, sl.u, rng, tb_s, tb_f, s, cphi, tr, capped);
            trials += tr;
            if (capped) flags |= PTG_FLAG_REJECT_CAP;
            // positron in the zero-momentum frame (pair_creation/mod.rs:63-104)
            double jn = double(n);
            double e_zmf = sqrt(0.5 * jn * eta);
            double p_zmf = sqrt(0.5 * jn * eta - 1.0 - a * a);
            double cos_theta_zmf = (1.0 - 2.0 * s) * e_zmf / p_zmf;
            d4 lk{k.t + jn * k0, k.x, k.y, k.z + jn * k0};
            double inv = 1.0 / sqrt(mdot(lk, lk));
            d4 u_zmf{lk.t * inv, lk.x * inv, lk.y * inv, lk.z * inv};
            d3 along = -unit(sp(boost4(d4{jn * k0, 0.0, 0.0, jn * k0}, u_zmf)));
            d3 eps = unit(sp(boost4(d4{0.0, M.cos_pa, M.sin_pa, 0.0}, u_zmf)));
            eps = unit(cross(along, cross(eps, along)));
            double sp_, cp_;
            sincos(cphi, &sp_, &cp_);
            d3 perp = rotate(eps, along, sp_, cp_);
            d3 q3 = p_zmf * (cos_theta_zmf * along + sqrt(1.0 - cos_theta_zmf * cos_theta_zmf) * perp);
            d4 qp{sqrt(1.0 + a * a + nsq(q3)), q3.x, q3.y, q3.z};
            qp = boost4(qp, d4{u_zmf.t, -u_zmf.x, -u_zmf.y, -u_zmf.z});
            double s_new = 1.0 - (k0 * (qp.t - qp.z)) / eta;
            if (!(fabs(s - s_new) / s < 1.0e-3)) flags |= PTG_FLAG_SANITY;
            d4 qe{k.t + jn * k0 - qp.t, k.x - qp.x, k.y - qp.y, k.z + jn * k0 - qp.z};
            // electron and positron (src/main.rs:220-243): weight frac * w, half the absorbed energy each
            unsigned long long j = 0;
            uint64_t id = 0;
            if (writer) {
                j = wave_reserve(M.pairs_ctr, 2u);
                id = A->id_base + wave_reserve(A->id_ctr, 2u);
            }
            if (writer && j + 1 < M.pairs_cap) {
                sv4 unpol{1.0, 0.0, 0.0, 0.0};
                d4 u0{1.0, 0.0, 0.0, 0.0};
                uint64_t base = rng_child(q.stream[i], nev);
                write_new(M.pairs, j, 0u, r, qe, u0, unpol, a, a * eta, frac * weight, id, q.id[i], rng_child(base, 0), q.shower[i]);
                write_new(M.pairs, j + 1, 1u, r, qp, u0, unpol, a, a * eta, frac * weight, id + 1, q.id[i], rng_child(base, 1),
                          q.shower[i]);
                M.pairs.d[P_WORK][j] = 0.5 * jn * k0;
                M.pairs.d[P_WORK][j + 1] = 0.5 * jn * k0;
            }
            ++n_pairs;
            made = true;
        }
        // complete the photon's step (src/main.rs:244-257)
        bool decayed = made && frac == 1.0;
        if (made && !decayed) weight = weight * (1.0 - frac);
        if (A->opt.rotate_stokes_pars && A->opt.pol_resolved) pol = pnew;
        icount += prob;
        if (!writer) continue;
        if (decayed) {
            finalize_photon_from(*A, q, i, r, k, pol, weight, icount, meta, true);
        } else {
            unsigned long long j = wave_reserve(M.next_ctr, 1u);
            const PBuf& o = M.next;
            copy_record(o, j, q, i);
            o.d[P_RT][j] = r.t; o.d[P_RX][j] = r.x; o.d[P_RY][j] = r.y; o.d[P_RZ][j] = r.z;
            o.d[P_UT][j] = k.t; o.d[P_UX][j] = k.x; o.d[P_UY][j] = k.y; o.d[P_UZ][j] = k.z;
            o.d[P_SI][j] = pol.i; o.d[P_SQ][j] = pol.q; o.d[P_SU][j] = pol.u; o.d[P_SV][j] = pol.v;
            o.d[P_WORK][j] = q.d[P_WORK][i];
            o.d[P_ICOUNT][j] = icount;
            o.d[P_WEIGHT][j] = weight;
            o.nev[j] = nev + 1;
        }
    }
    if (!writer) return;
    wave_add(A->stats + STAT_PAIRS, n_pairs);
    wave_add(A->stats + STAT_TRIALS, trials);
    if (flags) atomicOr(A->flags, flags);
}

void launch_photon_step_lma(const TransportArgs* a, const LmaArgs& m, int grid, hipStream_t s) {
    if (m.mode == LMA_CP || m.mode == LMA_CP_CLASSICAL) k_photon_step_lma<true><<<grid, 256, 0, s>>>(a, m);
    else k_photon_step_lma<false><<<grid, 256, 0, s>>>(a, m);
}
void launch_pair_event_lma(const TransportArgs* a, const LmaArgs& m, int lanes, hipStream_t s) {
    if (m.mode == LMA_CP || m.mode == LMA_CP_CLASSICAL) k_pair_event_lma<true><<<lanes / 64, 64, 0, s>>>(a, m);
    else k_pair_event_lma<false><<<lanes / 64, 64, 0, s>>>(a, m);
}

void launch_step_lma(const TransportArgs* a, const LmaArgs& m, int grid, hipStream_t s, int block) {
    switch (m.mode) {
    case LMA_CP:
        if (m.mixed) k_step_lma<LMA_CP, true><<<grid, block, 0, s>>>(a, m);
        else k_step_lma<LMA_CP, false><<<grid, block, 0, s>>>(a, m);
        break;
    case LMA_CP_CLASSICAL: k_step_lma<LMA_CP_CLASSICAL, false><<<grid, block, 0, s>>>(a, m); break;
    case LMA_LP_CLASSICAL: k_step_lma<LMA_LP_CLASSICAL, false><<<grid, block, 0, s>>>(a, m); break;
    default:
        if (m.mixed) k_step_lma<LMA_LP, true><<<grid, block, 0, s>>>(a, m);
        else k_step_lma<LMA_LP, false><<<grid, block, 0, s>>>(a, m);
        break;
    }
}
void launch_event_lma(const TransportArgs* a, const LmaArgs& m, int lanes, hipStream_t s) {
    switch (m.mode) {
    case LMA_CP: k_event_lma<LMA_CP><<<lanes / 64, 64, 0, s>>>(a, m); break;
    case LMA_CP_CLASSICAL: k_event_lma<LMA_CP_CLASSICAL><<<lanes / 64, 64, 0, s>>>(a, m); break;
    case LMA_LP_CLASSICAL: k_event_lma_wave<LMA_LP_CLASSICAL><<<lanes / 64, 64, 0, s>>>(a, m); break;
    default: k_event_lma_wave<LMA_LP><<<lanes / 64, 64, 0, s>>>(a, m); break;
    }
}

// Field-only push in the plane wave (test support for PlaneWave::push).
__global__ void k_push_lma(LaserDev L, double dt, int ll, const uint8_t* species, double* r4, double* u4, uint32_t n,
                           uint32_t n_steps, uint32_t* taken, double* work) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    d4 r{r4[4 * i], r4[4 * i + 1], r4[4 * i + 2], r4[4 * i + 3]};
    d4 u{u4[4 * i], u4[4 * i + 1], u4[4 * i + 2], u4[4 * i + 3]};
    double rqm = species[i] == 1 ? -RQM_ELECTRON : RQM_ELECTRON;
    double w = 0.0;
    uint32_t k = 0;
    for (; k < n_steps && field_contains(L, r); ++k) w += lma_push(L, r, u, rqm, dt, ll);
    r4[4 * i] = r.t; r4[4 * i + 1] = r.x; r4[4 * i + 2] = r.y; r4[4 * i + 3] = r.z;
    u4[4 * i] = u.t; u4[4 * i + 1] = u.x; u4[4 * i + 2] = u.y; u4[4 * i + 3] = u.z;
    taken[i] = k;
    if (work) work[i] = w;
}
void launch_push_lma(const LaserDev& L, double dt, int ll, const uint8_t* sp_, double* r4, double* u4, uint32_t n,
                     uint32_t n_steps, uint32_t* taken, double* work, hipStream_t s) {
    k_push_lma<<<(n + 255) / 256, 256, 0, s>>>(L, dt, ll, sp_, r4, u4, n, n_steps, taken, work);
}

// ================================================================ learned sampler (surrogate)
// The network maps (ln a, ln eta, u1, u2, u3), scaled to its domain box, to
// (ln nu / 5, v, t): the harmonic n = ceil(nu) (nu stands in for the CDF-table
// inversion of lp/mod.rs:478-553, cp/tables/mod.rs:46-121), the lightfront
// fraction s = v s_max(n) and, for LP, the first-quadrant azimuth 2 t / pi
// (stand-ins for the rejection sampling of lp/mod.rs:586-676, cp/mod.rs:181-234).
// The photon's Stokes parameters are computed from the exact rate at the sampled
// point, and the photon is built by lma_emit_finish as in the exact kernels.
// a is the sampler's amplitude: sqrt(2) a_rms for LP, a_rms for CP.
__device__ __forceinline__ bool surr_in_domain(const SurrArgs& S, double la, double le) {
    return fabs(la - S.la_c) <= S.la_h && fabs(le - S.le_c) <= S.le_h;
}
// The network is trained at the levels u2 = (k + 1) / 65, k = 0..63 (the means of
// 64 ordered uniforms); outside [1/65, 64/65] the quantile function of v is
// continued from the network's value at the end level: towards v = 0 as
// sqrt(u2) (a density rising linearly from zero, as the rate of every harmonic
// above the first vanishes at s = 0) and linearly to v = 1 at u2 = 1.
constexpr double SURR_U2_LO = 1.0 / 65.0, SURR_U2_HI = 64.0 / 65.0;
__device__ __forceinline__ double surr_u2_level(double u2) { return fmin(fmax(u2, SURR_U2_LO), SURR_U2_HI); }
__device__ __forceinline__ double surr_v_tail(double u2, double v_net) {
    const double v = fmin(fmax(v_net, 0.0), 1.0);
    if (u2 < SURR_U2_LO) return v * sqrt(u2 / SURR_U2_LO);
    if (u2 > SURR_U2_HI) return 1.0 - (1.0 - v) * ((1.0 - u2) / (1.0 - SURR_U2_HI));
    return v;
}
__device__ __forceinline__ void surr_features(const SurrArgs& S, double la, double le, double u1, double u2, double u3, float* x) {
    x[0] = float((la - S.la_c) / S.la_h);
    x[1] = float((le - S.le_c) / S.le_h);
    x[2] = float(2.0 * u1 - 1.0);
    x[3] = float(2.0 * u2 - 1.0);
    x[4] = float(2.0 * u3 - 1.0);
    x[5] = 0.0f; x[6] = 0.0f; x[7] = 0.0f;
}
template <bool CIRC>
__device__ __forceinline__ double surr_nu(const LmaArgs& M, double a, double eta, double frac) {
    return CIRC ? cp_harmonic_cont(M.cp, M.csc, a, eta, frac) : lp_harmonic_cont(M.tables + lma_pack::CDF, M.sc, a, eta, frac);
}

// Queued events -> network rows (inside the domain), the exact sampler in this
// generation (below the amplitude or eta from which the network is used: where
// the exact sampler is the cheaper, or the network falls short of it), or the
// exact sampler's deferred pass (the rest: above the domain, where an exact
// event is expensive).
// u1 is the exact path's harmonic draw (counter 0); u2, u3 come from counter 2.
template <bool CIRC>
__global__ void k_surr_inputs(const TransportArgs* __restrict__ A, LmaArgs M, SurrArgs S) {
    const uint64_t e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= (S.n_ev_dev ? *S.n_ev_dev : S.n_ev)) return;
    const PBuf& q = M.ev;
    const uint32_t i = uint32_t(e);
    const double ut = q.d[P_UT][i], ux = q.d[P_UX][i], uy = q.d[P_UY][i], uz = q.d[P_UZ][i];
    const double a_rms = sqrt(ut * ut - ux * ux - uy * uy - uz * uz - 1.0);
    const double eta = M.ev_k0[i] * (ut - uz);
    const double amp = CIRC ? a_rms : a_rms * 1.4142135623730951;
    Rng base = rng_make(A->key, q.stream[i]);
    base.c1 = q.nev[i] | EVENT_BLOCK;
    Rng g0 = base;
    const double u1 = rng_uniform(g0);
    Rng g2 = base;
    g2.c0 = 2u;
    double u2, u3;
    rng_uniform2(g2, u2, u3);
    const double la = log(amp), le = log(eta);
    if (la < S.la_use || le < S.le_use) {
        S.now[wave_reserve(S.n_now, 1u)] = i;
    } else if (surr_in_domain(S, la, le)) {
        const unsigned long long j = wave_reserve(S.n_rows, 1u);
        S.list[j] = i;
        surr_features(S, la, le, u1, surr_u2_level(u2), u3, S.x + 8 * j);
    } else {
        S.fb[wave_reserve(S.n_fb, 1u)] = i;
    }
}

// Network outputs -> photons. One lane per row; LP Stokes parameters need one
// J(n, x, y) evaluation per event.
template <bool CIRC>
__global__ __launch_bounds__(64) void k_surr_event(const TransportArgs* __restrict__ A, LmaArgs M, SurrArgs S) {
    const uint32_t lane = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n_lanes = gridDim.x * blockDim.x;
    const uint64_t n_rows = *S.n_rows;
    uint64_t n_emit = 0;
    uint32_t flags = 0;
    M.ev_in = M.ev;
    for (uint64_t rr = lane; rr < n_rows; rr += n_lanes) {
        const uint64_t r = S.perm ? S.perm[rr] : rr;
        const uint32_t i = S.list[r];
        const PBuf& q = M.ev;
        const double ut = q.d[P_UT][i], ux = q.d[P_UX][i], uy = q.d[P_UY][i], uz = q.d[P_UZ][i];
        const double a_rms = sqrt(ut * ut - ux * ux - uy * uy - uz * uz - 1.0);
        const double k0 = M.ev_k0[i];
        const double eta = k0 * (ut - uz);
        const double amp = CIRC ? a_rms : a_rms * 1.4142135623730951;
        const float* o = S.out + 32 * r;
        const double nu = exp(5.0 * double(o[0]));
        // the network's harmonic is cut at 1.5x the reference's harmonic sum limit
        // at the event's amplitude (+16), where the tables end: beyond it a network
        // output is an extrapolation (and 4 n^2 would wrap in the CP Stokes parameters)
        const double n_cap = ceil(1.5 * (6.0 * (1.0 + amp * amp) + 2.0 * amp * amp * amp)) + 16.0;
        int n = nu > 1.0 ? int(fmin(ceil(nu), n_cap)) : 1;
        Rng base = rng_make(A->key, q.stream[i]);
        base.c1 = q.nev[i] | EVENT_BLOCK;
        Rng g2 = base;  // the level u2 of the inputs (counter 2)
        g2.c0 = 2u;
        double u2, u3;
        rng_uniform2(g2, u2, u3);
        // 0 < v < 1 as the exact samplers' draws: the photon construction divides by s,
        // and sqrt(1 - cos^2) (photon) and sqrt(w (1 - w)) (CP Stokes, w -> 1 as v -> 1)
        // must not see a rounding past 1
        // the network's v at the (clamped) level is kept inside [0.005, 0.995]: an
        // output at or beyond 0 (a small error where the true end quantile is ~0.01)
        // would otherwise become v ~ 0 and emit photons of the laser frequency co-moving
        // with the pulse, each stepped for millions of steps
        const double v_net = double(o[1]);
        if (!(v_net == v_net)) {
            S.fb[wave_reserve(S.n_fb, 1u)] = i;
            continue;
        }
        double v = fmin(fmax(surr_v_tail(u2, fmin(fmax(v_net, 0.005), 0.995)), 1.0e-12), 1.0 - 1.0e-9);
        double s = 0.0, theta, sn;
        sv4 sv;
        if (CIRC) {
            sn = 2.0 * double(n) * eta / (1.0 + amp * amp);
            Rng g3 = base;
            g3.c0 = 3u;
            theta = 2.0 * PI_D * rng_uniform(g3);
        } else {
            sn = 2.0 * double(n) * eta / (1.0 + 0.5 * amp * amp);
            const double th = fmin(fmax(double(o[2]), 0.0), 1.0) * (PI_D / 2.0);
            Rng gq = base;  // counter 1: the quadrant, as the exact sampler
            gq.c0 = 1u;
            const int quadrant = int(rng_uniform(gq) * 4.0);
            theta = th;
            if (quadrant == 1) theta = PI_D - th;
            else if (quadrant == 2) theta = PI_D + th;
            else if (quadrant == 3) theta = 2.0 * PI_D - th;
        }
        // At a point where the exact rate underflows to zero (a high harmonic at v near
        // 0 or 1) the Stokes parameters are 0/0: the exact sampler never draws such a
        // point. The network's v is then moved halfway to 1/2 until the rate is
        // resolved (a few tries; such draws are ~1e-5 of all), else the event goes to
        // the exact kernels.
        bool ok = false;
        for (int tr = 0; tr < 6 && !ok; ++tr) {
            if (tr) v = 0.5 + 0.5 * (v - 0.5);
            s = v * (sn / (1.0 + sn));
            if (CIRC) {
                double st[4];
                cp_stokes(M.lcfa, n, amp, s, sn, v, theta, false, st);
                sv = {st[0], st[1], st[2], st[3]};
            } else {
                sv = lp_stokes(dbessel_of(n), amp, eta, s, theta);
            }
            ok = isfinite(sv.q) && isfinite(sv.u) && isfinite(sv.v);
        }
        if (!ok) {
            S.fb[wave_reserve(S.n_fb, 1u)] = i;
            continue;
        }
        lma_emit_finish<false>(A, M, i, k0, a_rms, eta, true, n, s, theta, sv, n_emit, flags);
    }
    wave_add(A->stats + STAT_EMISSIONS, n_emit);
    wave_add(A->stats + STAT_SURROGATE, n_emit);
    if (flags) atomicOr(A->flags, flags);
}

// Training data from the exact sampler: per wave one (a, eta, u1) drawn in the
// domain box, its harmonic n = ceil(nu) from the tables, and 64 samples of
// (v, t) at that harmonic. LP: the wave evaluates the ceiling points, then
// 64 candidates per round, keeping every accepted one (rejection sampling
// under the ceiling of lp/mod.rs:53-140 within the theta bound of :143-302);
// CP: every lane runs the exact sampler (cp/mod.rs:181-234) at n.
template <bool CIRC>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(PTG_LMA_WAVE_WPE))) void k_surr_groups(LmaArgs M, SurrArgs S) {
    const int lane = threadIdx.x;
    __shared__ double tb_s[16], tb_f[16];
    for (uint32_t g = blockIdx.x; g < S.n_groups; g += gridDim.x) {
        __syncthreads();
        Rng gr = rng_make(S.key, g);
        gr.c1 = 0x40000000u;
        double ua, ue;
        rng_uniform2(gr, ua, ue);
        const double u1 = rng_uniform(gr);
        const double la = S.la_c + S.la_h * (2.0 * ua - 1.0), le = S.le_c + S.le_h * (2.0 * ue - 1.0);
        const double amp = exp(la), eta = exp(le);
        const double nu = surr_nu<CIRC>(M, amp, eta, u1);
        const int n = nu > 1.0 ? int(ceil(nu)) : 1;
        bool ok = nu > 0.0 && n <= 10000000;
        float* out = S.gs + size_t(g) * 128;
        if (CIRC) {
            if (ok) {
                Rng rl = rng_make(S.key, (uint64_t(g) << 6) | uint64_t(lane));
                rl.c1 = 0x50000000u;
                double s, theta;
                sv4 sv;
                uint32_t tr = 0;
                bool capped = false;
                cp_sample_at(M.lcfa, M.cp, n, amp, eta, false, rl, s, theta, sv, tr, capped);
                const double sn = 2.0 * double(n) * eta / (1.0 + amp * amp);
                out[2 * lane] = float(s / (sn / (1.0 + sn)));
                out[2 * lane + 1] = 0.0f;
                if (capped) ok = false;
            }
            ok = __all(ok);
        } else if (ok) {
            const DBessel D = dbessel_of(n);
            const double sn = 2.0 * double(n) * eta / (1.0 + 0.5 * amp * amp);
            const double smax = sn / (1.0 + sn);
            if (lane < 16) theta_bound_point(n, amp, eta, lane, tb_s[lane], tb_f[lane]);
            __syncthreads();
            if (lane == 0) theta_bound_running_min(tb_f);
            __syncthreads();
            const int m = lp_ceiling_count(n, amp, eta);
            double f = -INFINITY;
            for (int k = lane; k < m; k += 64) {
                double sk, tk;
                lp_ceiling_arg(n, amp, eta, k, sk, tk);
                f = fmax(f, lp_ddpr(D, amp, eta, sk, tk));
            }
            const double mx = 1.15 * wave_fmax(f);
            int filled = 0;
            for (uint32_t round = 0; filled < 64 && round < 4096; ++round) {
                Rng gc = gr;
                gc.c1 = 0x48000000u;
                gc.c0 = 16u + 2u * (round * 64u + uint32_t(lane));
                double v, th;
                rng_uniform2(gc, v, th);
                th *= PI_D / 2.0;
                bool acc = false;
                if (th <= theta_bound_at(tb_s, tb_f, smax * v)) {
                    const double uz = rng_uniform(gc);
                    acc = mx * uz < lp_ddpr(D, amp, eta, smax * v, th);
                }
                const uint64_t mask = __ballot(acc);
                const int slot = filled + __popcll(mask & ((1ull << lane) - 1ull));
                if (acc && slot < 64) {
                    out[2 * slot] = float(v);
                    out[2 * slot + 1] = float(th / (PI_D / 2.0));
                }
                filled += __popcll(mask);
            }
            ok = filled >= 64;
        } else {
            ok = false;
        }
        if (lane == 0) {
            double* h = S.gh + size_t(g) * 4;
            h[0] = la; h[1] = le; h[2] = u1; h[3] = ok ? nu : 0.0;
        }
    }
}

// One training batch: rows [0, n_table_rows) drawn fresh from the harmonic
// tables (target nu only), the rest drawn from the exact-sampler rows (targets
// nu, the rank-coupled v and, for LP, t at a fresh quantile level u3).
template <bool CIRC>
__global__ void k_surr_batch(LmaArgs M, SurrArgs S) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= S.batch) return;
    Rng rr = rng_make(S.key, (uint64_t(S.step) << 32) | uint64_t(r));
    rr.c1 = 0x60000000u;
    float* x = S.tx + 8 * size_t(r);
    float* t = S.tt + 8 * size_t(r);
    double a1, a2, u1, u2, u3, dummy;
    rng_uniform2(rr, a1, a2);
    rng_uniform2(rr, u1, u2);
    rng_uniform2(rr, u3, dummy);
    if (r < S.n_table_rows) {
        const double la = S.la_c + S.la_h * (2.0 * a1 - 1.0), le = S.le_c + S.le_h * (2.0 * a2 - 1.0);
        const double nu = fmax(surr_nu<CIRC>(M, exp(la), exp(le), u1), 0.9);
        surr_features(S, la, le, u1, u2, u3, x);
        t[0] = float(log(nu) / 5.0); t[1] = 0.0f; t[2] = 0.0f; t[3] = 0.0f; t[4] = 0.0f; t[5] = float(u3);
    } else {
        const uint32_t idx = uint32_t(fmin(a1 * double(S.n_sampler_rows), double(S.n_sampler_rows - 1)));
        const float* row = S.rows + 8 * size_t(idx);
        surr_features(S, row[0], row[1], row[2], row[3], u3, x);
        t[0] = row[4]; t[1] = row[5]; t[2] = row[6]; t[3] = 1.0f; t[4] = CIRC ? 0.0f : 1.0f; t[5] = float(u3);
    }
    t[6] = 0.0f; t[7] = 0.0f;
}

// Rows ordered by the power-of-two bucket of their harmonic, largest first: a
// lane's time is that of its J(n, x, y) recurrence, ~n, so a wave of lanes with
// like harmonics does not wait for one large one, and the longest waves start first.
__device__ __forceinline__ int surr_bucket(const float* o) {
    const double nu = exp(5.0 * double(o[0]));
    return nu <= 1.0 ? 0 : min(31, int(log2(nu)) + 1);
}
__global__ void k_surr_count(SurrArgs S, uint32_t n_rows) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n_rows) atomicAdd(S.bucket + surr_bucket(S.out + 32 * size_t(r)), 1u);
}
__global__ void k_surr_scan(SurrArgs S) {
    if (threadIdx.x != 0) return;
    unsigned int off = 0;
    for (int b = 31; b >= 0; --b) {  // largest harmonics first
        S.bucket[32 + b] = off;
        off += S.bucket[b];
    }
}
__global__ void k_surr_scatter(SurrArgs S, uint32_t n_rows) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n_rows) S.perm[atomicAdd(S.bucket + 32 + surr_bucket(S.out + 32 * size_t(r)), 1u)] = r;
}
void launch_surr_order(const SurrArgs& s, uint32_t n_rows, hipStream_t st) {
    if (!n_rows) return;
    const unsigned g = (n_rows + 255) / 256;
    k_surr_count<<<g, 256, 0, st>>>(s, n_rows);
    k_surr_scan<<<1, 64, 0, st>>>(s);
    k_surr_scatter<<<g, 256, 0, st>>>(s, n_rows);
}

void launch_surr_inputs(const TransportArgs* a, const LmaArgs& m, const SurrArgs& s, hipStream_t st) {
    const unsigned blocks = unsigned((s.n_ev + 255) / 256);
    if (!blocks) return;
    if (m.mode == LMA_CP) k_surr_inputs<true><<<blocks, 256, 0, st>>>(a, m, s);
    else k_surr_inputs<false><<<blocks, 256, 0, st>>>(a, m, s);
}
void launch_surr_event(const TransportArgs* a, const LmaArgs& m, const SurrArgs& s, int lanes, hipStream_t st) {
    if (m.mode == LMA_CP) k_surr_event<true><<<lanes / 64, 64, 0, st>>>(a, m, s);
    else k_surr_event<false><<<lanes / 64, 64, 0, st>>>(a, m, s);
}
void launch_surr_groups(const LmaArgs& m, const SurrArgs& s, int lanes, hipStream_t st) {
    if (m.mode == LMA_CP) k_surr_groups<true><<<lanes / 64, 64, 0, st>>>(m, s);
    else k_surr_groups<false><<<lanes / 64, 64, 0, st>>>(m, s);
}
void launch_surr_batch(const LmaArgs& m, const SurrArgs& s, hipStream_t st) {
    if (m.mode == LMA_CP) k_surr_batch<true><<<(s.batch + 255) / 256, 256, 0, st>>>(m, s);
    else k_surr_batch<false><<<(s.batch + 255) / 256, 256, 0, st>>>(m, s);
}

// Moves events out of the event queue (list of indices) into another queue with
// their kappa^0 and sqrt(<a^2>), at [off, off + n).
__global__ void k_gather_events(PBuf src, const double* __restrict__ k0s, const double* __restrict__ as_,
                                const uint32_t* __restrict__ list, uint64_t n, PBuf dst, double* k0d, double* ad, uint64_t off) {
    const uint64_t j = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t i = list[j];
    const uint64_t o = off + j;
#pragma unroll
    for (int c = 0; c < P_NDOUBLE; ++c) dst.d[c][o] = src.d[c][i];  // the whole record, as queued
    dst.id[o] = src.id[i];
    dst.parent_id[o] = src.parent_id[i];
    dst.stream[o] = src.stream[i];
    dst.shower[o] = src.shower[i];
    dst.meta[o] = src.meta[i];
    dst.nev[o] = src.nev[i];
    k0d[off + j] = k0s[i];
    ad[off + j] = as_[i];
}
void launch_gather_events(const PBuf& src, const double* k0s, const double* as_, const uint32_t* list, uint64_t n, const PBuf& dst,
                          double* k0d, double* ad, uint64_t off, hipStream_t s) {
    if (n) k_gather_events<<<unsigned((n + 255) / 256), 256, 0, s>>>(src, k0s, as_, list, n, dst, k0d, ad, off);
}

// ---------------------------------------------------------------- captured generations
// Kernels whose counts live in device memory, so that a sequence of
// generations can be captured once and replayed without the host reading a
// count between launches.

__global__ void k_gen_begin(unsigned long long* ctr, TransportArgs* A, unsigned int* bucket) {
    if (threadIdx.x == 0) {
        ctr[C_EVRUN] = 0ull;
        ctr[C_SROWS] = 0ull;
        ctr[C_SFB] = 0ull;
        ctr[C_SNOW] = 0ull;
        if (ctr[C_STALL]) {
            A->n_in = 0u;
            ctr[C_NCUR] = 0ull;
        } else {
            const unsigned long long n = ctr[C_LNEXT];
            A->n_in = uint32_t(n);
            ctr[C_NCUR] = n;
            if (n) ctr[C_GENS] += 1ull;
            ctr[C_WORK] = 0ull;
            ctr[C_EV] = 0ull;
            ctr[C_LNEXT] = 0ull;
        }
    }
    if (bucket) bucket[threadIdx.x] = 0u;
}
void launch_gen_begin(unsigned long long* ctr, TransportArgs* A, unsigned int* bucket, hipStream_t s) {
    k_gen_begin<<<1, 64, 0, s>>>(ctr, A, bucket);
}

__global__ void k_event_room(unsigned long long* ctr, EventRoom k) {
    if (threadIdx.x != 0 || ctr[C_STALL]) return;
    const unsigned long long n = ctr[C_EV];
    const bool room = ctr[C_OUT] + k.out_per * n <= k.out_cap && ctr[k.ph_slot] + k.ph_per * n <= k.ph_cap &&
                      ctr[C_PAIRS] + k.pairs_per * n <= k.pairs_cap && ctr[C_INTER] + k.inter_per * n <= k.inter_cap;
    if (room) ctr[C_EVRUN] = n;
    else if (n) ctr[C_STALL] = 1ull;
}
void launch_event_room(unsigned long long* ctr, const EventRoom& k, hipStream_t s) { k_event_room<<<1, 64, 0, s>>>(ctr, k); }

__global__ void k_surr_count_dev(SurrArgs S) {
    const uint64_t r = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r < *S.n_rows) atomicAdd(S.bucket + surr_bucket(S.out + 32 * size_t(r)), 1u);
}
__global__ void k_surr_scatter_dev(SurrArgs S) {
    const uint64_t r = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r < *S.n_rows) S.perm[atomicAdd(S.bucket + 32 + surr_bucket(S.out + 32 * size_t(r)), 1u)] = uint32_t(r);
}
void launch_surr_order_dev(const SurrArgs& s, uint32_t max_rows, hipStream_t st) {
    if (!max_rows) return;
    const unsigned g = (max_rows + 255) / 256;
    k_surr_count_dev<<<g, 256, 0, st>>>(s);
    k_surr_scan<<<1, 64, 0, st>>>(s);
    k_surr_scatter_dev<<<g, 256, 0, st>>>(s);
}

// rows [n_rows, n_rows rounded up to 32) of the network's inputs, 8 floats each
__global__ void k_surr_pad(SurrArgs S) {
    const uint64_t n = *S.n_rows, m = (n + 31) / 32 * 32;
    const uint64_t k = n * 8 + threadIdx.x;
    if (k < m * 8) S.x[k] = 0.0f;
}
void launch_surr_pad(const SurrArgs& s, hipStream_t st) { k_surr_pad<<<1, 256, 0, st>>>(s); }

__global__ void k_gather_deferred(PBuf src, const double* __restrict__ k0s, const double* __restrict__ as_,
                                  const uint32_t* __restrict__ list, const unsigned long long* n_list, PBuf dst, double* k0d,
                                  double* ad, const unsigned long long* n_dst, uint64_t cap, uint32_t* now,
                                  unsigned long long* n_now) {
    const uint64_t j = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool have = j < *n_list;
    const uint64_t o = *n_dst + j;
    const uint32_t i = have ? list[j] : 0u;
    // the lanes without room take their places in the exact pass of this generation
    const bool spill = have && o >= cap;
    if (__ballot(spill)) {
        if (spill) now[wave_reserve(n_now, 1u)] = i;
    }
    if (!have || spill) return;
#pragma unroll
    for (int c = 0; c < P_NDOUBLE; ++c) dst.d[c][o] = src.d[c][i];
    dst.id[o] = src.id[i];
    dst.parent_id[o] = src.parent_id[i];
    dst.stream[o] = src.stream[i];
    dst.shower[o] = src.shower[i];
    dst.meta[o] = src.meta[i];
    dst.nev[o] = src.nev[i];
    k0d[o] = k0s[i];
    ad[o] = as_[i];
}
void launch_gather_deferred(const PBuf& src, const double* k0s, const double* as_, const uint32_t* list,
                            const unsigned long long* n_list, uint64_t max_list, const PBuf& dst, double* k0d, double* ad,
                            const unsigned long long* n_dst, uint64_t cap, uint32_t* now, unsigned long long* n_now,
                            hipStream_t s) {
    if (max_list)
        k_gather_deferred<<<unsigned((max_list + 255) / 256), 256, 0, s>>>(src, k0s, as_, list, n_list, dst, k0d, ad, n_dst, cap,
                                                                          now, n_now);
}
__global__ void k_commit_deferred(const unsigned long long* n_list, unsigned long long* n_dst, uint64_t cap,
                                  const unsigned long long* n_now, unsigned long long* stats) {
    if (threadIdx.x != 0) return;
    const unsigned long long n = *n_dst + *n_list;
    const unsigned long long spilled = n > cap ? n - cap : 0ull;
    *n_dst = n - spilled;
    stats[STAT_SURR_FB] += *n_list + *n_now - spilled;
}
void launch_commit_deferred(const unsigned long long* n_list, unsigned long long* n_dst, uint64_t cap,
                            const unsigned long long* n_now, unsigned long long* stats, hipStream_t s) {
    k_commit_deferred<<<1, 64, 0, s>>>(n_list, n_dst, cap, n_now, stats);
}

// Probes of the LMA device functions (parity tests).
__global__ void k_eval_lma(int fn, const double* in, double* out, uint32_t n, uint32_t in_w, uint32_t out_w, LmaArgs m,
                           uint64_t key) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* tables = m.tables;
    const LmaScalars& sc = m.sc;
    const double* bw = m.bw;
    const BwScalars& bsc = m.bsc;
    const double* x = in + uint64_t(i) * in_w;
    double* y = out + uint64_t(i) * out_w;
    switch (fn) {
    case PTG_FN_LP_RATE: {
        double f = 0.0;
        y[1] = lp_rate(tables + lma_pack::RATE, sc, x[0], x[1], f) ? 1.0 : 0.0;
        y[0] = f;
        break;
    }
    case PTG_FN_LP_HARMONIC: y[0] = lp_harmonic(tables + lma_pack::CDF, sc, x[0], x[1], x[2]); break;
    case PTG_FN_DOUBLE_BESSEL: {
        // mode x[3]: 0 the kernels' evaluation (saddle point or quadrature), 1 the
        // reference's recurrence (the host checks 3 <= n, |x| <= n sqrt 2,
        // 0 <= y <= n / 2 and sizes m.scratch for it), 2 the quadrature alone,
        // 3 the reference's evaluation (saddle point or recurrence), 4 the
        // quadrature shared by a wave (k_eval_dbessel_wave)
        if (x[3] == 4.0) break;
        const int nn = int(x[0]);
        const DBesselRec R = dbessel_rec_at(nn, m.scratch + i, n);
        double j[5];
        if (x[3] == 1.0) dbessel_recurrence(R, x[1], x[2], j);
        else if (x[3] == 2.0) dbessel_quadrature(nn, x[1], x[2], j);
        else if (x[3] == 3.0) dbessel_eval(R, x[1], x[2], j);
        else dbessel_eval(dbessel_of(nn), x[1], x[2], j);
        for (int q = 0; q < 5; ++q) y[q] = j[q];
        break;
    }
    case PTG_FN_BW_RATE: {
        double f[2] = {0.0, 0.0};
        double a = x[0], eta = x[1];
        if (bw_is_negligible(a, eta)) y[2] = 0.0;
        else if (bw_mid_contains(bsc, a, eta)) { bw_mid_interpolate(bw + bw_pack::MID_TOTAL, bsc, a, eta, f); y[2] = 1.0; }
        else if (bw_hi_contains(bsc, a, eta)) y[2] = bw_hi_interpolate(bw + bw_pack::HI_TOTAL, bsc, a, eta, f) ? 2.0 : -1.0;
        else y[2] = -1.0;
        y[0] = f[0]; y[1] = f[1];
        break;
    }
    case PTG_FN_BW_HARMONIC: y[0] = bw_harmonic(bw, bsc, x[0], x[1], x[2], x[3]); break;
    case PTG_FN_LP_CL_RATE: {
        double f = 0.0;
        y[1] = lp_cl_rate(tables, sc, x[0], x[1], f) ? 1.0 : 0.0;
        y[0] = f;
        break;
    }
    case PTG_FN_LP_CL_HARMONIC: y[0] = lp_cl_harmonic(tables, sc, x[0], x[1]); break;
    case PTG_FN_LP_CL_CEILING: {
        const int nn = int(x[0]);  // the host checks 1 <= n
        y[0] = lp_cl_ceiling(dbessel_of(nn), x[1]);
        break;
    }
    case PTG_FN_CPBW_RATE: {
        double f[2] = {0.0, 0.0};
        double a = x[0], eta = x[1];
        bool ok = cpbw_total(m.lcfa, m.cp, m.csc, a, eta, f);
        y[0] = f[0]; y[1] = f[1];
        y[2] = !ok ? -1.0 : ((a < 0.02 || cpbw_is_too_small(a, eta)) ? 0.0 : (a < exp(m.csc.bw_ln_min_a) ? 1.0 : 2.0));
        break;
    }
    case PTG_FN_CPBW_PARTIAL_RATE: {
        cplx r = cpbw_partial_rate(m.lcfa, m.cp + cp_pack::G32N, m.cp + cp_pack::G32W, int(x[0]), x[1], x[2]);
        y[0] = r.re; y[1] = r.im;
        break;
    }
    case PTG_FN_BESSEL_J: {
        double j[3];
        bessel_j_pm(m.lcfa, int(x[0]), x[1], j);  // the host checks 1 <= n <= 1e7, x >= 0
        y[0] = j[0]; y[1] = j[1]; y[2] = j[2];
        break;
    }
    case PTG_FN_CP_RATE: {
        double f = 0.0;
        bool ok = x[2] != 0.0 ? cp_classical_rate(m.cp, m.csc, x[0], x[1], f) : cp_rate(m.cp + cp_pack::RATE, m.cp, m.csc, x[0], x[1], f);
        y[0] = f;
        y[1] = ok ? 1.0 : 0.0;
        break;
    }
    case PTG_FN_CP_HARMONIC:
        y[0] = x[3] != 0.0 ? cp_classical_harmonic(m.cp, m.csc, x[0], x[2]) : cp_harmonic(m.cp, m.csc, x[0], x[1], x[2]);
        break;
    case PTG_FN_CP_PARTIAL_RATE: {
        int nn = int(x[0]);  // the host checks 1 <= n <= 1e6
        if (x[3] != 0.0) {
            y[0] = cp_integrated_spectrum_low_eta(m.lcfa, m.cp + cp_pack::G32N, m.cp + cp_pack::G32W, nn, x[1]);
            y[1] = cp_classical_ceiling(m.lcfa, nn, x[1]);
        } else {
            y[0] = cp_partial_rate(m.lcfa, m.cp + cp_pack::G32N, m.cp + cp_pack::G32W, nn, x[1], x[2]);
            y[1] = cp_ceiling(m.lcfa, m.cp + cp_pack::G16N, nn, x[1], x[2]);
        }
        break;
    }
    default: break;
    }
}
// The Breit-Wheeler sampler probe, one row per wave through the pair kernel's
// sampler: (a, eta, S1, S2, stream, fixed_n) -> (n, s, theta, trials), trials
// -2 outside the tables.
__global__ __launch_bounds__(64) void k_eval_bw_sample(const double* in, double* out, uint32_t n, uint32_t in_w, uint32_t out_w,
                                                       LmaArgs m, uint64_t key) {
    __shared__ double tb_s[16], tb_f[16];
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        const double* x = in + uint64_t(i) * in_w;
        double* y = out + uint64_t(i) * out_w;
        Rng rng = rng_make(key, uint64_t(x[4]));
        const int nh = x[5] > 0.0 ? int(x[5]) : bw_harmonic(m.bw, m.bsc, x[0], x[1], x[2], rng_uniform(rng));
        if (nh < 1) {
            if (threadIdx.x == 0) { y[0] = 0.0; y[3] = -2.0; }
            continue;
        }
        double s, th;
        uint32_t tr = 0;
        bool capped = false;
        bw_sample_wave(dbessel_of(nh), x[0], x[1], x[2], x[3], rng, tb_s, tb_f, s, th, tr, capped);
        if (threadIdx.x == 0) { y[0] = nh; y[1] = s; y[2] = th; y[3] = tr; }
    }
}

// The LP emission sampler probes (quantum and classical), one row per wave
// through the emission kernel's sampler: (a, eta, stream, fixed_n) -> (n, s,
// theta, S1, S2, trials), trials -2 outside the table.
template <bool classical>
__global__ __launch_bounds__(64) void k_eval_lp_sample(const double* in, double* out, uint32_t n, uint32_t in_w, uint32_t out_w,
                                                       LmaArgs m, uint64_t key) {
    __shared__ double tb_s[16], tb_f[16];
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        const double* x = in + uint64_t(i) * in_w;
        double* y = out + uint64_t(i) * out_w;
        Rng rng = rng_make(key, uint64_t(x[2]));
        int nh = int(x[3]);
        if (!(x[3] > 0.0)) {
            const double fr = rng_uniform(rng);
            nh = classical ? lp_cl_harmonic(m.tables, m.sc, x[0], fr) : lp_harmonic_wave(m.tables + lma_pack::CDF, m.sc, x[0], x[1], fr);
        }
        if (nh < 1) {
            if (threadIdx.x == 0) { y[0] = 0.0; y[5] = -2.0; }
            continue;
        }
        double s, th;
        sv4 pol;
        uint64_t tr = 0;
        uint32_t flags = 0;
        uint64_t restarts = 0;
        lp_sample_wave<classical>(nh, x[0], x[1], rng, tb_s, tb_f, m.lp_ceil, s, th, pol, tr, flags, restarts);
        if (threadIdx.x == 0) {
            y[0] = nh; y[1] = s; y[2] = th; y[3] = pol.q; y[4] = pol.u; y[5] = double(tr);
            if (out_w > 6) y[6] = double(restarts);
        }
    }
}

// The LP ceiling probe, one row per wave: (n, a, eta) -> the largest of the
// reference's ceiling samples of the double-differential rate (lp/mod.rs:53-140,
// without its factor 1.15; J(n, x, y) shared by the wave as in the sampler), and
// the tabulated value at (n, a, eta), -1 where the table does not cover it.
__global__ __launch_bounds__(64) void k_eval_lp_ceiling(const double* in, double* out, uint32_t n, uint32_t in_w, uint32_t out_w,
                                                        LmaArgs m) {
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        const double* x = in + uint64_t(i) * in_w;
        double* y = out + uint64_t(i) * out_w;
        const int nh = int(x[0]);
        const double a = x[1], eta = x[2];
        const DBesselWave W{nh};
        const int mm = lp_ceiling_count(nh, a, eta);
        double f = -INFINITY;
        for (int k = 0; k < mm; ++k) {
            double sk, tk;
            lp_ceiling_arg(nh, a, eta, k, sk, tk);
            f = fmax(f, lp_ddpr(W, a, eta, sk, tk));
        }
        double t = -1.0;
        if (!(m.lp_ceil && lp_ceiling_lookup(m.lp_ceil, nh, a, eta, t))) t = -1.0;
        if (threadIdx.x == 0) { y[0] = f; y[1] = t; }
    }
}

// The CP emission sampler probe, one row per wave through the emission
// kernel's sampler: (a, eta, stream, fixed_n, classical) -> (n, s, theta, S1,
// S2, S3, trials), trials -2 outside the table, -1 at the trial cap.
__global__ __launch_bounds__(64) void k_eval_cp_sample(const double* in, double* out, uint32_t n, uint32_t in_w, uint32_t out_w,
                                                       LmaArgs m, uint64_t key) {
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        const double* x = in + uint64_t(i) * in_w;
        double* y = out + uint64_t(i) * out_w;
        Rng rng = rng_make(key, uint64_t(x[2]));
        const bool cl = x[4] != 0.0;
        int nh = int(x[3]);
        if (!(x[3] > 0.0)) {
            const double fr = rng_uniform(rng);
            nh = cl ? cp_classical_harmonic(m.cp, m.csc, x[0], fr) : cp_harmonic_wave(m.cp, m.csc, x[0], x[1], fr);
        }
        if (nh < 1) {
            if (threadIdx.x == 0) { y[0] = 0.0; y[6] = -2.0; }
            continue;
        }
        double s, th;
        sv4 sv;
        uint64_t tr = 0;
        uint32_t flags = 0;
        if (cl) cp_sample_wave<true>(m.lcfa, m.cp, nh, x[0], x[1], rng, s, th, sv, tr, flags);
        else cp_sample_wave<false>(m.lcfa, m.cp, nh, x[0], x[1], rng, s, th, sv, tr, flags);
        if (threadIdx.x == 0) {
            y[0] = nh; y[1] = s; y[2] = th; y[3] = sv.q; y[4] = sv.u; y[5] = sv.v;
            y[6] = (flags & PTG_FLAG_REJECT_CAP) ? -1.0 : double(tr);
        }
    }
}

// The CP Breit-Wheeler harmonic and (n, s, theta) probes, one row per wave as the
// pair-event kernel: the harmonic by the wave's shared sum (cpbw_harmonic_wave),
// (s, theta) by cpbw_sample_at on every lane.
__global__ __launch_bounds__(64) void k_eval_cpbw_wave(int fn, const double* in, double* out, uint32_t n, uint32_t in_w,
                                                       uint32_t out_w, LmaArgs m, uint64_t key) {
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        const double* x = in + uint64_t(i) * in_w;
        double* y = out + uint64_t(i) * out_w;
        bool fb = false;
        if (fn == PTG_FN_CPBW_HARMONIC) {
            const int nh = cpbw_harmonic_wave(m.lcfa, m.cp, m.csc, x[0], x[1], x[2], x[3], fb);
            if (threadIdx.x == 0) { y[0] = nh; y[1] = fb ? 1.0 : 0.0; }
            continue;
        }
        Rng rng = rng_make(key, uint64_t(x[5]));
        const int nh = x[6] > 0.0 ? int(x[6]) : cpbw_harmonic_wave(m.lcfa, m.cp, m.csc, x[0], x[1], x[4], rng_uniform(rng), fb);
        if (nh < 1) {
            if (threadIdx.x == 0) { y[0] = 0.0; y[3] = -2.0; }
            continue;
        }
        double sv, th;
        uint32_t tr = 0;
        bool capped = false;
        cpbw_sample_at(m.lcfa, m.cp, nh, x[0], x[1], x[2], x[3], x[4], rng, sv, th, tr, capped);
        if (threadIdx.x == 0) { y[0] = nh; y[1] = sv; y[2] = th; y[3] = capped ? -1.0 : double(tr); }
    }
}

// The J(n, x, y) probe in mode 4: the quadrature shared by a wave
// (dbessel_quadrature_wave), one row per wave; every lane's result is checked
// against lane 0's and a disagreement returns NaN.
__global__ __launch_bounds__(64) void k_eval_dbessel_wave(const double* in, double* out, uint32_t n, uint32_t in_w,
                                                          uint32_t out_w) {
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        const double* x = in + uint64_t(i) * in_w;
        double* y = out + uint64_t(i) * out_w;
        if (x[3] != 4.0) continue;
        double j[5];
        dbessel_quadrature_wave(int(x[0]), x[1], x[2], j);
        bool same = true;
        for (int q = 0; q < 5; ++q) same = same && __shfl(j[q], 0, 64) == j[q];
        same = __all(same);
        if (threadIdx.x == 0)
            for (int q = 0; q < 5; ++q) y[q] = same ? j[q] : NAN;
    }
}

void launch_eval_lma(int fn, const double* in, double* out, uint32_t n, uint32_t in_w, uint32_t out_w, const LmaArgs& m, uint64_t key,
                     hipStream_t s) {
    if (!n) return;
    const uint32_t rows = std::min<uint32_t>(n, 65536u);
    if (fn == PTG_FN_BW_SAMPLE) {
        k_eval_bw_sample<<<rows, 64, 0, s>>>(in, out, n, in_w, out_w, m, key);
        return;
    }
    if (fn == PTG_FN_CP_SAMPLE) {
        k_eval_cp_sample<<<rows, 64, 0, s>>>(in, out, n, in_w, out_w, m, key);
        return;
    }
    if (fn == PTG_FN_CPBW_HARMONIC || fn == PTG_FN_CPBW_SAMPLE) {
        k_eval_cpbw_wave<<<rows, 64, 0, s>>>(fn, in, out, n, in_w, out_w, m, key);
        return;
    }
    if (fn == PTG_FN_LP_CEILING) {
        k_eval_lp_ceiling<<<rows, 64, 0, s>>>(in, out, n, in_w, out_w, m);
        return;
    }
    if (fn == PTG_FN_LP_SAMPLE || fn == PTG_FN_LP_CL_SAMPLE) {
        if (fn == PTG_FN_LP_SAMPLE) k_eval_lp_sample<false><<<rows, 64, 0, s>>>(in, out, n, in_w, out_w, m, key);
        else k_eval_lp_sample<true><<<rows, 64, 0, s>>>(in, out, n, in_w, out_w, m, key);
        return;
    }
    k_eval_lma<<<(n + 63) / 64, 64, 0, s>>>(fn, in, out, n, in_w, out_w, m, key);
    if (fn == PTG_FN_DOUBLE_BESSEL) k_eval_dbessel_wave<<<rows, 64, 0, s>>>(in, out, n, in_w, out_w);
}

}  // namespace ptg
