// Shared pieces of the transport kernels: kernel arguments, output of final
// states and secondaries, statistics.
#pragma once
#include "../../include/ptg.h"
#include "dev_common.h"
#include "dev_fields.h"
#include "dev_lcfa.h"
#include "dev_particles.h"
#include "dev_tables.h"

namespace ptg {

enum { STAT_LEPTON_STEPS = 0, STAT_PHOTON_STEPS, STAT_EMISSIONS, STAT_PAIRS, STAT_TRIALS, STAT_PHOTON_STEPS_CF, STAT_SURROGATE,
       STAT_SURR_FB, STAT_CEIL_RESTART, STAT_N };

// Counter block of a context (unsigned long long words): work and queue counters
// of the transport, then the statistics. C_NCUR, C_GENS, C_NDEF, C_EVRUN and
// C_STALL belong to generations run from a captured graph: the input count of
// the generation running, the generations that ran their step kernel, the
// events held for the learned sampler's deferred exact pass, the events the
// event kernels of the generation take, and the stall flag (a generation found
// too little room in the output queues for its events: its events stay queued
// and the rest of the sequence does nothing).
enum { C_WORK = 0, C_NEXT, C_OUT, C_INTER, C_ID, C_FLAGS, C_EV, C_LNEXT, C_PH, C_PAIRS, C_SROWS, C_SFB, C_SNOW,
       C_NCUR, C_GENS, C_NDEF, C_EVRUN, C_STALL, C_STATS, C_N = C_STATS + STAT_N };

// CollideOptions as the kernels need them (src/main.rs:72-105).
struct DevOptions {
    double rate_increase, t_stop, min_energy, cos_max_angle;
    int tracking_photons, tracking_secondaries, keep_decayed_photons, pol_resolved;
    int rotate_stokes_pars, discard_bg_e, discard_bg_ph, multiplicity_on;
};

struct TransportArgs {
    LaserDev laser;
    LcfaScalars sc;
    DevOptions opt;
    double dt;
    int rr_mode;          // equation of motion: 0 Lorentz, 1 LL, 2 modified LL (src/main.rs:143-151)
    int classical_mode;   // RadiationMode::Classical (src/main.rs:153-157)
    int recoil;           // electron_recoils (src/main.rs:159)
    int precision;        // ptg_precision
    uint32_t step_cap;    // steps after which a particle is stopped and PTG_FLAG_STEP_CAP raised
    // Photons with energy below ff_kt_max (units of m c^2) cannot reach chi = 0.01 anywhere
    // in the field, where the pair rate is exactly zero (lcfa/pair_creation/mod.rs:13-18);
    // after ff_after steps their remaining straight-line steps are taken in closed form.
    double ff_kt_max;
    uint32_t ff_after;
    uint64_t key;         // RNG key derived from control:rng_seed
    const double* lcfa_tables;
    const double* lma_tables;
    PBuf in;
    uint32_t n_in;
    uint32_t* work_ctr;
    PBuf out;             // final-state particles (Shower::secondaries)
    unsigned long long* out_ctr;
    PBuf inter;           // decayed photons (Shower::intermediates)
    unsigned long long* inter_ctr;
    PBuf next;            // secondaries to be transported by the next kernel
    unsigned long long* next_ctr;
    unsigned long long next_cap;
    unsigned long long* id_ctr;
    uint64_t id_base;
    const double* prim_n0;      // [3][n_prim] unit momentum direction of each primary
    uint32_t n_prim;
    unsigned long long* stats;
    uint32_t* flags;
};

__device__ __forceinline__ void wave_add(unsigned long long* p, uint64_t v) {
    if (v) atomicAdd(p, (unsigned long long)v);
}

// Lane-uniform reads: the builtins take and return int, which must not be
// sign-extended when widened to 64 bits.
__device__ __forceinline__ uint32_t readlane_u32(uint32_t v, int lane) {
    return uint32_t(__builtin_amdgcn_readlane(int(v), lane));
}
__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, int lane) {
    return (uint64_t(readlane_u32(uint32_t(v >> 32), lane)) << 32) | uint64_t(readlane_u32(uint32_t(v), lane));
}

// Slot reservation with one atomic per wave and counter: every pass of the loop
// takes the lowest lane still waiting as leader, reads its counter and count
// (readlane at a lane index that is uniform, being derived from a ballot), and
// serves every waiting lane that names the same pair: their ranks come from mbcnt
// over the ballot of matches, the leader adds the group's total and the base is
// read back from the leader. The loop runs while a ballot of waiting lanes is
// non-zero, so its control flow is uniform and every ballot inside it sees the
// lanes it was written for (lanes reaching this point from different call sites,
// which the compiler may merge, are served in turn; when all agree one pass
// does). Each pass serves at least its leader, so 64 passes serve a whole wave;
// the per-lane atomic after them only bounds the loop.
// Slots within a wave follow lane order; across waves they follow the order in
// which the waves' atomics land.
__device__ __forceinline__ uint64_t wave_reserve(unsigned long long* ctr, uint32_t cnt) {
    const uint64_t addr = reinterpret_cast<uint64_t>(ctr);
    const uint32_t lane = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
    bool pending = true;
    uint64_t slot = 0;
    for (int pass = 0; pass < 64; ++pass) {
        const uint64_t waiting = __ballot(pending);
        if (waiting == 0) break;
        const int leader = __ffsll((long long)waiting) - 1;
        const uint64_t lead_addr = readlane_u64(addr, leader);
        const uint32_t lead_cnt = readlane_u32(cnt, leader);
        const bool match = pending && addr == lead_addr && cnt == lead_cnt;
        const uint64_t group = __ballot(match);
        const uint32_t rank = __builtin_amdgcn_mbcnt_hi(uint32_t(group >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(group), 0u));
        unsigned long long base = 0;
        if (lane == uint32_t(leader)) base = atomicAdd(ctr, (unsigned long long)(__popcll(group)) * lead_cnt);
        base = readlane_u64(base, leader);
        if (match) {
            slot = base + uint64_t(rank) * cnt;
            pending = false;
        }
    }
    if (pending) slot = atomicAdd(ctr, (unsigned long long)cnt);
    return slot;
}
// 32-bit counter form (work counters): same scheme.
__device__ __forceinline__ uint32_t wave_reserve32(uint32_t* ctr, uint32_t cnt) {
    const uint64_t addr = reinterpret_cast<uint64_t>(ctr);
    const uint32_t lane = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
    bool pending = true;
    uint32_t slot = 0;
    for (int pass = 0; pass < 64; ++pass) {
        const uint64_t waiting = __ballot(pending);
        if (waiting == 0) break;
        const int leader = __ffsll((long long)waiting) - 1;
        const uint64_t lead_addr = readlane_u64(addr, leader);
        const uint32_t lead_cnt = readlane_u32(cnt, leader);
        const bool match = pending && addr == lead_addr && cnt == lead_cnt;
        const uint64_t group = __ballot(match);
        const uint32_t rank = __builtin_amdgcn_mbcnt_hi(uint32_t(group >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(group), 0u));
        uint32_t base = 0;
        if (lane == uint32_t(leader)) base = atomicAdd(ctr, uint32_t(__popcll(group)) * lead_cnt);
        base = readlane_u32(base, leader);
        if (match) {
            slot = base + rank * cnt;
            pending = false;
        }
    }
    if (pending) slot = atomicAdd(ctr, cnt);
    return slot;
}

// The final energy/angle cut of collide() (src/main.rs:268-275).
__device__ __forceinline__ bool passes_cut(const TransportArgs& a, uint32_t shower, double e_mev, d3 p) {
    if (!(e_mev > a.opt.min_energy)) return false;
    d3 n0{a.prim_n0[shower], a.prim_n0[a.n_prim + shower], a.prim_n0[2 * a.n_prim + shower]};
    double inv = 1.0 / sqrt(nsq(p));
    return dot(n0, p) * inv > a.opt.cos_max_angle;
}

__device__ __forceinline__ void copy_record(const PBuf& dst, uint64_t j, const PBuf& src, uint32_t i) {
#pragma unroll
    for (int c = P_R0T; c <= P_PCHI; ++c) dst.d[c][j] = src.d[c][i];
    dst.d[P_WEIGHT][j] = src.d[P_WEIGHT][i];
    dst.id[j] = src.id[i];
    dst.parent_id[j] = src.parent_id[i];
    dst.stream[j] = src.stream[i];
    dst.shower[j] = src.shower[i];
    dst.meta[j] = src.meta[i];
    dst.nev[j] = 0;
}

__device__ __forceinline__ void finalize_lepton(const TransportArgs& a, uint32_t i, d4 r, double ut, d3 u, double work,
                                                double icount, uint32_t meta) {
    bool primary = meta & META_PRIMARY;
    if (primary && a.opt.discard_bg_e && !(icount > 0.0)) return;
    uint32_t shower = a.in.shower[i];
    if (!passes_cut(a, shower, M_E_MEV * ut, u)) return;
    uint64_t j = wave_reserve(a.out_ctr, 1u);
    const PBuf& o = a.out;
    copy_record(o, j, a.in, i);
    o.d[P_RT][j] = r.t; o.d[P_RX][j] = r.x; o.d[P_RY][j] = r.y; o.d[P_RZ][j] = r.z;
    o.d[P_UT][j] = ut; o.d[P_UX][j] = u.x; o.d[P_UY][j] = u.y; o.d[P_UZ][j] = u.z;
    o.d[P_WORK][j] = work;
    o.d[P_ICOUNT][j] = icount;
}

__device__ __forceinline__ void finalize_photon_from(const TransportArgs& a, const PBuf& src, uint32_t i, d4 r, d4 k, sv4 pol,
                                                     double weight, double icount, uint32_t meta, bool decayed) {
    bool primary = meta & META_PRIMARY;
    const PBuf* o = nullptr;
    unsigned long long* ctr = nullptr;
    uint32_t shower = src.shower[i];
    if (!decayed) {
        if (primary && a.opt.discard_bg_ph) return;
        if (!passes_cut(a, shower, M_E_MEV * k.t, sp(k))) return;
        o = &a.out;
        ctr = a.out_ctr;
    } else {
        if (!a.opt.keep_decayed_photons) return;
        o = &a.inter;
        ctr = a.inter_ctr;
    }
    uint64_t j = wave_reserve(ctr, 1u);
    copy_record(*o, j, src, i);
    o->d[P_RT][j] = r.t; o->d[P_RX][j] = r.x; o->d[P_RY][j] = r.y; o->d[P_RZ][j] = r.z;
    o->d[P_UT][j] = k.t; o->d[P_UX][j] = k.x; o->d[P_UY][j] = k.y; o->d[P_UZ][j] = k.z;
    o->d[P_SI][j] = pol.i; o->d[P_SQ][j] = pol.q; o->d[P_SU][j] = pol.u; o->d[P_SV][j] = pol.v;
    o->d[P_WEIGHT][j] = weight;
    o->d[P_ICOUNT][j] = icount;
    o->d[P_WORK][j] = src.d[P_WORK][i];
}
__device__ __forceinline__ void finalize_photon(const TransportArgs& a, uint32_t i, d4 r, d4 k, sv4 pol, double weight,
                                                double icount, uint32_t meta, bool decayed) {
    finalize_photon_from(a, a.in, i, r, k, pol, weight, icount, meta, decayed);
}

// Writes one new particle into the secondary queue at slot j.
__device__ __forceinline__ void write_new(const PBuf& q, uint64_t j, uint32_t species, d4 r, d4 u, d4 u0, sv4 pol,
                                          double payload, double chi, double weight, uint64_t id, uint64_t parent,
                                          uint64_t stream, uint32_t shower) {
    q.d[P_RT][j] = r.t; q.d[P_RX][j] = r.x; q.d[P_RY][j] = r.y; q.d[P_RZ][j] = r.z;
    q.d[P_R0T][j] = r.t; q.d[P_R0X][j] = r.x; q.d[P_R0Y][j] = r.y; q.d[P_R0Z][j] = r.z;
    q.d[P_UT][j] = u.t; q.d[P_UX][j] = u.x; q.d[P_UY][j] = u.y; q.d[P_UZ][j] = u.z;
    q.d[P_U0T][j] = u0.t; q.d[P_U0X][j] = u0.x; q.d[P_U0Y][j] = u0.y; q.d[P_U0Z][j] = u0.z;
    q.d[P_SI][j] = pol.i; q.d[P_SQ][j] = pol.q; q.d[P_SU][j] = pol.u; q.d[P_SV][j] = pol.v;
    q.d[P_PAYLOAD][j] = payload;
    q.d[P_PCHI][j] = chi;
    q.d[P_ICOUNT][j] = 0.0;
    q.d[P_WORK][j] = 0.0;
    q.d[P_WEIGHT][j] = weight;
    q.id[j] = id;
    q.parent_id[j] = parent;
    q.stream[j] = stream;
    q.shower[j] = shower;
    q.meta[j] = species;
    q.nev[j] = 0;
}

// Photon spawned by an emission event (src/main.rs:172-189): Particle::create
// leaves the creation momentum at zero for photons.
__device__ __forceinline__ void emit_photon(const TransportArgs& a, d4 r, d4 k, sv4 pol, double a_eff, double chi,
                                            double weight, uint64_t parent, uint64_t stream, uint32_t shower,
                                            uint32_t& flags) {
    uint64_t j = wave_reserve(a.next_ctr, 1u);
    uint64_t id = a.id_base + wave_reserve(a.id_ctr, 1u);
    if (j >= a.next_cap) return;  // counted; the host grows the queue and repeats the launch
    write_new(a.next, j, 2u, r, k, d4{0.0, 0.0, 0.0, 0.0}, pol, a_eff, chi, weight, id, parent, stream, shower);
}

// Electron-positron pair from a decay event (src/main.rs:220-246).
__device__ __forceinline__ void emit_pair(const TransportArgs& a, d4 r, d3 ue, d3 up, double weight, double a_eff,
                                          double chi, uint64_t parent, uint64_t base_stream, uint32_t shower,
                                          uint32_t& flags) {
    uint64_t j = wave_reserve(a.next_ctr, 2u);
    uint64_t id = a.id_base + wave_reserve(a.id_ctr, 2u);
    if (j + 1 >= a.next_cap) return;
    sv4 unpol{1.0, 0.0, 0.0, 0.0};
    d4 u0{1.0, 0.0, 0.0, 0.0};
    write_new(a.next, j, 0u, r, d4{sqrt(1.0 + nsq(ue)), ue.x, ue.y, ue.z}, u0, unpol, a_eff, chi, weight, id, parent,
              rng_child(base_stream, 0), shower);
    write_new(a.next, j + 1, 1u, r, d4{sqrt(1.0 + nsq(up)), up.x, up.y, up.z}, u0, unpol, a_eff, chi, weight, id + 1,
              parent, rng_child(base_stream, 1), shower);
}

// Hands a particle of the other kind (photon in the lepton stage or vice versa)
// on to the next stage unchanged.
__device__ __forceinline__ void forward_record(const TransportArgs& a, uint32_t i) {
    uint64_t j = wave_reserve(a.next_ctr, 1u);
    if (j >= a.next_cap) return;
    const PBuf& q = a.next;
#pragma unroll
    for (int c = 0; c < P_NDOUBLE; ++c) q.d[c][j] = a.in.d[c][i];
    q.id[j] = a.in.id[i];
    q.parent_id[j] = a.in.parent_id[i];
    q.stream[j] = a.in.stream[i];
    q.shower[j] = a.in.shower[i];
    q.meta[j] = a.in.meta[i];
    q.nev[j] = a.in.nev[i];
}

// launchers (kern_*.hip)
// `a` is a device pointer to the launch arguments.
void launch_leptons_lcfa(const TransportArgs* a, int kind, bool cp, int grid, hipStream_t s);
void launch_photons_lcfa(const TransportArgs* a, int kind, bool cp, int grid, hipStream_t s);
int lcfa_blocks_per_cu(int kind, bool cp, bool photons);
void launch_push_lcfa(const LaserDev& L, double dt, int rr_mode, const uint8_t* sp_, double* r4, double* u4, uint32_t n,
                      uint32_t n_steps, uint32_t* taken, double* work, hipStream_t s);

}  // namespace ptg
