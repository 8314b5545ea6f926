// Arguments and launchers of the LMA (PlaneWave) kernels.
#pragma once
#include "dev_bw.h"
#include "dev_cp.h"
#include "dev_lma.h"
#include "kern_common.h"

namespace ptg {

// Emission model of the lepton kernels: linear or circular polarisation of the
// wave, quantum or classical radiation mode (nonlinear_compton/mod.rs:31-36)
enum LmaMode : int { LMA_LP = 0, LMA_CP = 1, LMA_CP_CLASSICAL = 2, LMA_LP_CLASSICAL = 3 };

struct LmaArgs {
    const double* tables;        // LMA pack (dev_lma.h)
    LmaScalars sc;
    const double* lp_ceil;       // tabulated LP emission ceiling (lp_ceil:: in dev_lma.h), or null
    const double* cp;            // CP pack (dev_cp.h)
    CpScalars csc;
    const double* lcfa;          // LCFA pack (Airy functions for J_n at n >= 1000)
    int mode;                    // LmaMode
    int mixed;                   // mixed precision: FP32 rate table and pusher algebra (LP, CP quantum)
    const float* tables_f;       // FP32 copy of the LMA pack (mixed precision: its rate table)
    // a batch of collisions: laser of collision coll[shower] for each particle (null: A->laser)
    const LaserDev* lasers;
    const uint8_t* coll;
    double bandwidth;            // fractional bandwidth (plane_wave.rs:49-63); 0 = monochromatic
    double pol_angle, sin_pa, cos_pa;
    // step kernel -> event queue
    PBuf ev;
    double* ev_k0;               // kappa^0 at the event (kappa = k0 (1, 0, 0, 1))
    double* ev_a;                // sqrt(<a^2>) at the event position
    unsigned long long* ev_ctr;
    unsigned long long ev_cap;
    // event kernel input: events ev_list[0..n_ev) of ev_in, or 0..n_ev when ev_list is null;
    // n_ev is read from *n_ev_dev when that is set (generations run from a captured graph)
    PBuf ev_in;
    const uint32_t* ev_list;
    uint64_t n_ev;
    const unsigned long long* n_ev_dev;
    // event kernel outputs
    PBuf next;                   // electrons continuing after their event
    unsigned long long* next_ctr;
    PBuf ph;                     // emitted photons when photons are tracked
    unsigned long long* ph_ctr;
    unsigned long long ph_cap;
    int photons_to_out;          // photons are final state (photon tracking off)
    double* scratch;             // the J(n, x, y) probe's recurrence workspace (k_eval_lma only)
    // LP Breit-Wheeler (photon phase)
    const double* bw;            // BW pack (dev_bw.h)
    BwScalars bsc;
    PBuf pairs;                  // electrons and positrons created by photons
    unsigned long long* pairs_ctr;
    unsigned long long pairs_cap;
};

// The learned sampler (surrogate) of LMA emission: inputs built from queued
// events, outputs turned into photons; training data from the tables and the
// exact sampler.
struct SurrArgs {
    double la_c, la_h, le_c, le_h;   // domain box: centre and half width of ln a and ln eta
    double la_use, le_use;           // ln a, ln eta below which events are sampled exactly in this generation
    // inference
    float* x;                        // [rows][8] network inputs
    const float* out;                // [rows][32] network outputs
    uint32_t* list;                  // row -> event index (events the network samples)
    unsigned long long* n_rows;
    uint32_t* fb;                    // events left to the exact sampler's deferred pass
    unsigned long long* n_fb;
    uint32_t* now;                   // events below the box's amplitude: the exact sampler in this generation
    unsigned long long* n_now;
    unsigned long long n_ev;         // events queued by the step kernel
    const unsigned long long* n_ev_dev;  // or the count in device memory (captured generations)
    uint32_t* perm;                  // rows ordered by harmonic bucket (largest first), or null
    unsigned int* bucket;            // [2][32] bucket counts and cursors
    // training data
    uint64_t key;                    // Philox key of the training draws
    uint32_t n_groups;
    double* gh;                      // [groups][4]: ln a, ln eta, u1, nu (nu = 0: no samples)
    float* gs;                       // [groups][64][2]: v = s / s_max, t = 2 theta / pi
    const float* rows;               // [n_sampler_rows][8]: ln a, ln eta, u1, u2, y_nu, v, t, 0
    uint32_t n_sampler_rows;
    float* tx;                       // batch inputs [batch][8]
    float* tt;                       // batch targets [batch][8]
    int batch, n_table_rows;
    uint32_t step;
};
void launch_surr_inputs(const TransportArgs* a, const LmaArgs& m, const SurrArgs& s, hipStream_t st);
void launch_surr_event(const TransportArgs* a, const LmaArgs& m, const SurrArgs& s, int lanes, hipStream_t st);
// orders the rows (n_rows on the host) by their harmonic's power-of-two bucket, largest first
void launch_surr_order(const SurrArgs& s, uint32_t n_rows, hipStream_t st);
// the same with the row count *s.n_rows in device memory, at most max_rows
void launch_surr_order_dev(const SurrArgs& s, uint32_t max_rows, hipStream_t st);
// zeroes the network input rows from *s.n_rows up to the next multiple of 32
void launch_surr_pad(const SurrArgs& s, hipStream_t st);
void launch_surr_groups(const LmaArgs& m, const SurrArgs& s, int lanes, hipStream_t st);
void launch_surr_batch(const LmaArgs& m, const SurrArgs& s, hipStream_t st);
void launch_gather_events(const PBuf& src, const double* k0s, const double* as_, const uint32_t* list, uint64_t n, const PBuf& dst,
                          double* k0d, double* ad, uint64_t off, hipStream_t s);
// Generations run from a captured graph: the deferred events of one generation
// (list[0..*n_list)) appended to dst at *n_dst while dst has room (capacity
// cap); those beyond it go to the exact pass of the same generation
// (now[*n_now ...]). The count is added to *n_dst by launch_commit_deferred,
// after the gather. max_list bounds *n_list.
void launch_gather_deferred(const PBuf& src, const double* k0s, const double* as_, const uint32_t* list,
                            const unsigned long long* n_list, uint64_t max_list, const PBuf& dst, double* k0d, double* ad,
                            const unsigned long long* n_dst, uint64_t cap, uint32_t* now, unsigned long long* n_now,
                            hipStream_t s);
// ... and the events sampled exactly in this generation (the deferred and the
// spilled ones) are added to stats[STAT_SURR_FB].
void launch_commit_deferred(const unsigned long long* n_list, unsigned long long* n_dst, uint64_t cap,
                            const unsigned long long* n_now, unsigned long long* stats, hipStream_t s);

// Start of a captured generation. Unless an earlier generation of the sequence
// stalled: its input count (ctr[C_LNEXT], what the last generation queued) goes
// to A->n_in and ctr[C_NCUR], the generation is counted in ctr[C_GENS] when it
// has input, and C_WORK, C_EV and C_LNEXT are zeroed; after a stall A->n_in = 0
// and the queue counters are left as they are. Either way C_EVRUN and the
// learned sampler's C_SROWS, C_SFB, C_SNOW are zeroed, and bucket (64 words)
// when given.
void launch_gen_begin(unsigned long long* ctr, TransportArgs* A, unsigned int* bucket, hipStream_t s);

// Room the events of a captured generation need: per event at most out_per
// final-state records, ph_per photons queued at ctr[ph_slot], pairs_per pair
// records and inter_per decayed photons; the capacities of those queues.
struct EventRoom {
    unsigned long long out_cap, ph_cap, pairs_cap, inter_cap;
    uint32_t out_per, ph_per, pairs_per, inter_per;
    int ph_slot;
};
// Between the step kernel and the event kernels of a captured generation:
// ctr[C_EVRUN] = ctr[C_EV] when the queues have room for that many events;
// otherwise C_EVRUN = 0 and C_STALL = 1, the events stay queued (C_EV) for the
// host, and the rest of the sequence does nothing.
void launch_event_room(unsigned long long* ctr, const EventRoom& k, hipStream_t s);

// Fills the CP pack's CDF_LN and CL_CDF_LN from its CDF rows (once, after upload).
void launch_cp_cdf_logs(double* cp, hipStream_t s);
// Fills the LMA pack's CDF_LN from its CDF rows (once, after upload).
void launch_lp_cdf_logs(double* lma, hipStream_t s);

// block: 64, 128 or 256 threads (the kernel's bound); grid: persistent blocks
void launch_step_lma(const TransportArgs* a, const LmaArgs& m, int grid, hipStream_t s, int block = 256);
void launch_event_lma(const TransportArgs* a, const LmaArgs& m, int lanes, hipStream_t s);
void launch_photon_step_lma(const TransportArgs* a, const LmaArgs& m, int grid, hipStream_t s);
void launch_pair_event_lma(const TransportArgs* a, const LmaArgs& m, int lanes, hipStream_t s);
void launch_push_lma(const LaserDev& L, double dt, int ll, const uint8_t* sp_, double* r4, double* u4, uint32_t n,
                     uint32_t n_steps, uint32_t* taken, double* work, hipStream_t s);
// probes: the table pointers, scalars and scratch (the recurrence probe's J(n, x, y)
// workspace, dbessel_len(n) orders per row) are taken from m
void launch_eval_lma(int fn, const double* in, double* out, uint32_t n, uint32_t in_w, uint32_t out_w, const LmaArgs& m, uint64_t key,
                     hipStream_t s);

}  // namespace ptg
