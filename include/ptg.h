/*
 * ptg.h -- C ABI of the MI355X particle-shower engine (libptg.so).
 *
 * Drop-in boundary for Ptarmigan's per-particle Monte Carlo loop
 * (tgblackburn/ptarmigan v1.5.0). The reference has no FFI: the path sits behind
 *   fn collide<F: Field, R: Rng>(field: &F, incident: Particle, rng: &mut R,
 *                                current_id: &mut u64, options: CollideOptions) -> Shower
 * (src/main.rs:134) and the `Field` trait (src/field/mod.rs:126-192), driven by
 * the chunked map in src/main.rs:1118-1135. Each entry point below names the
 * reference interface it replaces.
 *
 * Conventions
 *  - Plain C types only. Host buffers are owned by the caller; device buffers,
 *    tables and results are owned by the context until the next ptg_run/ptg_collide.
 *  - Four-vectors are n x 4 row-major doubles, as the reference's FourVector
 *    arrays ((ct, x, y, z) in metres; (gamma, u_x, u_y, u_z) normalised momentum;
 *    Stokes (I, Q, U, V)).
 *  - Every call returns PTG_OK (0) or a negative PTG_E_* code and never aborts;
 *    ptg_last_error() gives the message. Device-side consistency failures (the
 *    reference's assert!/panic! sites) are accumulated in ptg_counters.error_flags.
 *  - One context per host thread. Work runs on the context's HIP stream; calls
 *    return when their results are complete.
 */
#ifndef PTG_H
#define PTG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    PTG_OK = 0,
    PTG_E_ARG = -1,      /* invalid argument (null pointer, bad enum, inconsistent options) */
    PTG_E_TABLES = -2,   /* table blob missing or malformed */
    PTG_E_HIP = -3,      /* HIP runtime error */
    PTG_E_OOM = -4,      /* device allocation failed */
    PTG_E_DOMAIN = -5,   /* parameters outside the tabulated domain (reference: prints and returns None) */
    PTG_E_UNSUPPORTED = -6, /* feature not built into this library (stated in the message) */
    PTG_E_SANITY = -7,   /* a reference assert! would have fired (see error_flags) */
    PTG_E_STATE = -8     /* call out of order (e.g. ptg_run before ptg_set_laser) */
};

/* Device error-flag bits (ptg_counters.error_flags) */
enum {
    PTG_FLAG_REJECT_CAP = 1u << 0,   /* a rejection loop hit its trial cap and used the fallback */
    PTG_FLAG_TABLE_RANGE = 1u << 1,  /* a table lookup fell outside its domain (probability set to 0) */
    PTG_FLAG_SANITY = 1u << 2,       /* photon/pair construction failed the reference's sanity check */
    PTG_FLAG_STEP_CAP = 1u << 3      /* a particle hit the per-particle step cap */
};

typedef enum { PTG_ELECTRON = 0, PTG_POSITRON = 1, PTG_PHOTON = 2 } ptg_species;   /* particle/mod.rs:35-40 */

/* src/field/mod.rs:117-123 (enum Laser) */
typedef enum { PTG_PLANE_WAVE = 0, PTG_FOCUSED_LASER = 1, PTG_FAST_PLANE_WAVE = 2, PTG_FAST_FOCUSED_LASER = 3 } ptg_laser_kind;
typedef enum { PTG_COS_SQUARED = 0, PTG_FLATTOP = 1, PTG_GAUSSIAN = 2 } ptg_envelope;       /* field/mod.rs:43-48 */
typedef enum { PTG_LINEAR = 0, PTG_CIRCULAR = 1 } ptg_polarization;                        /* field/mod.rs:24-28 */

/* Mirrors the laser construction in src/main.rs:1042-1060. n_cycles is the
 * reference's n_cycles (for gaussian envelopes: c * fwhm_duration / wavelength). */
typedef struct {
    int32_t kind;            /* ptg_laser_kind */
    int32_t envelope;        /* ptg_envelope */
    int32_t polarization;    /* ptg_polarization */
    int32_t finite_bandwidth;/* control:bandwidth_correction (LMA fields only) */
    double a0;
    double wavelength;       /* m */
    double waist;            /* m (focused kinds only) */
    double n_cycles;
    double pol_angle;        /* rad, linear polarisation only */
    double chirp_b;          /* plane waves only */
} ptg_laser_desc;

/* Precision of the per-step arithmetic (mixed mode keeps phase accumulation,
 * momentum update and cumulative probability in FP64). */
typedef enum { PTG_PREC_FP64 = 0, PTG_PREC_MIXED = 1 } ptg_precision;

/* Field-for-field mirror of CollideOptions (src/main.rs:72-105). */
typedef struct {
    double dt_multiplier;
    double rate_increase;
    double t_stop;           /* s; +inf for none */
    double min_energy;       /* MeV */
    double max_angle;        /* rad */
    int64_t multiplicity;    /* -1 = None */
    int32_t discard_bg_e;
    int32_t discard_bg_ph;
    int32_t rr;
    int32_t tracking_photons;
    int32_t tracking_secondaries;
    int32_t keep_decayed_photons;
    int32_t pol_resolved;
    int32_t rotate_stokes_pars;
    int32_t classical;
    int32_t gaunt_factor;
    int32_t precision;       /* ptg_precision */
    int32_t surrogate;       /* 1: sample LMA emission with the loaded surrogate inside its domain
                                (ptg_surrogate_load); no counterpart in CollideOptions */
} ptg_collide_options;

/* Particle set, structure of arrays (particle/mod.rs:66-78). Any pointer except
 * species, r and u may be NULL on input (defaults: unpolarised, weight 1, ids 0..n-1). */
typedef struct {
    uint8_t* species;          /* n */
    double* r;                 /* n x 4 current position */
    double* u;                 /* n x 4 current normalised momentum */
    double* r0;                /* n x 4 position at creation */
    double* u0;                /* n x 4 momentum at creation */
    double* pol;               /* n x 4 Stokes parameters */
    double* payload;           /* n  a_eff at creation */
    double* parent_chi;        /* n */
    double* interaction_count; /* n */
    double* work;              /* n  absorbed energy / mc^2 */
    double* weight;            /* n */
    uint64_t* id;              /* n */
    uint64_t* parent_id;       /* n */
} ptg_particles;

/* Counters for one ptg_run. One particle-step is one iteration of collide()'s
 * electron/positron loop (src/main.rs:169) or photon loop (src/main.rs:214). */
typedef struct {
    uint64_t lepton_steps;
    uint64_t photon_steps;
    uint64_t emissions;
    uint64_t pair_events;
    uint64_t n_secondaries;     /* final-state particles held by the context */
    uint64_t n_intermediates;   /* decayed photons kept (keep_decayed_photons) */
    uint64_t generations;
    uint64_t rejection_trials;
    uint32_t error_flags;
    uint32_t graph_stalls;      /* captured sequences whose events outgrew a queue and finished on the host's launches */
    double kernel_ms;           /* device time of all transport kernels (HIP events on the context stream) */
    double lepton_ms;           /* ... of the electron/positron kernels */
    double photon_ms;           /* ... of the photon kernels */
    uint64_t lepton_launches;
    uint64_t photon_launches;
    uint64_t photon_steps_closed_form; /* photon steps with provably zero pair probability taken in closed
                                          form (not included in photon_steps) */
    uint64_t surrogate_events;  /* emissions sampled by the surrogate */
    uint64_t surrogate_fallbacks; /* emissions outside its domain, sampled exactly */
    uint64_t graph_launches;    /* captured sequences of generations launched (PTG_LAUNCH_GRAPH) */
    uint64_t ceiling_restarts;  /* LP emissions whose tabulated ceiling a candidate passed: sampled again under the
                                   reference's per-event ceiling */
} ptg_counters;

/* Beam description, mirroring BeamBuilder (src/particle/builder.rs:173-219) for
 * the normal-gamma case (particle/dstr.rs:161-176). */
typedef struct {
    uint64_t n;
    int32_t species;
    int32_t radial;            /* 0 = normal, 1 = truncated normal, 2 = uniform (dstr.rs:26-51) */
    double gamma, sigma;       /* mean Lorentz factor, standard deviation */
    double energy_chirp;       /* rho (sign as in builder.rs:117-120) */
    double radius, max_radius; /* sigma_xy (normal) or r_max (uniform); truncation radius */
    double length;             /* sigma_z, m */
    double rms_div;            /* rad */
    double angle;              /* collision angle, rad */
    double collision_plane_angle;
    double offset[3];
    double pol[4];             /* Stokes */
    double weight;
} ptg_beam_desc;

/* Quantities a histogram can bin: the accessors of src/output/functions.rs:10-263
 * (same names and definitions; momenta in MeV, lengths in m, angles in rad). */
typedef enum {
    PTG_Q_ANGLE_X = 0, PTG_Q_ANGLE_Y, PTG_Q_THETA, PTG_Q_POLAR_ANGLE, PTG_Q_PX, PTG_Q_PY, PTG_Q_PZ,
    PTG_Q_P_PERP, PTG_Q_R_PERP, PTG_Q_R_X, PTG_Q_R_Y, PTG_Q_P_MINUS, PTG_Q_P_PLUS, PTG_Q_GAMMA,
    PTG_Q_ENERGY, PTG_Q_UNIT, PTG_Q_X, PTG_Q_Y, PTG_Q_Z, PTG_Q_RADIUS, PTG_Q_BIRTH_A, PTG_Q_PARENT_CHI,
    PTG_Q_N_INTER, PTG_Q_ABSORPTION, PTG_Q_S1, PTG_Q_S2, PTG_Q_S3,
    PTG_Q_POL_X, PTG_Q_POL_Y,  /* polarization projected on the x / y axis (particle/mod.rs:344-358) */
    PTG_Q_WEIGHT, PTG_Q_HELICITY,  /* weight; weight times S_3 */
    PTG_Q_COUNT
} ptg_quantity;

/* Histogram weights (functions.rs:232-263). */
typedef enum { PTG_W_NUMBER = 0, PTG_W_ENERGY = 1, PTG_W_HELICITY = 2, PTG_W_POL_X = 3, PTG_W_POL_Y = 4 } ptg_weight_fn;

/* 1-D weighted histogram of the resident final-state particles of one species,
 * as Histogram::generate_1d (src/output/hgram.rs:209-323): bins span the global
 * min..max of the (optionally log-scaled) quantity, the particle at exactly max
 * is not binned, the total counts every weight. */
typedef struct {
    int32_t species;         /* ptg_species; 3 = the intermediate (decayed) photons */
    int32_t quantity;        /* ptg_quantity */
    int32_t weight_fn;       /* ptg_weight_fn */
    int32_t log_scale;       /* BinSpec::LogScaled: bin ln(quantity) */
    uint32_t nbins;          /* BinSpec::FixedNumber */
    int32_t fixed_range;     /* use min/max below instead of the data range */
    double min, max;         /* in/out: range used (of ln(q) when log_scale) */
    /* Range cut (Filter, src/output/mod.rs:60-140): with cut = 1 only particles
     * with cut_min <= cut_quantity < cut_max (default units) set the range and
     * are binned; the total still counts every particle of the species. The
     * 2-D histogram takes the cut of spec_x. */
    int32_t cut;
    int32_t cut_quantity;    /* ptg_quantity */
    double cut_min, cut_max;
    /* Unit conversion (Unit::convert, src/output/units.rs): the quantity is
     * multiplied by scale before the log and the binning; 0 means 1. */
    double scale;
} ptg_hist_spec;

/* A summary statistic's reductions over the resident final-state particles of
 * one species (SummaryStatistic, src/output/stats.rs:265-410): v = variable,
 * w = weight quantity (PTG_Q_UNIT for none), "in" = cut_min < cut_quantity <
 * cut_max (strict) or every particle when cut = 0. Default units (MeV, m, rad). */
typedef struct {
    int32_t species;         /* ptg_species; 3 = the intermediate (decayed) photons */
    int32_t variable;        /* ptg_quantity */
    int32_t weight;          /* ptg_quantity */
    int32_t cut;
    int32_t cut_quantity;    /* ptg_quantity */
    int32_t reserved;
    double cut_min, cut_max;
    double shift;            /* subtracted from v in the last two sums */
} ptg_stats_spec;
/* out[PTG_STATS_N]: sum v w (all), sum v w (in), sum w (in), min v w (in, +inf if
 * none), max v w (in, -inf if none), sum w cos v, sum w sin v, sum w (v - shift)^2,
 * sum w (v - shift) (in). */
#define PTG_STATS_N 9

typedef struct ptg_ctx ptg_ctx;

/* Creates a context on HIP device `device`, loading tables from the blob at
 * `table_path` (NULL = the blob installed next to libptg.so). */
int ptg_create(ptg_ctx** out, const char* table_path, int device);
void ptg_destroy(ptg_ctx* ctx);
const char* ptg_last_error(const ptg_ctx* ctx);

/* Selects the field: replaces the laser construction at src/main.rs:1042-1060. */
int ptg_set_laser(ptg_ctx* ctx, const ptg_laser_desc* laser);
/* Field::max_timestep / ideal_initial_z (src/field/mod.rs:130,138). */
int ptg_field_info(ptg_ctx* ctx, double* max_timestep, double* ideal_initial_z);
/* Field::energy (src/field/mod.rs:191): the pulse energy of `laser` in J for the
 * focused kinds, per unit area in J/m^2 for the plane waves (*per_area = 1), as
 * fast_focused_laser.rs:220-261 and fast_plane_wave.rs:197-240 (PlaneWave and
 * FocusedLaser use the same formulas, plane_wave.rs:327-333, focused_laser.rs:317-323);
 * a chirped plane wave integrates its on-axis energy density (fast_plane_wave.rs:56-85).
 * Host arithmetic only: needs no context or device. */
int ptg_laser_energy(const ptg_laser_desc* laser, double* energy, int32_t* per_area);

/* Generates `beam->n` primaries on the device: replaces BeamBuilder::build
 * (src/particle/builder.rs:173-219). ids are 0..n-1. */
int ptg_build_beam(ptg_ctx* ctx, const ptg_beam_desc* beam, double initial_z, uint64_t seed);
/* The energy spectrum of the beams built next (GammaDistribution, particle/dstr.rs:60-214):
 * kind 0 the normal spectrum of ptg_beam_desc (the default); 1 a tabulated CDF,
 * `table` = n (gamma, cumulative fraction) pairs, sampled by the inverse of its
 * monotone cubic interpolant from a uniform made of the beam's correlated
 * normals (so energy_chirp still applies); 2 a bremsstrahlung source between
 * brem_min and brem_max. Replaces the spectrum branch at src/main.rs:610-693. */
int ptg_set_beam_spectrum(ptg_ctx* ctx, int kind, const double* table, size_t n, double brem_min, double brem_max);
/* Uploads host primaries to the device (replaces the Vec<Particle> handed to collide). */
int ptg_load_primaries(ptg_ctx* ctx, const ptg_particles* primaries, size_t n);

/* Runs collide() over every resident primary: replaces the chunked map at
 * src/main.rs:1118-1135 and collide() itself (src/main.rs:134-289). Primaries
 * must already be resident (ptg_load_primaries / ptg_build_beam). Results stay on
 * the device until ptg_fetch. `seed` is control:rng_seed. */
int ptg_run(ptg_ctx* ctx, const ptg_collide_options* opt, uint64_t seed, ptg_counters* counters);
/* ptg_load_primaries + ptg_run. */
int ptg_collide(ptg_ctx* ctx, const ptg_particles* primaries, size_t n, const ptg_collide_options* opt,
                uint64_t seed, ptg_counters* counters);
/* Copies results to caller buffers sized from counters (n_secondaries / n_intermediates).
 * Either output may be NULL. Replaces the Shower/ParticleBunch fields (particle/shower.rs:9-46). */
int ptg_fetch(ptg_ctx* ctx, ptg_particles* secondaries, ptg_particles* intermediates);
/* The index of the primary whose shower holds each final-state particle, in
 * ptg_fetch's order (n_secondaries values); in a batch (ptg_set_batch) it
 * names the collision. No counterpart in the reference (one shower per
 * collide() call there). */
int ptg_fetch_showers(ptg_ctx* ctx, uint32_t* shower);

/* Field-only push of n particles for up to n_steps steps each (stops when a particle
 * leaves the field): repeated Field::push (src/field/mod.rs:147-152 and the LMA
 * override plane_wave.rs:196-249). r, u are n x 4, updated in place (host);
 * work (may be NULL) receives the summed work returned by each push. */
int ptg_push(ptg_ctx* ctx, const uint8_t* species, double* r, double* u, size_t n, uint32_t n_steps,
             const ptg_collide_options* opt, uint32_t* steps_taken, double* work);

/* Weighted histogram of the resident final-state particles (hgram.rs:209-323).
 * bins_out has nbins entries, nbins <= 2^20 (min/max written back when computed). */
int ptg_histogram(ptg_ctx* ctx, ptg_hist_spec* spec, double* bins_out, double* total_weight);
/* ptg_histogram without allocating anything: no counterpart in the reference.
 * For callers that cannot allocate device or host memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). The bins
 * accumulator and range keys live in a fixed device array of the library,
 * (PTG_NOALLOC_MAX_BINS + 3) x 8 bytes per GPU, held until the process ends;
 * results go straight into the caller's bins_out / total_weight. nbins above
 * PTG_NOALLOC_MAX_BINS is refused with PTG_E_ARG. Every call on one GPU shares
 * that array and clears it first: calls that overlap on one GPU (two host
 * threads, two contexts on different streams) race, so the caller serialises them. */
#define PTG_NOALLOC_MAX_BINS 8192
int ptg_histogram_noalloc(ptg_ctx* ctx, ptg_hist_spec* spec, double* bins_out, double* total_weight);
/* Weighted 2-D histogram (Histogram::generate_2d, hgram.rs:326-446): species and
 * weight from spec_x; bins_out has nbins_x * nbins_y entries, flat index
 * by * nbins_y + bx as the reference (rows overlap when the counts differ). */
int ptg_histogram_2d(ptg_ctx* ctx, ptg_hist_spec* spec_x, ptg_hist_spec* spec_y, double* bins_out, double* total_weight);
/* The driver's post-processing of the resident final state (src/main.rs:1150-1196):
 * stop_mode 0 leaves positions, 1 drifts every particle along its momentum to
 * the largest ct among them (stop_at_time: auto), 2 to c t_stop; then, with
 * beam_basis = 1, rotates positions and momenta (current and at creation) into
 * the frame where the beam centroid, at collision angle theta in the plane at
 * phi, moves along +z (Particle::to_beam_coordinate_basis, particle/mod.rs:256-300).
 * t_common (may be NULL) receives the common time in s (+inf for mode 0). */
int ptg_finish(ptg_ctx* ctx, int stop_mode, double t_stop, int beam_basis, double theta, double phi, double* t_common);
/* The reductions of one summary statistic (stats.rs:265-410): replaces the
 * per-particle loops of total / variance / min / max / directmean. */
int ptg_stats(ptg_ctx* ctx, const ptg_stats_spec* spec, double* out);

/* Batched evaluation of one device function for parity testing. fn ids are
 * listed in ptg_eval_fn; in is n x in_w, out is n x out_w (host buffers). */
typedef enum {
    PTG_FN_LCFA_RATE = 0,        /* (chi, gamma) -> rate      lcfa/photon_emission/mod.rs:13 */
    PTG_FN_LCFA_SAMPLE = 1,      /* (chi, gamma, r1, r2, r3) -> (omega, theta_ok, theta, phi)  :80 */
    PTG_FN_AUX_T = 2,            /* chi -> (T_par, T_perp)   lcfa/pair_creation/mod.rs:13 */
    PTG_FN_AIRY = 3,             /* x -> (Ai, Ai', K1/3, K2/3) special_functions/airy.rs:41-136 */
    PTG_FN_FIELDS = 4,           /* (ct, x, y, z) -> (E[3], B[3], a) current laser */
    PTG_FN_GAUNT = 5,            /* chi -> g(chi)  lcfa/photon_emission/mod.rs:330 */
    PTG_FN_PAIR_SPECTRUM = 6,    /* (s, chi, sv1) -> spectrum  lcfa/pair_creation/mod.rs:80 */
    PTG_FN_LCFA_STOKES = 7,      /* (k[4], chi, gamma, v[3], w[3]) -> Stokes[4]  mod.rs:239 */
    PTG_FN_LP_RATE = 8,          /* (a, eta) -> (f, in_table)  nonlinear_compton/lp/mod.rs:430 */
    PTG_FN_LP_HARMONIC = 9,      /* (a, eta, frac) -> n  nonlinear_compton/lp/mod.rs:478 */
    PTG_FN_DOUBLE_BESSEL = 10,   /* (n, x, y, mode) -> J(n-2..n+2, x, y); mode 0: as the kernels evaluate it (saddle point or quadrature), 1: the reference's recurrence, 2: quadrature, 3: the reference's evaluation (saddle point or recurrence), 4: quadrature shared by a wave  special_functions/double_bessel.rs */
    PTG_FN_LP_SAMPLE = 11,       /* (a, eta, stream, fixed_n) -> (n, s, theta, S1, S2, trials)  lp/mod.rs:586 */
    PTG_FN_BW_RATE = 12,         /* (a, eta) -> (par, perp, table: 0 negligible, 1 mid, 2 high, -1 outside)
                                    pair_creation/lp/mod.rs:385-401 */
    PTG_FN_BW_HARMONIC = 13,     /* (a, eta, S1, frac) -> n  pair_creation/lp/mod.rs:424-436 */
    PTG_FN_BW_SAMPLE = 14,       /* (a, eta, S1, S2, stream, fixed_n) -> (n, s, theta, trials)  lp/mod.rs:261-301 */
    PTG_FN_BESSEL_J = 15,        /* (n, x) -> (J_{n-1}, J_n, J_{n+1})(x)  special_functions/bessel.rs:26-33 */
    PTG_FN_CP_RATE = 16,         /* (a, eta, classical) -> (f, in_table)  nonlinear_compton/cp/mod.rs:162-176, cp/classical.rs:22-44 */
    PTG_FN_CP_HARMONIC = 17,     /* (a, eta, frac, classical) -> n  cp/tables/mod.rs:46-121, cp/classical.rs:83-123 */
    PTG_FN_CP_PARTIAL_RATE = 18, /* (n, a, eta, classical) -> (partial rate, rejection ceiling)  cp/mod.rs:42-155,
                                    cp/classical.rs:46-69 (classical: rate per unit eta) */
    PTG_FN_CP_SAMPLE = 19,       /* (a, eta, stream, fixed_n, classical) -> (n, s, theta, S1, S2, S3, trials)  cp/mod.rs:181-234 */
    PTG_FN_LP_CL_RATE = 20,      /* (a, eta) -> (f, in_table)  nonlinear_compton/lp/classical.rs:189-213 */
    PTG_FN_LP_CL_HARMONIC = 21,  /* (a, frac) -> n  lp/classical.rs:215-252 */
    PTG_FN_LP_CL_CEILING = 22,   /* (n, a) -> rejection ceiling  lp/classical.rs:43-110 */
    PTG_FN_LP_CL_SAMPLE = 23,    /* (a, eta, stream, fixed_n) -> (n, s, theta, S1, S2, trials)  lp/classical.rs:254-333 */
    PTG_FN_CPBW_RATE = 24,       /* (a, eta) -> (f(S3 = +1), f(S3 = -1), method: 0 zero, 1 summation, 2 table, -1 outside)
                                    pair_creation/cp/mod.rs:271-281 */
    PTG_FN_CPBW_PARTIAL_RATE = 25, /* (n, a, eta) -> (S3 = +1, S3 = -1)  pair_creation/cp/mod.rs:98-184 */
    PTG_FN_CPBW_HARMONIC = 26,   /* (a, eta, S3, frac) -> (n, fell_back)  pair_creation/cp/mod.rs:327-365 */
    PTG_FN_CPBW_SAMPLE = 27,     /* (a, eta, S1, S2, S3, stream, fixed_n) -> (n, s, theta, trials)  pair_creation/cp/mod.rs:188-245 */
    PTG_FN_LP_CEILING = 28,      /* (n, a, eta) -> (largest sample of the reference's ceiling points, its tabulated value or -1)
                                    lp/mod.rs:53-140 */
    PTG_FN_COUNT = 29
} ptg_eval_fn;
int ptg_eval(ptg_ctx* ctx, int fn, const double* in, size_t n, size_t in_w, double* out, size_t out_w);

/* ------------------------------------------------------------------ batches and launch tuning
 * A batch of collisions in one particle population (no counterpart in the
 * reference, which runs one collision per input deck): n_coll lasers that may
 * differ only in a0 (an a0 scan; LMA fields) and n_coll beams; the primaries of
 * collision k are built as ptg_build_beam would (initial_z from the laser and
 * the beam length, src/main.rs:1062-1066) at consecutive indices, and each
 * particle's shower index (its primary) names its collision. ptg_run then
 * transports every collision with one sequence of launches. A later
 * ptg_set_laser, ptg_build_beam or ptg_load_primaries ends the batch. */
int ptg_set_batch(ptg_ctx* ctx, const ptg_laser_desc* lasers, const ptg_beam_desc* beams, int n_coll, uint64_t seed);

/* Launch-shape autotuner for the LMA lepton step kernel: runs the loaded
 * primaries (at most max_primaries of them; 0: all) at each block size (64,
 * 128, 256 threads) and persistent grid (2, 4, 8 blocks per CU), keeps the
 * fastest for later runs of this context. Results do not depend on the launch
 * shape; `consistent` reports that every candidate produced the same number
 * of emissions. No counterpart in the reference. */
typedef struct {
    int32_t block, grid_per_cu;      /* chosen */
    int32_t n_candidates, consistent;
    int32_t cand_block[16], cand_grid_per_cu[16];
    double cand_ms[16];              /* lepton step kernel time of each candidate */
} ptg_tune_result;
int ptg_autotune(ptg_ctx* ctx, const ptg_collide_options* opt, uint64_t seed, uint64_t max_primaries, ptg_tune_result* out);

/* How ptg_run drives the generations of the LMA loop, which replaces the
 * per-primary stack loop of collide() (src/main.rs:161-289) and has no
 * counterpart there. PTG_LAUNCH_GRAPH (the default): after the first
 * generation of each lepton phase, and from the first generation of each
 * photon phase, generations run in sequences of 16 captured in a HIP graph and
 * launched at once; each launch reads its counts from device memory and the
 * host reads the counters once per sequence (ptg_counters.graph_launches).
 * PTG_LAUNCH_EAGER: the host reads the counts after every launch, and
 * lepton_ms / photon_ms time every kernel. The particles transported do not
 * depend on the mode (the order of the final-state records and the ids given
 * to secondaries follow the device's scheduling in both). */
typedef enum { PTG_LAUNCH_EAGER = 0, PTG_LAUNCH_GRAPH = 1 } ptg_launch_mode;
int ptg_set_launch_mode(ptg_ctx* ctx, int mode);

/* ------------------------------------------------------------------ learned sampler
 * A surrogate for the sampling of LMA nonlinear Compton emission. No counterpart
 * in the reference, which always samples from its tables; it replaces the
 * harmonic CDF inversion (nonlinear_compton/lp/mod.rs:478-553,
 * cp/tables/mod.rs:46-121) and the rejection sampling of the lightfront
 * fraction s and, for LP, of the azimuth theta (lp/mod.rs:586-676,
 * cp/mod.rs:181-234). The photon's Stokes parameters are still computed from
 * the exact rate at the sampled point. A 4-layer perceptron 5 -> 128 -> 128 ->
 * 128 -> 3, SiLU, BF16 weights with FP32 accumulation (layer 1 in FP32):
 * inputs (ln a, ln eta, u1, u2, u3) scaled to the domain box, outputs
 * (ln nu / 5, v, t): n = ceil(nu), s = v s_max(n), theta = pi t / 2 in the
 * first quadrant. a is the sampler's amplitude: sqrt(2) a_rms (LP), a_rms (CP).
 * Events outside the box are sampled exactly: those below the larger of a_min and
 * a_use_min, or below the larger of eta_min and eta_use_min, in the same
 * generation (where the exact sampler is the cheaper, or the network's spectra
 * fall short of the exact sampler's), the others (above a_max or eta_max) in one
 * deferred exact pass. */
#define PTG_MLP_NPARAM 38304
typedef struct {
    int32_t polarization;    /* ptg_polarization sampled */
    int32_t valid;           /* passed the validation of ptg_surrogate_train */
    double a_min, a_max;     /* domain box: the network's inputs are scaled to it */
    double eta_min, eta_max;
    double a_use_min;        /* events below this amplitude are sampled exactly (0: a_min) */
    double eta_use_min;      /* events below this eta are sampled exactly (0: eta_min) */
    float params[PTG_MLP_NPARAM]; /* FP32 weights: W1[128][8] b1[128] W2[128][128] b2 W3[128][128] b3 W4[32][128] b4[32] */
} ptg_surrogate_model;

typedef struct {
    int32_t polarization;
    int32_t groups;          /* exact-sampler groups of 64 (s, theta) samples at one (a, eta, n); 1 in 8 held out */
    int32_t steps;           /* optimiser (Adam) steps */
    int32_t batch;           /* rows per step, a multiple of 64: half drawn fresh from the harmonic tables */
    double lr;               /* peak learning rate (cosine decay) */
    double a_min, a_max, eta_min, eta_max;
    uint64_t seed;
} ptg_surrogate_train_desc;

typedef struct {
    double loss[3];          /* mean losses over the last 10% of steps: harmonic, fraction (mean squares), angle (pinball) */
    double val_harmonic_rel; /* held-out groups: mean |n_pred - n| / n */
    double val_fraction_abs; /* held-out groups: mean |v_pred(u2) - v| at the rank levels (k + 1) / 65 */
    double ks_fraction;      /* held-out samples: Kolmogorov-Smirnov distance of the network's CDF of v at the
                                samples (probability integral transform, by bisection in u2) from uniform */
    double ks_angle;         /* the same for theta at the samples' v (LP; 0 for CP) */
    int32_t valid;           /* val_harmonic_rel < 0.05, ks_fraction < 0.02 and ks_angle < 0.02 */
    int32_t groups_used;     /* groups that produced 64 samples */
    uint64_t table_rows;     /* training rows drawn from the harmonic tables */
    uint64_t sampler_rows;   /* training rows from the exact sampler */
    double gen_ms, train_ms;
} ptg_surrogate_report;

/* Generates the training data on the device (the tables and the exact sampler),
 * trains the network with the library's own forward/backward MFMA kernels, loads
 * it into the context and returns it in model_out (may be NULL). */
int ptg_surrogate_train(ptg_ctx* ctx, const ptg_surrogate_train_desc* desc, ptg_surrogate_model* model_out,
                        ptg_surrogate_report* report);
/* Loads a model for its polarization (one per polarization). */
int ptg_surrogate_load(ptg_ctx* ctx, const ptg_surrogate_model* model);
/* Test probe: forward and backward of the network with params (PTG_MLP_NPARAM FP32)
 * on n rows (a multiple of 32): x n x 8 inputs, t n x 8 targets (y_nu, v, t,
 * mask_v, mask_t, tau, 0, 0), w_t the angle-loss weight; returns outputs n x 32,
 * the gradient of the mean loss (PTG_MLP_NPARAM) and the loss sums [3]. */
int ptg_mlp_probe(ptg_ctx* ctx, const float* params, const float* x, const float* t, size_t n, float w_t, float* out,
                  float* grads, float* loss);

#ifdef __cplusplus
}
#endif
#endif /* PTG_H */
