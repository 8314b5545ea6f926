// Arguments and launchers of the beam, histogram, selection and probe kernels.
#pragma once
#include "kern_common.h"

namespace ptg {

struct BeamDev {
    uint32_t n;
    int species, radial;
    double gamma, sigma, rho, radius, max_radius, length, rms_div, angle, plane_angle;
    double off_x, off_y, off_z;
    double pol[4];
    double weight;
    double initial_z;
    uint64_t key;
    uint32_t off, n_total;   // this beam's primaries at [off, off + n) of n_total (a batch of collisions)
    int spectrum;            // 0 normal, 1 tabulated CDF, 2 bremsstrahlung source (particle/dstr.rs:161-214)
    int n_cdf;
    const double* cdf;       // [n_cdf][2] (gamma, cumulative fraction)
    double brem_min, brem_max;
};

struct HistArgs {
    PBuf p;
    uint64_t n;
    int species, q, wfn, log_scale;
    uint32_t nbins;
    double gmin, bin_vol;
    int cut, cq;             // range cut on quantity cq: cmin <= q < cmax
    double cmin, cmax, scale;
};

struct EvalArgs {
    int fn;
    const double* in;
    double* out;
    uint32_t n, in_w, out_w;
    const double* lcfa;
    const double* lma;
    LcfaScalars sc;
    LaserDev laser;
    uint64_t key;
};

void launch_build_beam(const BeamDev& b, const PBuf& p, double* n0, hipStream_t s);
void launch_primary_dirs(const PBuf& p, uint32_t n, double* n0, hipStream_t s);
void launch_hist_minmax(const HistArgs& h, unsigned long long* mm, int grid, hipStream_t s);
// bins held in LDS per workgroup up to this many (64 KB), global atomics above
constexpr uint32_t HIST_LDS_BINS = 8192;
void launch_hist_bin(const HistArgs& h, double* bins, double* total, int grid, hipStream_t s);
void launch_fill_u64(unsigned long long* p, uint64_t n, unsigned long long v, hipStream_t s);
// 2-D weighted binning (hgram.rs:326-446): hx gives species, weight and the x axis,
// hy the y axis; flat bin by * ny + bx as the reference, dropped past nx * ny.
// drift to c t (when drift) then rotate by R_x(-vangle) R_y(hangle) (when rotate)
void launch_finish(const PBuf& p, uint64_t n, int drift, double ct, int rotate, double hangle, double vangle,
                   hipStream_t s);
void launch_max_ct(const PBuf& p, uint64_t n, unsigned long long* key, hipStream_t s);
struct StatArgs {
    PBuf p;
    uint64_t n;
    int species, qv, qw, cut, qf;
    double fmin, fmax, shift;
};
// acc[7]: sums (all vw, in vw, in w, in w cos v, in w sin v, in w (v-s)^2, in w (v-s)); mm[2]: order keys of min/max vw
void launch_stats(const StatArgs& a, double* acc, unsigned long long* mm, int grid, hipStream_t s);
void launch_hist_bin2(const HistArgs& hx, const HistArgs& hy, double* bins, double* total, int grid, hipStream_t s);
double host_order_key_inv(unsigned long long k);
void launch_count_showers(const PBuf& p, uint64_t n, uint32_t* count, hipStream_t s);
void launch_select_multiplicity(const PBuf& src, uint64_t n, const PBuf& dst, unsigned long long* ctr, const uint32_t* count,
                                uint32_t want, hipStream_t s);
void launch_eval(const EvalArgs& e, hipStream_t s);

}  // namespace ptg
