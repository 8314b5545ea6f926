// Beam generation, weighted histograms, multiplicity compaction and the
// batched function probes used by the parity tests.
#include <cstring>

#include "kern_misc.h"

namespace ptg {

// ---------------------------------------------------------------- beam
__device__ __forceinline__ void normal2(Rng& g, double& a, double& b) {
    double u1, u2;
    rng_uniform2(g, u1, u2);
    double r = sqrt(-2.0 * log(1.0 - u1));
    double s, c;
    sincos(2.0 * PI_D * u2, &s, &c);
    a = r * c;
    b = r * s;
}

// BeamBuilder::build (src/particle/builder.rs:173-219) with GammaDistribution::Normal
// (particle/dstr.rs:161-176) and RadialDistribution (dstr.rs:26-51). Each primary
// draws from its own counter-based stream, so the beam is independent of the
// launch shape.
__global__ void k_build_beam(BeamDev b, PBuf p, double* n0) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= b.n) return;
    Rng g = rng_make(b.key, k);
    const uint32_t i = b.off + k;  // index in the primary buffer
    const uint32_t nt = b.n_total;
    double gamma, dz;
    if (b.spectrum == 1) {
        // correlated normals mapped to a uniform by the reference's erf
        // approximation (NORTA), then the inverse of the tabulated CDF
        double n0_, n1_;
        normal2(g, n0_, n1_);
        const double n2 = b.rho * n0_ + sqrt(1.0 - b.rho * b.rho) * n1_;
        const double x = n2 * 0.70710678118654752440;
        const double u1 = 0.5 * (1.0 + tanh(167.0 * x / 148.0 + 11.0 * x * x * x / 109.0));
        dz = b.length * n0_;
        if (!pwmci_invert(b.cdf, b.n_cdf, u1, true, gamma)) gamma = b.cdf[0];
    } else if (b.spectrum == 2) {
        // x = gamma / gamma_max from 4/(3x) - 4/3 + x by rejection
        const double x_min = b.brem_min / b.brem_max;
        const double y_max = 4.0 / (3.0 * x_min) - 4.0 / 3.0 + x_min;
        double x = 1.0;
        for (int t = 0; t < 100000; ++t) {
            double r1, u;
            rng_uniform2(g, r1, u);
            x = x_min + (1.0 - x_min) * r1;
            if (u <= (4.0 / (3.0 * x) - 4.0 / 3.0 + x) / y_max) break;
        }
        double na_, nb_;
        normal2(g, na_, nb_);
        dz = b.length * na_;
        gamma = x * b.brem_max;
    } else {
        for (int t = 0;; ++t) {
            double n0_, n1_;
            normal2(g, n0_, n1_);
            double n2 = b.rho * n0_ + sqrt(1.0 - b.rho * b.rho) * n1_;
            dz = b.length * n0_;
            gamma = b.gamma + b.sigma * n2;
            if (gamma > 1.0 || t > 1000) break;
        }
    }
    double uu = b.species == 2 ? -gamma : -sqrt(gamma * gamma - 1.0);
    double na, nb;
    normal2(g, na, nb);
    double theta_x = b.angle + b.rms_div * na;
    double theta_y = b.rms_div * nb;
    double sx, cx, sy, cy;
    sincos(theta_x, &sx, &cx);
    sincos(theta_y, &sy, &cy);
    double sp_, cp_;
    sincos(b.plane_angle, &sp_, &cp_);
    d3 u3 = rot_z(d3{uu * sx * cy, uu * sy, uu * cx * cy}, sp_, cp_);
    double ut = b.species == 2 ? sqrt(nsq(u3)) : sqrt(1.0 + nsq(u3));
    double t, z;
    if (b.off_z >= 0.0) { t = -b.initial_z; z = b.initial_z + b.off_z + dz; }
    else { t = -b.initial_z - fabs(b.off_z); z = b.initial_z + dz; }
    double x, y;
    if (b.radial == 0) {
        normal2(g, x, y);
        x *= b.radius;
        y *= b.radius;
    } else if (b.radial == 1) {
        for (int k = 0;; ++k) {
            normal2(g, x, y);
            x *= b.radius;
            y *= b.radius;
            if (x * x / (b.max_radius * b.max_radius) + y * y / (b.max_radius * b.max_radius) <= 1.0 || k > 1000) break;
        }
    } else {
        double a1, a2;
        rng_uniform2(g, a1, a2);
        double rr = b.radius * sqrt(a1), s, c;
        sincos(2.0 * PI_D * a2, &s, &c);
        x = rr * c;
        y = rr * s;
    }
    x += b.off_x;
    y += b.off_y;
    double sa, ca;
    sincos(b.angle, &sa, &ca);
    d3 r3{ca * x + sa * z, y, -sa * x + ca * z};  // rotate about y by the collision angle
    r3 = rot_z(r3, sp_, cp_);
    p.d[P_RT][i] = t; p.d[P_RX][i] = r3.x; p.d[P_RY][i] = r3.y; p.d[P_RZ][i] = r3.z;
    p.d[P_R0T][i] = t; p.d[P_R0X][i] = r3.x; p.d[P_R0Y][i] = r3.y; p.d[P_R0Z][i] = r3.z;
    p.d[P_UT][i] = ut; p.d[P_UX][i] = u3.x; p.d[P_UY][i] = u3.y; p.d[P_UZ][i] = u3.z;
    p.d[P_U0T][i] = b.species == 2 ? 0.0 : 1.0; p.d[P_U0X][i] = 0.0; p.d[P_U0Y][i] = 0.0; p.d[P_U0Z][i] = 0.0;
    p.d[P_SI][i] = b.pol[0]; p.d[P_SQ][i] = b.pol[1]; p.d[P_SU][i] = b.pol[2]; p.d[P_SV][i] = b.pol[3];
    p.d[P_PAYLOAD][i] = 0.0; p.d[P_PCHI][i] = 0.0; p.d[P_ICOUNT][i] = 0.0; p.d[P_WORK][i] = 0.0;
    p.d[P_WEIGHT][i] = b.weight;
    p.id[i] = i;
    p.parent_id[i] = i;
    p.stream[i] = i;
    p.shower[i] = i;
    p.meta[i] = uint32_t(b.species) | META_PRIMARY;
    p.nev[i] = 0;
    double inv = 1.0 / sqrt(nsq(u3));
    n0[i] = u3.x * inv; n0[nt + i] = u3.y * inv; n0[2 * nt + i] = u3.z * inv;
}

void launch_build_beam(const BeamDev& b, const PBuf& p, double* n0, hipStream_t s) {
    k_build_beam<<<(b.n + 255) / 256, 256, 0, s>>>(b, p, n0);
}

// Unit momentum direction of uploaded primaries (for the final angle cut).
__global__ void k_primary_dirs(PBuf p, uint32_t n, double* n0) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    d3 u{p.d[P_UX][i], p.d[P_UY][i], p.d[P_UZ][i]};
    double m = sqrt(nsq(u));
    double inv = m > 0.0 ? 1.0 / m : 0.0;
    n0[i] = u.x * inv; n0[n + i] = u.y * inv; n0[2 * n + i] = u.z * inv;
}
void launch_primary_dirs(const PBuf& p, uint32_t n, double* n0, hipStream_t s) {
    k_primary_dirs<<<(n + 255) / 256, 256, 0, s>>>(p, n, n0);
}

// ---------------------------------------------------------------- histograms
// Accessors of src/output/functions.rs:10-263.
// Fraction of the intensity polarized along a unit axis (ax = 0: x, 1: y): the
// Stokes vector as a Jones vector (e_x real, e_y complex) in the basis x = (n_z,
// 0, -n_x)/|..|, y = x cross n about the direction n, projected on the axis and
// mixed with the unpolarized part (geometry/polarization.rs:159-211).
__device__ double pol_along(const PBuf& p, uint64_t i, int ax) {
    double ux = p.d[P_UX][i], uy = p.d[P_UY][i], uz = p.d[P_UZ][i];
    double un = sqrt(ux * ux + uy * uy + uz * uz);
    double nx = ux / un, ny = uy / un, nz = uz / un;
    double mag = hypot(nx, nz);
    double x0 = 1.0, x2 = 0.0;
    if (mag != 0.0) { x0 = nz / mag; x2 = -nx / mag; }
    // y = x cross n with x = (x0, 0, x2)
    double y0 = -x2 * ny, y1 = x2 * nx - x0 * nz, y2 = x0 * ny;
    double xa = ax == 0 ? x0 : 0.0, ya = ax == 0 ? y0 : y1;
    double unpol = 0.5 * (xa * xa + ya * ya);
    double si = p.d[P_SI][i], sq = p.d[P_SQ][i], su = p.d[P_SU][i], sv = p.d[P_SV][i];
    double frac = hypot(hypot(sq, su), sv) / si;
    if (!(frac > 0.0)) return unpol;
    double q = sq / (frac * si), u = su / (frac * si), v = sv / (frac * si);
    double ex;
    if (1.0 + q < 1.0e-12) {
        double s2 = u * u + v * v;
        double delta = 0.5 * s2 + 0.125 * s2 * s2;
        ex = sqrt(0.5 * delta);
    } else {
        ex = sqrt(0.5 * (1.0 + q));
    }
    double eyr = 1.0, eyi = 0.0;
    if (ex != 0.0) { eyr = 0.5 * u / ex; eyi = -0.5 * v / ex; }
    double re = ex * xa + eyr * ya, im = eyi * ya;
    return frac * (re * re + im * im) + (1.0 - frac) * unpol;
}

__device__ __forceinline__ double quantity(const PBuf& p, uint64_t i, int q) {
    double ut = p.d[P_UT][i], ux = p.d[P_UX][i], uy = p.d[P_UY][i], uz = p.d[P_UZ][i];
    double p0 = M_E_MEV * ut, p1 = M_E_MEV * ux, p2 = M_E_MEV * uy, p3 = M_E_MEV * uz;
    switch (q) {
    case PTG_Q_ANGLE_X: return atan2(p1, p3);
    case PTG_Q_ANGLE_Y: return atan2(p2, p3);
    case PTG_Q_THETA: return atan2(hypot(p1, p2), -p3);
    case PTG_Q_POLAR_ANGLE: return atan2(hypot(p1, p2), p3);
    case PTG_Q_PX: return p1;
    case PTG_Q_PY: return p2;
    case PTG_Q_PZ: return p3;
    case PTG_Q_P_PERP: return hypot(p1, p2);
    case PTG_Q_R_PERP: return hypot(p1, p2) / (p0 - p3);
    case PTG_Q_R_X: return p1 / (p0 - p3);
    case PTG_Q_R_Y: return p2 / (p0 - p3);
    case PTG_Q_P_MINUS: return p0 - p3;
    case PTG_Q_P_PLUS: return p0 + p3;
    case PTG_Q_GAMMA: return ut;
    case PTG_Q_ENERGY: return ut * M_E_MEV;
    case PTG_Q_UNIT: return 1.0;
    case PTG_Q_X: return p.d[P_RX][i];
    case PTG_Q_Y: return p.d[P_RY][i];
    case PTG_Q_Z: return p.d[P_RZ][i];
    case PTG_Q_RADIUS: return hypot(p.d[P_RX][i], p.d[P_RY][i]);
    case PTG_Q_BIRTH_A: return p.d[P_PAYLOAD][i];
    case PTG_Q_PARENT_CHI: return p.d[P_PCHI][i];
    case PTG_Q_N_INTER: return p.d[P_ICOUNT][i];
    case PTG_Q_ABSORPTION: return p.d[P_WORK][i] * M_E_MEV;
    case PTG_Q_S1: return p.d[P_SQ][i];
    case PTG_Q_S2: return p.d[P_SU][i];
    case PTG_Q_S3: return p.d[P_SV][i];
    case PTG_Q_POL_X: return pol_along(p, i, 0);
    case PTG_Q_POL_Y: return pol_along(p, i, 1);
    case PTG_Q_WEIGHT: return p.d[P_WEIGHT][i];
    case PTG_Q_HELICITY: return p.d[P_WEIGHT][i] * p.d[P_SV][i];
    default: return 0.0;
    }
}
__device__ __forceinline__ double hweight(const PBuf& p, uint64_t i, int w) {
    double wt = p.d[P_WEIGHT][i];
    if (w == PTG_W_ENERGY) return wt * p.d[P_UT][i] * M_E_MEV;
    if (w == PTG_W_HELICITY) return wt * p.d[P_SV][i];
    if (w == PTG_W_POL_X) return wt * pol_along(p, i, 0);
    if (w == PTG_W_POL_Y) return wt * pol_along(p, i, 1);
    return wt;
}
// the binned value: converted to the output unit, then log-scaled if asked
__device__ __forceinline__ double hvalue(const HistArgs& h, uint64_t i) {
    double v = quantity(h.p, i, h.q) * h.scale;
    return h.log_scale ? log(v) : v;
}
__device__ __forceinline__ bool hcut(const HistArgs& h, uint64_t i) {
    if (!h.cut) return true;
    double c = quantity(h.p, i, h.cq);
    return c >= h.cmin && c < h.cmax;
}

__device__ __forceinline__ double order_key_inv(unsigned long long k) {
    return __longlong_as_double((long long)(k & 0x8000000000000000ull ? k & 0x7fffffffffffffffull : ~k));
}
__device__ __forceinline__ unsigned long long order_key(double v) {
    unsigned long long b = (unsigned long long)__double_as_longlong(v);
    return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}

// global min/max of the finite, selected values (hgram.rs:150-167)
__global__ void k_hist_minmax(HistArgs h, unsigned long long* mm) {
    unsigned long long lo = ~0ull, hi = 0ull;
    for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < h.n; i += uint64_t(gridDim.x) * blockDim.x) {
        if ((h.p.meta[i] & 3u) != uint32_t(h.species)) continue;
        double v = hvalue(h, i);
        if (!isfinite(v) || !hcut(h, i)) continue;
        unsigned long long k = order_key(v);
        lo = k < lo ? k : lo;
        hi = k > hi ? k : hi;
    }
    // wave reduction, then one atomic per wave
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long olo = __shfl_xor(lo, off), ohi = __shfl_xor(hi, off);
        lo = olo < lo ? olo : lo;
        hi = ohi > hi ? ohi : hi;
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(mm, lo);
        atomicMax(mm + 1, hi);
    }
}

// Weighted binning with per-workgroup bins in LDS, merged with one global atomic
// per bin per workgroup (hgram.rs:268-298). The total counts all weights of the
// species, binned or not.
// Above HIST_LDS_BINS bins the kernel bins straight into global memory.
template <bool LDS>
__global__ void k_hist_bin(HistArgs h, double* bins, double* total) {
    extern __shared__ double lbins[];
    if (LDS) {
        for (uint32_t b = threadIdx.x; b < h.nbins; b += blockDim.x) lbins[b] = 0.0;
        __syncthreads();
    }
    double tot = 0.0;
    for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < h.n; i += uint64_t(gridDim.x) * blockDim.x) {
        if ((h.p.meta[i] & 3u) != uint32_t(h.species)) continue;
        double v = hvalue(h, i);
        double w = hweight(h.p, i, h.wfn);
        tot += w;
        if (!isfinite(v) || !hcut(h, i)) continue;
        double fb = floor((v - h.gmin) / h.bin_vol);
        if (!(fb >= 0.0)) fb = 0.0;  // Rust `as usize` saturates at 0
        if (fb < double(h.nbins)) {
            if (LDS) atomicAdd(&lbins[uint32_t(fb)], w);
            else atomicAdd(&bins[uint32_t(fb)], w);
        }
    }
    for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off);
    if ((threadIdx.x & 63) == 0 && tot != 0.0) atomicAdd(total, tot);
    if (LDS) {
        __syncthreads();
        for (uint32_t b = threadIdx.x; b < h.nbins; b += blockDim.x)
            if (lbins[b] != 0.0) atomicAdd(&bins[b], lbins[b]);
    }
}

// 2-D: a value pair is binned only if both are finite; the flat index is
// by * ny + bx (hgram.rs:418, as the reference: row index times the number of
// y bins), so with nx != ny rows overlap and indices past nx * ny are dropped.
// Bins in LDS when they fit (<= 6144 doubles), else straight to global atomics.
template <bool LDS>
__global__ void k_hist_bin2(HistArgs hx, HistArgs hy, double* bins, double* total) {
    extern __shared__ double lbins[];
    const uint32_t nb = hx.nbins * hy.nbins;
    if (LDS) {
        for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) lbins[b] = 0.0;
        __syncthreads();
    }
    double tot = 0.0;
    for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < hx.n; i += uint64_t(gridDim.x) * blockDim.x) {
        if ((hx.p.meta[i] & 3u) != uint32_t(hx.species)) continue;
        double vx = hvalue(hx, i), vy = hvalue(hy, i);
        const double w = hweight(hx.p, i, hx.wfn);
        tot += w;
        if (!isfinite(vx) || !isfinite(vy) || !hcut(hx, i)) continue;
        double bx = hx.bin_vol == 0.0 ? 0.0 : floor((vx - hx.gmin) / hx.bin_vol);
        double by = hy.bin_vol == 0.0 ? 0.0 : floor((vy - hy.gmin) / hy.bin_vol);
        if (!(bx >= 0.0)) bx = 0.0;  // Rust `as usize` saturates at 0
        if (!(by >= 0.0)) by = 0.0;
        const double fb = by * double(hy.nbins) + bx;
        if (fb < double(nb)) {
            if (LDS) atomicAdd(&lbins[uint32_t(fb)], w);
            else atomicAdd(&bins[uint32_t(fb)], w);
        }
    }
    for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off);
    if ((threadIdx.x & 63) == 0 && tot != 0.0) atomicAdd(total, tot);
    if (LDS) {
        __syncthreads();
        for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x)
            if (lbins[b] != 0.0) atomicAdd(&bins[b], lbins[b]);
    }
}
void launch_hist_bin2(const HistArgs& hx, const HistArgs& hy, double* bins, double* total, int grid, hipStream_t s) {
    const uint32_t nb = hx.nbins * hy.nbins;
    if (nb <= 6144) k_hist_bin2<true><<<grid, 256, nb * sizeof(double), s>>>(hx, hy, bins, total);
    else k_hist_bin2<false><<<grid, 256, 0, s>>>(hx, hy, bins, total);
}

// ---------------------------------------------------------------- post-processing
__global__ void k_max_ct(PBuf p, uint64_t n, unsigned long long* key) {
    unsigned long long hi = 0ull;
    for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) {
        const unsigned long long k = order_key(p.d[P_RT][i]);
        hi = k > hi ? k : hi;
    }
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o = __shfl_xor(hi, off);
        hi = o > hi ? o : hi;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(key, hi);
}
void launch_max_ct(const PBuf& p, uint64_t n, unsigned long long* key, hipStream_t s) {
    int grid = int(std::min<uint64_t>(1024, (n + 255) / 256));
    if (n) k_max_ct<<<grid, 256, 0, s>>>(p, n, key);
}

// r -> r + u (ct - r_t) / u_t (src/main.rs:1165-1169), then the rotation of
// Particle::to_beam_coordinate_basis: about y by hangle, then about x by -vangle,
// for the position and momentum now and at creation (times unchanged)
__global__ void k_finish(PBuf p, uint64_t n, int drift, double ct, int rotate, double ch, double sh, double cv,
                         double sv) {
    const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    if (drift) {
        const double s = (ct - p.d[P_RT][i]) / p.d[P_UT][i];
        p.d[P_RX][i] += p.d[P_UX][i] * s;
        p.d[P_RY][i] += p.d[P_UY][i] * s;
        p.d[P_RZ][i] += p.d[P_UZ][i] * s;
        p.d[P_RT][i] += p.d[P_UT][i] * s;
    }
    if (rotate) {
        for (int base : {P_RX, P_UX, P_R0X, P_U0X}) {
            const double x = p.d[base][i], y = p.d[base + 1][i], z = p.d[base + 2][i];
            const double x1 = ch * x + sh * z, z1 = ch * z - sh * x;  // about y by hangle
            p.d[base][i] = x1;
            p.d[base + 1][i] = cv * y + sv * z1;                      // about x by -vangle
            p.d[base + 2][i] = cv * z1 - sv * y;
        }
    }
}
void launch_finish(const PBuf& p, uint64_t n, int drift, double ct, int rotate, double hangle, double vangle,
                   hipStream_t s) {
    if (!n) return;
    k_finish<<<int((n + 255) / 256), 256, 0, s>>>(p, n, drift, ct, rotate, cos(hangle), sin(hangle), cos(vangle),
                                                   sin(vangle));
}

// One pass of every reduction a summary statistic needs (stats.rs:265-410),
// wave-reduced then one atomic per wave and sum.
__global__ void k_stats(StatArgs a, double* acc, unsigned long long* mm) {
    double s[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    unsigned long long lo = ~0ull, hi = 0ull;
    for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < a.n; i += uint64_t(gridDim.x) * blockDim.x) {
        if ((a.p.meta[i] & 3u) != uint32_t(a.species)) continue;
        const double v = quantity(a.p, i, a.qv), w = quantity(a.p, i, a.qw);
        const double vw = v * w;
        s[0] += vw;
        if (a.cut) {
            const double f = quantity(a.p, i, a.qf);
            if (!(f > a.fmin && f < a.fmax)) continue;
        }
        s[1] += vw;
        s[2] += w;
        s[3] += w * cos(v);
        s[4] += w * sin(v);
        const double d = v - a.shift;
        s[5] += w * d * d;
        s[6] += w * d;
        if (!isnan(vw)) {
            const unsigned long long k = order_key(vw);
            lo = k < lo ? k : lo;
            hi = k > hi ? k : hi;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        for (int j = 0; j < 7; ++j) s[j] += __shfl_xor(s[j], off);
        unsigned long long olo = __shfl_xor(lo, off), ohi = __shfl_xor(hi, off);
        lo = olo < lo ? olo : lo;
        hi = ohi > hi ? ohi : hi;
    }
    if ((threadIdx.x & 63) == 0) {
        for (int j = 0; j < 7; ++j)
            if (s[j] != 0.0) atomicAdd(&acc[j], s[j]);
        atomicMin(mm, lo);
        atomicMax(mm + 1, hi);
    }
}
void launch_stats(const StatArgs& a, double* acc, unsigned long long* mm, int grid, hipStream_t s) {
    k_stats<<<grid, 256, 0, s>>>(a, acc, mm);
}

void launch_hist_minmax(const HistArgs& h, unsigned long long* mm, int grid, hipStream_t s) {
    k_hist_minmax<<<grid, 256, 0, s>>>(h, mm);
}
void launch_hist_bin(const HistArgs& h, double* bins, double* total, int grid, hipStream_t s) {
    if (h.nbins <= HIST_LDS_BINS) k_hist_bin<true><<<grid, 256, h.nbins * sizeof(double), s>>>(h, bins, total);
    else k_hist_bin<false><<<grid, 256, 0, s>>>(h, bins, total);
}

// n copies of v (clears counters and accumulators with a kernel, which graph
// capture records as an ordinary kernel node)
__global__ void k_fill_u64(unsigned long long* p, uint64_t n, unsigned long long v) {
    for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) p[i] = v;
}
void launch_fill_u64(unsigned long long* p, uint64_t n, unsigned long long v, hipStream_t s) {
    if (n) k_fill_u64<<<int(std::min<uint64_t>(1024, (n + 255) / 256)), 256, 0, s>>>(p, n, v);
}
double host_order_key_inv(unsigned long long k) {
    unsigned long long b = (k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k;
    double v;
    memcpy(&v, &b, 8);
    return v;
}

// ---------------------------------------------------------------- multiplicity selection
// Final-state particles per shower, counted over the finished final state (the
// length of Shower::secondaries after the cuts, src/main.rs:268-282).
__global__ void k_count_showers(PBuf p, uint64_t n, uint32_t* count) {
    const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
    if (i < n) atomicAdd(count + p.shower[i], 1u);
}
void launch_count_showers(const PBuf& p, uint64_t n, uint32_t* count, hipStream_t s) {
    if (n) k_count_showers<<<int((n + 255) / 256), 256, 0, s>>>(p, n, count);
}

// Drops every final-state particle of showers whose kept count differs from
// multiplicity + 1 (src/main.rs:277-282), compacting into `dst`.
__global__ void k_select_multiplicity(PBuf src, uint64_t n, PBuf dst, unsigned long long* ctr, const uint32_t* count,
                                      uint32_t want) {
    uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    if (count[src.shower[i]] != want) return;
    uint64_t j = wave_reserve(ctr, 1u);
#pragma unroll
    for (int c = 0; c < P_NDOUBLE; ++c) dst.d[c][j] = src.d[c][i];
    dst.id[j] = src.id[i];
    dst.parent_id[j] = src.parent_id[i];
    dst.stream[j] = src.stream[i];
    dst.shower[j] = src.shower[i];
    dst.meta[j] = src.meta[i];
    dst.nev[j] = src.nev[i];
}
void launch_select_multiplicity(const PBuf& src, uint64_t n, const PBuf& dst, unsigned long long* ctr, const uint32_t* count,
                                uint32_t want, hipStream_t s) {
    if (n) k_select_multiplicity<<<(n + 255) / 256, 256, 0, s>>>(src, n, dst, ctr, count, want);
}

// ---------------------------------------------------------------- function probes
__global__ void k_eval(EvalArgs e) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= e.n) return;
    const double* x = e.in + uint64_t(i) * e.in_w;
    double* y = e.out + uint64_t(i) * e.out_w;
    TP tb = e.lcfa;
    switch (e.fn) {
    case PTG_FN_LCFA_RATE: y[0] = lcfa_rate(tb, e.sc, x[0], x[1]); break;
    case PTG_FN_LCFA_SAMPLE: {
        double om, th, ph;
        bool ok;
        lcfa_sample(tb, e.sc, x[0], x[1], x[2], x[3], x[4], om, ok, th, ph);
        y[0] = om; y[1] = ok ? 1.0 : 0.0; y[2] = th; y[3] = ph;
        break;
    }
    case PTG_FN_AUX_T: auxiliary_t(tb, e.sc, x[0], y[0], y[1]); break;
    case PTG_FN_AIRY:
        y[0] = x[0] >= 0.0 ? airy_ai(tb, x[0]) : 0.0;
        y[1] = x[0] >= 0.0 ? airy_ai_prime(tb, x[0]) : 0.0;
        y[2] = bessel_k13(tb, x[0]);
        y[3] = bessel_k23(tb, x[0]);
        break;
    case PTG_FN_FIELDS: {
        d3 E, B;
        d4 r{x[0], x[1], x[2], x[3]};
        if (e.laser.kind == K_FAST_FOCUSED_LASER) ffl_fields(e.laser, r, E, B);
        else if (e.laser.kind == K_FAST_PLANE_WAVE) fpw_fields(e.laser, r, E, B);
        else { E = {0, 0, 0}; B = {0, 0, 0}; }
        y[0] = E.x; y[1] = E.y; y[2] = E.z; y[3] = B.x; y[4] = B.y; y[5] = B.z;
        y[6] = e.laser.a_norm * sqrt(nsq(E));
        break;
    }
    case PTG_FN_GAUNT: y[0] = gaunt_factor(tb, x[0]); break;
    case PTG_FN_PAIR_SPECTRUM: y[0] = pair_spectrum(tb + lcfa_pack::PAIR_QN, x[0], x[1], x[2]); break;
    case PTG_FN_LCFA_STOKES: {
        sv4 s = lcfa_stokes(tb, d4{x[0], x[1], x[2], x[3]}, x[4], x[5], d3{x[6], x[7], x[8]}, d3{x[9], x[10], x[11]}, false);
        y[0] = s.i; y[1] = s.q; y[2] = s.u; y[3] = s.v;
        break;
    }
    default: break;
    }
}
void launch_eval(const EvalArgs& e, hipStream_t s) {
    if (e.n) k_eval<<<(e.n + 127) / 128, 128, 0, s>>>(e);
}

}  // namespace ptg
