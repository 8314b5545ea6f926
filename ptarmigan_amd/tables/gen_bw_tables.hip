// Host program: the cumulative harmonic tables of LP nonlinear Breit-Wheeler pair
// creation for the mid range of (a, eta) (0.04 < a < 2), which the reference
// generates with pair_creation/lp/mod.rs:1065-1163 (TotalRate::by_summation) and
// whose output is not part of its checkout.
//
// Every node sums the partial rates of all harmonics in sum_limits; each
// partial rate integrates the double-differential rate of dev_bw.h over s and
// theta, by fixed Gauss-Legendre rules (lp/mod.rs:125-185) or by the adaptive
// Clenshaw-Curtis scheme of quadrature/adaptive.rs:1-380 when the theta bound
// is narrow (lp/mod.rs:187-228). The row kept per node is 16 log-spaced
// harmonics with their cumulative (par, perp) rates (lp/mod.rs:445-512).
//
// Usage: gen_bw_tables <in.bin> <out.bin> <threads>
//   in.bin:  u64 n_nodes, f64 gl16 nodes[16], weights[16], gl32 nodes[32], weights[32],
//            then n_nodes x (a, eta)
//   out.bin: n_nodes x 16 x (n, par, perp) f64
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

#include "../csrc/dev_bw.h"

using namespace ptg;

namespace {

struct Cx { double re, im; };
Cx operator+(Cx a, Cx b) { return {a.re + b.re, a.im + b.im}; }
Cx operator-(Cx a, Cx b) { return {a.re - b.re, a.im - b.im}; }
Cx operator*(double s, Cx a) { return {s * a.re, s * a.im}; }

double G16N[16], G16W[16], G32N[32], G32W[32];

// Partial rate of harmonic n: (par, perp) integrated over the whole (s, theta) domain.
struct Partial {
    int n;
    double a, eta;
    DBesselRec D;
    BwThetaBound tb;
    std::vector<double> ws;
    Partial(int n_, double a_, double eta_) : n(n_), a(a_), eta(eta_) {
        ws.assign(size_t(dbessel_len(n)) * 5, 0.0);
        D = dbessel_rec_at(n, ws.data(), 1);
        tb = bw_theta_bound(n, a, eta);
    }
    Cx at(double s, double theta) {
        double u, d1, d2;
        bw_elements(D, a, eta, s, theta, u, d1, d2);
        return {(u - d1) / (2.0 * PI_D), (u + d1) / (2.0 * PI_D)};
    }
    Cx over_theta(double s, double max_th) {
        Cx sum{0.0, 0.0};
        for (int i = 0; i < 32; ++i) sum = sum + (0.5 * G32W[i] * max_th) * at(s, 0.5 * (G32N[i] + 1.0) * max_th);
        return sum;
    }
    Cx segment(double s0, double s1, const double* N, const double* W, int m) {
        Cx acc{0.0, 0.0};
        for (int i = 0; i < m; ++i) {
            double s = s0 + 0.5 * (N[i] + 1.0) * (s1 - s0);
            acc = acc + (0.5 * W[i] * (s1 - s0)) * over_theta(s, bw_theta_at(tb, s));
        }
        return acc;
    }
    Cx direct() {
        double s_min, s_max;
        bw_s_bounds(n, a, eta, s_min, s_max);
        double s_peak = fmin(0.5 + sqrt(0.25 - (1.0 + a * a) / (2.0 * double(n) * eta)), 0.5);
        Cx integral;
        if (s_peak < s_min + 2.0 * (s_max - s_min) / 3.0) {
            s_peak = fmin(s_peak, s_min + (s_max - s_min) / 3.0);
            integral = segment(s_min, s_peak, G32N, G32W, 32) + segment(s_peak, s_max, G32N, G32W, 32);
        } else {
            double d = s_max - s_peak;
            integral = segment(s_min, s_max - 2.0 * d, G16N, G16W, 16) + segment(s_max - 2.0 * d, s_peak, G32N, G32W, 32) +
                       segment(s_peak, s_max, G32N, G32W, 32);
        }
        return 8.0 * integral;
    }

    // ---- adaptive 2-D Clenshaw-Curtis integration
    struct Region { Cx integral, error; double x[2], y[2]; Cx cache[3][3]; };
    Cx f(double s, double th) { return th < bw_theta_at(tb, s) ? at(s, th) : Cx{0.0, 0.0}; }
    static constexpr double R5[5][3] = {{0.0, 1.0 / 30.0, 1.0 / 6.0}, {0.1464466094067262, 4.0 / 15.0, 0.0}, {0.5, 2.0 / 5.0, 2.0 / 3.0},
                                        {0.8535533905932738, 4.0 / 15.0, 0.0}, {1.0, 1.0 / 30.0, 1.0 / 6.0}};
    static constexpr double X5[9] = {0.0, 0.1464466094067262, 0.5, 0.8535533905932738, 1.0, 1.1464466094067262, 1.5,
                                     1.8535533905932738, 2.0};
    static constexpr double R9[9][3] = {{0.0000000000000000, 0.0079365079365079, 1.0 / 30.0}, {0.0380602337443566, 0.0731093246080091, 0.0},
                                        {0.1464466094067262, 0.1396825396825397, 4.0 / 15.0}, {0.3086582838174551, 0.1808589293602449, 0.0},
                                        {0.5000000000000000, 0.1968253968253968, 2.0 / 5.0}, {0.6913417161825449, 0.1808589293602449, 0.0},
                                        {0.8535533905932738, 0.1396825396825397, 4.0 / 15.0}, {0.9619397662556434, 0.0731093246080091, 0.0},
                                        {1.0000000000000000, 0.0079365079365079, 1.0 / 30.0}};
    Region first(double x0, double x1, double y0, double y1) {
        Region g{};
        Cx res{0.0, 0.0}, err{0.0, 0.0};
        for (int i = 0; i < 5; ++i) {
            double y = y0 + R5[i][0] * (y1 - y0);
            for (int j = 0; j < 5; ++j) {
                Cx z = f(x0 + R5[j][0] * (x1 - x0), y);
                res = res + (R5[i][1] * R5[j][1] * (y1 - y0) * (x1 - x0)) * z;
                err = err + (R5[i][2] * R5[j][2] * (y1 - y0) * (x1 - x0)) * z;
                if (i % 2 == 0 && j % 2 == 0) g.cache[i / 2][j / 2] = z;
            }
        }
        g.integral = res;
        g.error = res - err;
        g.x[0] = x0; g.x[1] = x1; g.y[0] = y0; g.y[1] = y1;
        return g;
    }
    void split_at(const Region& g, const double* xs, const double* ys, Region out[4]) {
        double x0 = g.x[0], x1 = g.x[1], y0 = g.y[0], y1 = g.y[1];
        double xm = xs ? *xs : 0.5 * (x0 + x1), ym = ys ? *ys : 0.5 * (y0 + y1);
        Cx z[9][9];
        for (int i = 0; i < 9; ++i) {
            double y = i < 4 ? y0 + X5[i] * (ym - y0) : ym + (X5[i] - 1.0) * (y1 - ym);
            for (int j = 0; j < 9; ++j) {
                double x = j < 4 ? x0 + X5[j] * (xm - x0) : xm + (X5[j] - 1.0) * (x1 - xm);
                if (i % 8 == 0 && j % 8 == 0) z[i][j] = g.cache[i / 4][j / 4];
                else if (i % 4 == 0 && j % 4 == 0 && !xs && !ys) z[i][j] = g.cache[i / 4][j / 4];
                else z[i][j] = f(x, y);
            }
        }
        const double xr[2][2] = {{x0, xm}, {xm, x1}}, yr[2][2] = {{y0, ym}, {ym, y1}};
        for (int q = 0; q < 4; ++q) {
            int r = q / 2, c = q % 2;
            Cx res{0.0, 0.0}, err{0.0, 0.0};
            double dy = yr[r][1] - yr[r][0], dx = xr[c][1] - xr[c][0];
            for (int i = 0; i < 5; ++i)
                for (int j = 0; j < 5; ++j) {
                    Cx v = z[i + 4 * r][j + 4 * c];
                    res = res + (R5[j][1] * R5[i][1] * dx * dy) * v;
                    err = err + (R5[j][2] * R5[i][2] * dx * dy) * v;
                }
            Region& o = out[q];
            o.integral = res;
            o.error = res - err;
            o.x[0] = xr[c][0]; o.x[1] = xr[c][1]; o.y[0] = yr[r][0]; o.y[1] = yr[r][1];
            for (int p = 0; p < 3; ++p)
                for (int t = 0; t < 3; ++t) o.cache[p][t] = z[4 * r + 2 * p][4 * c + 2 * t];
        }
    }
    void quarter(const Region& g, Region out[4]) {
        double x0 = g.x[0], x1 = g.x[1], y0 = g.y[0], y1 = g.y[1];
        double xm = 0.5 * (x0 + x1), ym = 0.5 * (y0 + y1);
        static thread_local Cx z[17][17];
        static const double X9[17] = {0.0000000000000000, 0.0380602337443566, 0.1464466094067262, 0.3086582838174551,
                                      0.5000000000000000, 0.6913417161825449, 0.8535533905932738, 0.9619397662556434,
                                      1.0000000000000000, 1.0380602337443566, 1.1464466094067262, 1.3086582838174551,
                                      1.5000000000000000, 1.6913417161825449, 1.8535533905932738, 1.9619397662556434,
                                      2.0000000000000000};
        for (int i = 0; i < 17; ++i) {
            double y = y0 + 0.5 * X9[i] * (y1 - y0);
            for (int j = 0; j < 17; ++j) {
                double x = x0 + 0.5 * X9[j] * (x1 - x0);
                z[i][j] = (i % 8 == 0 && j % 8 == 0) ? g.cache[i / 8][j / 8] : f(x, y);
            }
        }
        const double xr[2][2] = {{x0, xm}, {xm, x1}}, yr[2][2] = {{y0, ym}, {ym, y1}};
        double dy = 0.5 * (y1 - y0), dx = 0.5 * (x1 - x0);
        for (int q = 0; q < 4; ++q) {
            int r = q / 2, c = q % 2;
            Cx res{0.0, 0.0}, err{0.0, 0.0};
            for (int i = 0; i < 9; ++i)
                for (int j = 0; j < 9; ++j) {
                    Cx v = z[i + 8 * r][j + 8 * c];
                    res = res + (R9[j][1] * R9[i][1] * dx * dy) * v;
                    err = err + (R9[j][2] * R9[i][2] * dx * dy) * v;
                }
            Region& o = out[q];
            o.integral = res;
            o.error = res - err;
            o.x[0] = xr[c][0]; o.x[1] = xr[c][1]; o.y[0] = yr[r][0]; o.y[1] = yr[r][1];
            for (int p = 0; p < 3; ++p)
                for (int t = 0; t < 3; ++t) o.cache[p][t] = z[8 * r + 4 * p][8 * c + 4 * t];
        }
    }
    Cx adaptive() {
        double s_min, s_max;
        bw_s_bounds(n, a, eta, s_min, s_max);
        double s_peak = 0.5 + sqrt(0.25 - (1.0 + a * a) / (2.0 * double(n) * eta));
        bool have_peak = s_peak > s_min + 2.0 * (s_max - s_min) / 3.0;
        double th_split = bw_theta_at(tb, have_peak ? s_peak : 0.5);
        std::vector<Region> regs;
        Region g = first(s_min, s_max, 0.0, 0.5 * PI_D);
        Region q[4];
        split_at(g, have_peak ? &s_peak : nullptr, &th_split, q);
        for (auto& r : q) regs.push_back(r);
        Cx integral{0.0, 0.0};
        const double tol = 1.0e-3;
        for (int it = 0; it < 16; ++it) {
            Cx err{0.0, 0.0};
            integral = Cx{0.0, 0.0};
            for (auto& r : regs) { integral = integral + r.integral; err = err + r.error; }
            if (fabs(err.re) <= tol * fabs(integral.re) && fabs(err.im) <= tol * fabs(integral.im)) break;
            Region worst = regs.back();
            regs.pop_back();
            quarter(worst, q);
            for (auto& r : q) regs.push_back(r);
            // ascending in |error|^2; the last region is the next to split
            for (size_t i = 1; i < regs.size(); ++i) {
                Region key = regs[i];
                double kv = key.error.re * key.error.re + key.error.im * key.error.im;
                size_t j = i;
                while (j > 0 && regs[j - 1].error.re * regs[j - 1].error.re + regs[j - 1].error.im * regs[j - 1].error.im > kv) {
                    regs[j] = regs[j - 1];
                    --j;
                }
                regs[j] = key;
            }
        }
        return 8.0 * integral;
    }
    Cx integrate() { return bw_theta_at(tb, 0.5) < 0.75 * 0.5 * PI_D ? adaptive() : direct(); }
};
constexpr double Partial::R5[5][3];
constexpr double Partial::X5[9];
constexpr double Partial::R9[9][3];

void node_row(double a, double eta, double* out48) {
    int n_min, n_max;
    bw_sum_limits(a, eta, n_min, n_max);
    std::vector<Cx> cum;
    Cx total{0.0, 0.0};
    for (int n = n_min; n <= n_max; ++n) {
        Partial p(n, a, eta);
        total = total + p.integrate();
        cum.push_back(total);
    }
    double u_max = log(1.0 + double(n_max - n_min));
    double nn[16];
    nn[0] = double(n_min);
    double u = 0.0, du = u_max / 15.0;
    for (int i = 1; i <= 15; ++i) {
        u += du;
        double n = double(n_min) + expm1(u);
        if (std::round(n) <= nn[i - 1]) {
            n = nn[i - 1] + 1.0;
            u = log(n - double(n_min) + 1.0);
            du = (u_max - u) / (15.0 - double(i));
            nn[i] = n;
        } else {
            nn[i] = std::round(n);
        }
    }
    for (int i = 0; i < 16; ++i) {
        size_t k = size_t(nn[i]) - size_t(n_min);
        Cx c = k < cum.size() ? cum[k] : cum.back();
        out48[3 * i] = nn[i];
        out48[3 * i + 1] = c.re;
        out48[3 * i + 2] = c.im;
    }
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s in.bin out.bin threads\n", argv[0]);
        return 2;
    }
    FILE* fi = std::fopen(argv[1], "rb");
    if (!fi) { std::perror(argv[1]); return 1; }
    uint64_t n_nodes = 0;
    bool ok = std::fread(&n_nodes, 8, 1, fi) == 1 && std::fread(G16N, 8, 16, fi) == 16 && std::fread(G16W, 8, 16, fi) == 16 &&
              std::fread(G32N, 8, 32, fi) == 32 && std::fread(G32W, 8, 32, fi) == 32;
    std::vector<double> nodes(2 * n_nodes);
    ok = ok && std::fread(nodes.data(), 8, nodes.size(), fi) == nodes.size();
    std::fclose(fi);
    if (!ok) { std::fprintf(stderr, "short input\n"); return 1; }
    std::vector<double> out(48 * n_nodes, 0.0);
    std::atomic<uint64_t> next{0}, done{0};
    int nt = std::max(1, std::atoi(argv[3]));
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&] {
            for (;;) {
                uint64_t i = next.fetch_add(1);
                if (i >= n_nodes) return;
                node_row(nodes[2 * i], nodes[2 * i + 1], &out[48 * i]);
                uint64_t d = done.fetch_add(1) + 1;
                if (d % 500 == 0) std::fprintf(stderr, "%llu / %llu nodes\n", (unsigned long long)d, (unsigned long long)n_nodes);
            }
        });
    for (auto& t : th) t.join();
    FILE* fo = std::fopen(argv[2], "wb");
    if (!fo || std::fwrite(out.data(), 8, out.size(), fo) != out.size()) { std::perror(argv[2]); return 1; }
    std::fclose(fo);
    return 0;
}
