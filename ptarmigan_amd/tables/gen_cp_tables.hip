// Host program: the cumulative harmonic tables of circularly polarised nonlinear
// Compton scattering, which the reference generates with
// src/nonlinear_compton/cp/mod.rs:474-612 (quantum, create_rate_table) and
// cp/classical.rs:220-300 (classical) and of which it ships only the classical one.
//
// Per node (a, eta): the partial rates of harmonics 1..n_max, n_max = int(10 (1 + a^3)),
// from the library's own cp_partial_rate / cp_integrated_spectrum_low_eta
// (csrc/dev_cp.h), summed in order; the row kept is 32 points [n, F(n) / F(n_max)]:
// every harmonic when n_max <= 32; the first 8 harmonics and 24 log-spaced points
// from n = 9 when n_max < 100; otherwise 32 log-spaced points from n = 1, the
// cumulative sum interpolated by PWMCI between harmonics.
//
// Usage: gen_cp_tables <in.bin> <out.bin> <threads>
//   in.bin:  u64 n_nodes, u64 classical (0/1), f64 Airy block (the LCFA pack's
//            AIRY_SERIES .. AIRYP_W4 slots), gl16 nodes[16], gl32 nodes[32], weights[32],
//            then n_nodes x (a, eta)
//   out.bin: n_nodes x (total, 32 x (n, F)) f64
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

#include "../csrc/dev_cp.h"

using namespace ptg;

namespace {

std::vector<double> g_lcfa;
double G16N[16], G32N[32], G32W[32];
bool g_classical = false;

double node_row(double a, double eta, double* out64) {
    const TP tb = g_lcfa.data();
    int n_max = int(10.0 * (1.0 + a * a * a));
    std::vector<double> rates(2 * size_t(n_max));
    double cs = 0.0;
    for (int n = 1; n <= n_max; ++n) {
        double r = g_classical ? cp_integrated_spectrum_low_eta(tb, G32N, G32W, n, a) : cp_partial_rate(tb, G32N, G32W, n, a, eta);
        cs = cs + r;
        rates[2 * size_t(n - 1)] = double(n);
        rates[2 * size_t(n - 1) + 1] = cs;
    }
    const double* R = rates.data();
    double rate = R[2 * size_t(n_max - 1) + 1];
    double limit = R[2 * size_t(n_max - 1)];
    auto eval = [&](double n) {
        double f = 0.0;
        if (!pwmci_eval(R, n_max, n, false, f)) {
            std::fprintf(stderr, "gen_cp_tables: PWMCI evaluate out of range at n = %.17g (a = %g, eta = %g)\n", n, a, eta);
            std::exit(1);
        }
        return f;
    };
    out64[0] = R[0];
    out64[1] = R[1] / rate;
    if (n_max <= 32) {
        for (int i = 1; i <= 31; ++i) {
            if (i < n_max) { out64[2 * i] = R[2 * i]; out64[2 * i + 1] = R[2 * i + 1] / rate; }
            else { out64[2 * i] = double(i + 1); out64[2 * i + 1] = 1.0; }
        }
    } else if (n_max < 100) {
        for (int i = 1; i <= 7; ++i) { out64[2 * i] = R[2 * i]; out64[2 * i + 1] = R[2 * i + 1] / rate; }
        double delta = (log(double(n_max)) - log(9.0)) / (31.0 - 8.0);
        for (int i = 8; i <= 31; ++i) {
            double n = fmin(exp(log(9.0) + double(i - 8) * delta), limit);
            out64[2 * i] = n;
            out64[2 * i + 1] = eval(n) / rate;
        }
    } else {
        double delta = log(double(n_max)) / 31.0;
        for (int i = 1; i <= 31; ++i) {
            double n = fmin(exp(double(i) * delta), limit);
            out64[2 * i] = n;
            out64[2 * i + 1] = eval(n) / rate;
        }
    }
    return rate;
}

bool read_all(FILE* f, void* p, size_t bytes) { return std::fread(p, 1, bytes, f) == bytes; }

}  // namespace

int main(int argc, char** argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s in.bin out.bin threads\n", argv[0]);
        return 2;
    }
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) { std::perror(argv[1]); return 1; }
    uint64_t n_nodes = 0, classical = 0;
    const int airy_len = lcfa_pack::AIRYP_W4 + 4 - lcfa_pack::AIRY_SERIES;
    g_lcfa.assign(size_t(lcfa_pack::AIRYP_W4 + 4), 0.0);
    bool ok = read_all(f, &n_nodes, 8) && read_all(f, &classical, 8) &&
              read_all(f, g_lcfa.data() + lcfa_pack::AIRY_SERIES, sizeof(double) * airy_len) && read_all(f, G16N, sizeof G16N) &&
              read_all(f, G32N, sizeof G32N) && read_all(f, G32W, sizeof G32W);
    std::vector<double> nodes(2 * n_nodes);
    ok = ok && read_all(f, nodes.data(), sizeof(double) * nodes.size());
    std::fclose(f);
    if (!ok) { std::fprintf(stderr, "gen_cp_tables: short input\n"); return 1; }
    g_classical = classical != 0;
    int threads = std::max(1, std::atoi(argv[3]));
    std::vector<double> out(n_nodes * 65);
    // the nodes with the most harmonics first, so the threads finish together
    std::vector<size_t> order(n_nodes);
    for (size_t i = 0; i < n_nodes; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return nodes[2 * x] > nodes[2 * y]; });
    std::atomic<size_t> next{0}, done{0};
    auto work = [&]() {
        for (;;) {
            size_t k = next.fetch_add(1);
            if (k >= n_nodes) return;
            size_t i = order[k];
            out[65 * i] = node_row(nodes[2 * i], nodes[2 * i + 1], &out[65 * i + 1]);
            size_t d = done.fetch_add(1) + 1;
            if (d % 200 == 0 || d == n_nodes) std::fprintf(stderr, "gen_cp_tables: %zu / %llu nodes\n", d, (unsigned long long)n_nodes);
        }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t) pool.emplace_back(work);
    for (auto& t : pool) t.join();
    FILE* g = std::fopen(argv[2], "wb");
    if (!g || std::fwrite(out.data(), sizeof(double), out.size(), g) != out.size()) { std::perror(argv[2]); return 1; }
    std::fclose(g);
    return 0;
}
